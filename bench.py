"""Benchmark of the sea-ice step on one MI355X.

Workload: the plain method as a GPU drop-in on the 1-degree global grid
(360 x 160 interior cells + 2-cell halo, FP64): one step = pre-EVP forcing,
400 fused EVP substeps, ocean stress, advection, clean-up, ridging and growth.
Inputs are the seeded synthetic fields of :mod:`veris_amd.synth`, resident on
the GPU before timing starts.

``value`` = interior cells x EVP substeps x steps / wall time of the timed
steps (cell-substeps/s).  ``roofline`` covers the dominant kernel, the fused
EVP substep: algorithmic bytes per launch (38 words per interior cell, 57,600
cells) over its average launch time measured with HIP events on the stream the
substeps are launched on.  ``cpu_baseline`` times the NumPy FP64 oracle (one
thread) on one 1-degree step.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
B_ALG_WORDS = 38  # algorithmic words per EVP cell-substep (33 reads + 5 writes)


def cpu_baseline(n_substeps: int) -> dict:
    """NumPy FP64 oracle on one full 1-degree step (single thread)."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    from oracle import veris_np as O
    from veris_amd import synth

    st = synth.config1()
    q0 = st.variables.Qnet.copy()
    sw = st.variables.SWdown.copy()
    nx, ny = st.settings.nx, st.settings.ny
    t0 = time.perf_counter()
    O.seaice_step(st, q0, sw, n_substeps=n_substeps)
    dt = time.perf_counter() - t0
    return {
        "value": nx * ny * n_substeps / dt,
        "unit": "cell-substeps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"1 full step of the 1-degree workload ({n_substeps} EVP substeps), "
                  f"NumPy FP64 oracle, {dt:.1f} s, host affinity {len(os.sched_getaffinity(0))} cpus",
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--substeps", type=int, default=400)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: every box in this setup has a single MI355X; "
              "run with --gpus 1", file=sys.stderr)
        return 2

    import torch

    from veris_amd import kernels as K
    from veris_amd import synth
    from veris_amd.step import seaice_step

    dev = torch.device("cuda:0")
    st = synth.config1()
    nx, ny = st.settings.nx, st.settings.ny
    q0 = torch.from_numpy(st.variables.Qnet.copy()).to(dev)
    sw = torch.from_numpy(st.variables.SWdown.copy()).to(dev)
    g = st.to(dev)
    nsub = args.substeps

    # one step first (pre-EVP fields in place), then pick the fastest exact EVP
    # workgroup shape on this state (every exact shape gives identical bits)
    from veris_amd import _native
    from veris_amd.autotune import tune_evp

    seaice_step(g, q0, sw, n_substeps=nsub)
    tile = tune_evp(g, exact=True, n_substeps=100)
    ri, rj, kk = _native.evp_tile_shapes()[tile]
    for _ in range(args.warmup):
        seaice_step(g, q0, sw, n_substeps=nsub)
    torch.cuda.synchronize()

    # timed region: K full steps; HIP events bracket the EVP subcycle on the
    # stream its kernels are launched on (the current stream)
    stream = torch.cuda.current_stream(dev)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    orig = K.IceVelocities
    counter = {"k": 0}

    def timed_ice_velocities(state, n_substeps=None):
        a, b = ev[counter["k"]]
        a.record(stream)
        out = orig(state, n_substeps)
        b.record(stream)
        counter["k"] += 1
        return out

    import veris_amd.step as step_mod

    step_mod.K.IceVelocities = timed_ice_velocities
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        seaice_step(g, q0, sw, n_substeps=nsub)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_mod.K.IceVelocities = orig

    evp_ms = sum(a.elapsed_time(b) for a, b in ev)
    per_launch_s = evp_ms * 1e-3 / (args.steps * nsub)
    cells = nx * ny
    value = cells * nsub * args.steps / wall
    bytes_per_launch = B_ALG_WORDS * 8 * cells
    achieved = bytes_per_launch / per_launch_s / 1e9
    finite = all(bool(torch.isfinite(g.variables[k]).all()) for k in ("uIce", "hIceMean", "Qnet"))

    line = {
        "metric": "EVP cell-substeps/s at 1 degree (plain FP64 GPU drop-in, full step)",
        "value": value,
        "unit": "cell-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded fields, see veris_amd/synth.py)",
        "config": {"workload": "1-degree global grid 360x160 + 2-cell halo, FP64, full sea-ice "
                               "step with 400 EVP substeps + advection + growth",
                   "nx": nx, "ny": ny, "evp_substeps": nsub, "parallelism": "none (1 GPU)"},
        "roofline": {
            "bound": "hbm",
            "kernel": f"evp_substep_kernel<double,double,{ri},{rj},{kk},false>",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "avg_launch_us": per_launch_s * 1e6,
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
        "finite": finite,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(nsub)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
