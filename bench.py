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
substeps are launched on.  ``cpu_baseline`` times the scalar-loop C
restatement of the step (on one thread and on every host CPU) and the NumPy
FP64 one (one thread) on about 8 s of full 1-degree steps each, and reports
the fastest with its thread count.

Beside the headline line the same JSON object carries the other measured
workloads (skip them with ``--no-extras``):

* ``batched``: 48 regional members of sizes 64..512 x 64..384 in one batched
  EVP launch sequence, FP32 bulk (Delta, zeta, press in FP64) and FP64, 400
  substeps, autotuned tile shapes; rates, the ratio to the plain FP64 1-degree
  drop-in, and the error of every EVP output of the FP32 run against the FP64
  run (the FP64 GPU path equals the NumPy oracle bit for bit) over interior
  ocean cells, plus the FP64 budget sums of the FP32 fields;
* ``adjoint``: pre-EVP + 120-substep EVP subcycle at 1 degree with the
  gradient of a velocity misfit w.r.t. a P* field, the drag coefficients and
  hIceMean (checkpoint every 10 substeps), FP64 and FP32: wall time of
  forward + backward over the plain drop-in forward and over the
  checkpointed forward, and central finite differences along 20 random
  directions (FP64);
* ``hires``: the 0.25-degree 1440 x 720 grid, one full step with 800 EVP
  substeps in FP32 and in FP64, the error of every output of the FP32 step
  against the FP64 step, FP64 budget sums, and the batched adjoint of 8
  members at 240 x 240;
* ``variants``: the 1-degree subcycle with adaptive EVP, free slip with
  lateral drag and the second-order boundary term, the residual-norm
  subcycle, free drift and the open-ocean bulk fluxes, timed per substep or
  per call.
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


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "evp_substep_1deg_fp64_traffic.md")


KERNEL_PROFILE_FILE = os.path.join(ROOT, "profiles", "evp_substep_1deg_fp64.md")


def evp_kernel_name(ri: int, rj: int, k: int, plain: bool = True, batched: bool = False,
                    t: str = "double", a: str = "double", s: str = "double") -> str:
    """The EVP substep instantiation a launch of shape ri x rj x k runs (plain:
    the default-settings one): its template arguments as rocprofv3 lists them,
    without the spaces (the key of the profile summaries)."""
    return (f"evp_substep_kernel<{t},double,{ri},{rj},{k},{str(batched).lower()},{a},"
            f"false,false,{s},{str(plain).lower()}>")


def _profile_value(path: str, key: str, kernel: str):
    """The number on the line ``<key> <kernel> <value>`` of a profile summary,
    or None if the summary has no such line."""
    try:
        with open(path) as fh:
            for line in fh:
                parts = line.split()
                if len(parts) == 3 and parts[0] == key and parts[1] == kernel:
                    return float(parts[2])
    except OSError:
        pass
    return None


def measured_traffic(kernel: str):
    """HBM bytes per launch of ``kernel`` from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (copy-calibrated), or None if it was not measured."""
    return _profile_value(TRAFFIC_FILE, "traffic_bytes_per_launch", kernel)


def cpu_baseline(n_substeps: int, min_seconds: float = 8.0, max_steps: int = 12) -> dict:
    """The CPU restatements on full 1-degree steps: the scalar-loop C one
    (gcc -O2) on one thread and on every host CPU this process may use
    (OpenMP over the cell loops), and the NumPy FP64 one (one thread), each
    repeated until at least ``min_seconds`` of CPU work; the value is the
    fastest, with the threads it used as ``cores``."""
    from oracle import c_ref
    from oracle import veris_np as O
    from veris_amd import synth

    st = synth.config1()
    q0 = st.variables.Qnet.copy()
    sw = st.variables.SWdown.copy()
    nx, ny = st.settings.nx, st.settings.ny
    def timed(step):
        n = 0
        t0 = time.perf_counter()
        while n < max_steps and (n == 0 or time.perf_counter() - t0 < min_seconds):
            step()
            n += 1
        return n, time.perf_counter() - t0

    def sample(n, dt, what):
        return (f"{n} full step{'s' if n > 1 else ''} of the 1-degree workload "
                f"({n_substeps} EVP substeps each) through {what}, {dt:.1f} s")

    ncpu = len(os.sched_getaffinity(0))
    legs = {}
    try:  # the C legs need gcc where the library is not built yet
        c_ref.build()
        for key, threads in (("c_port", 1), ("c_port_openmp", ncpu)):
            used = c_ref.set_threads(threads)
            n, dt = timed(lambda: c_ref.seaice_step(st, q0, sw, n_substeps=n_substeps))
            legs[key] = {"value": nx * ny * n_substeps * n / dt, "cores": used,
                         "sample": sample(n, dt, f"the scalar-loop C restatement (gcc -O2, FP64, "
                                                 f"{used} OpenMP thread{'s' if used > 1 else ''})")}
    except Exception as e:  # noqa: BLE001
        legs["c_port"] = {"error": repr(e)}
    n, dt = timed(lambda: O.seaice_step(st, q0, sw, n_substeps=n_substeps))
    legs["numpy_port"] = {"value": nx * ny * n_substeps * n / dt, "cores": 1,
                          "sample": sample(n, dt, "the NumPy FP64 restatement (one thread)")}
    # the value is the fastest CPU restatement
    best = max((v for v in legs.values() if "value" in v), key=lambda v: v["value"])
    return {
        "value": best["value"],
        "unit": "cell-substeps/s",
        "cores": best["cores"],
        "kind": "port",
        "sample": best["sample"] + f"; host affinity {ncpu} cpus",
        **legs,
    }


def _sync_time(fn, reps: int = 5) -> float:
    """Median wall seconds of ``fn()`` between device synchronisations."""
    import torch

    ts = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return ts[len(ts) // 2]


def _sync_times_interleaved(fns, reps: int) -> list:
    """Median wall seconds of each of ``fns`` between device
    synchronisations, the functions called in turn for ``reps`` rounds."""
    import torch

    ts = [[] for _ in fns]
    for _ in range(reps):
        for q, fn in enumerate(fns):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts[q].append(time.perf_counter() - t0)
    return [sorted(t)[len(t) // 2] for t in ts]


def _rel_l2(pairs):
    """sqrt(sum |got - ref|^2 / sum |ref|^2) over (got, ref, mask) triples of
    padded fields, interior cells with mask > 0."""
    import torch

    num = torch.zeros((), dtype=torch.float64, device=pairs[0][1].device)
    den = torch.zeros_like(num)
    for got, ref, mask in pairs:
        sel = mask[2:-2, 2:-2] > 0
        r = ref[2:-2, 2:-2].double()[sel]
        d = got[2:-2, 2:-2].double()[sel] - r
        num = num + (d * d).sum()
        den = den + (r * r).sum()
    return float((num / den).sqrt()) if float(den) > 0 else 0.0


def _rounded_to_fp32(state):
    """The state with every floating-point field rounded to the nearest FP32
    value (kept in FP64)."""
    import numpy as np

    vs = state.variables
    for k in list(vs.keys()):
        a = vs[k]
        if isinstance(a, np.ndarray) and a.dtype == np.float64:
            vs[k] = a.astype(np.float32).astype(np.float64)
    return state


EVP_FP_INPUTS = ("SeaIceStrength", "SeaIceMassC", "SeaIceMassU", "SeaIceMassV", "AreaW", "AreaS",
                 "WindForcingX", "WindForcingY", "uOcean", "vOcean", "fCori")
EVP_METRICS = ("recip_dxU", "recip_dyV", "recip_dyU", "recip_dxV", "dxU", "dyU", "dxV", "dyV",
               "recip_rAu", "recip_rAv", "rAz", "recip_rA")


def _fp32_field_sensitivity(e64, tile, nsub, ref, masks) -> dict:
    """Velocity error (relative L2 of uIce and vIce together, interior ocean
    cells) of the FP64 batched subcycle when exactly one of its floating-point
    inputs is rounded to FP32 (the masks, the zero initial state and every
    other field exact), of the twelve grid metrics rounded together, and of
    every input rounded except the three fields that move the result most."""
    import torch

    def run_with(rounded):
        saved = {}
        for name in rounded:
            buf = e64.fields[name]
            saved[name] = buf.clone()
            buf.copy_(buf.float().double())
        out = e64.evp(nsub, tile=tile)
        for name, keep in saved.items():
            e64.fields[name].copy_(keep)
        pairs = list(zip(e64.unpack(out[0]), e64.unpack(ref[0]), masks))
        pairs += list(zip(e64.unpack(out[1]), e64.unpack(ref[1]), masks))
        del out
        torch.cuda.synchronize()
        return _rel_l2(pairs)

    one = {name: run_with((name,)) for name in EVP_FP_INPUTS + EVP_METRICS}
    worst = sorted(one, key=one.get, reverse=True)[:3]
    return {
        "one_field": one,
        "metrics_together": run_with(EVP_METRICS),
        "all_but_top3": {"kept_fp64": worst,
                         "rel_l2": run_with(tuple(n for n in one if n not in worst))},
        "all": run_with(tuple(one)),
    }


def bench_batched(dev, nsub: int, fp64_evp_rate: float, fp64_step_rate: float) -> dict:
    import torch

    from veris_amd import synth
    from veris_amd.autotune import tune_batched
    from veris_amd.batched import EVP_OUT, Ensemble

    t_gen = time.perf_counter()
    states = [synth.config2_member(m) for m in range(len(synth.config2_sizes()))]
    e32 = Ensemble(states, dev, torch.float32)
    e64 = Ensemble(states, dev, torch.float64)
    e32.pre_evp()
    e64.pre_evp()
    t_gen = time.perf_counter() - t_gen
    cells = e32.interior_cells()
    # members are closed regional domains: every tile shape is exact
    tile32, times32 = tune_batched(e32, exact=False)
    tile64, _ = tune_batched(e64, exact=False)
    out32 = e32.evp(nsub, tile=tile32)
    out64 = e64.evp(nsub, tile=tile64)
    t32 = _sync_time(lambda: e32.evp(nsub, tile=tile32, out=out32), 7)
    t64 = _sync_time(lambda: e64.evp(nsub, tile=tile64, out=out64), 7)
    masks = e64.unpack("maskInC")
    err = {}
    for q, name in enumerate(EVP_OUT):
        err[name] = _rel_l2(list(zip(e32.unpack(out32[q]), e64.unpack(out64[q]), masks)))
    # FP32 storage with FP64 arithmetic
    e32a = Ensemble(states, dev, torch.float32, arith="fp64")
    e32a.pre_evp()
    tile32a, _ = tune_batched(e32a, exact=False)
    out32a = e32a.evp(nsub, tile=tile32a)
    t32a = _sync_time(lambda: e32a.evp(nsub, tile=tile32a, out=out32a))
    err_a = {name: _rel_l2(list(zip(e32a.unpack(out32a[q]), e64.unpack(out64[q]), masks)))
             for q, name in enumerate(EVP_OUT)}
    del e32a
    # FP32-stored step constants around an FP64 state, FP64 arithmetic
    e32s = Ensemble(states, dev, torch.float32, arith="fp64_state")
    e32s.pre_evp()
    tile32s, _ = tune_batched(e32s, exact=False)
    out32s = e32s.evp(nsub, tile=tile32s)
    t32s = _sync_time(lambda: e32s.evp(nsub, tile=tile32s, out=out32s))
    err_s = {name: _rel_l2(list(zip(e32s.unpack(out32s[q]), e64.unpack(out64[q]), masks)))
             for q, name in enumerate(EVP_OUT)}
    del e32s, out32s
    # the same FP64 subcycle on inputs rounded to FP32: what the rounding of
    # the inputs alone moves the result by, and the FP32 run against FP64 on
    # identical (FP32-representable) inputs
    states_q = [_rounded_to_fp32(synth.config2_member(m)) for m in range(len(states))]
    e64q = Ensemble(states_q, dev, torch.float64)
    e64q.pre_evp()
    out64q = e64q.evp(nsub, tile=tile64)
    err_inputs = {name: _rel_l2(list(zip(e64q.unpack(out64q[q]), e64.unpack(out64[q]), masks)))
                  for q, name in enumerate(EVP_OUT)}
    err_same = {name: _rel_l2(list(zip(e32.unpack(out32[q]), e64q.unpack(out64q[q]), masks)))
                for q, name in enumerate(EVP_OUT)}
    del e64q
    # how well the FP64 reference itself is determined: the FP64 subcycle
    # with the ice strength perturbed by 1e-15 (a few ulps), and where the
    # difference sits; the FP32 error over the cells that perturbation leaves
    # stable (below the 99th percentile of its velocity response)
    P = e64.fields["SeaIceStrength"]
    P0 = P.clone()
    gen = torch.Generator(device=dev).manual_seed(5)
    P.mul_(1 + 1e-15 * torch.randn(P.shape, generator=gen, device=dev, dtype=P.dtype))
    outp = e64.evp(nsub, tile=tile64)
    P.copy_(P0)
    err_pert = {name: _rel_l2(list(zip(e64.unpack(outp[q]), e64.unpack(out64[q]), masks)))
                for q, name in enumerate(EVP_OUT[:2])}
    def interior(buf, m):
        v = e64.view(buf, m)[2:-2, 2:-2]
        return v[e64.view(e64.fields["maskInC"], m)[2:-2, 2:-2] > 0]

    resp = torch.cat([interior((outp[0] - out64[0]).abs() + (outp[1] - out64[1]).abs(), m)
                      for m in range(e64.n)])
    cut = float(torch.quantile(resp.double()[:: max(1, resp.numel() // 2_000_000)], 0.99))
    stable = ((outp[0] - out64[0]).abs() + (outp[1] - out64[1]).abs() <= cut).to(
        out64[0].dtype) * e64.fields["maskInC"]
    err_stable = {name: _rel_l2(list(zip(e32.unpack(out32[q]), e64.unpack(out64[q]),
                                         e64.unpack(stable))))
                  for q, name in enumerate(EVP_OUT[:2])}
    d2 = torch.cat([interior((out32[0].double() - out64[0]) ** 2
                             + (out32[1].double() - out64[1]) ** 2, m) for m in range(e64.n)])
    top = torch.sort(d2, descending=True).values
    top_share = float(top[: max(1, top.numel() // 100)].sum() / top.sum().clamp_min(1e-300))
    del outp
    sens = _fp32_field_sensitivity(e64, tile64, nsub, out64, masks)
    b32 = e32.budgets()
    b64 = e64.budgets()
    budget_err = {k: float(((b32[k] - b64[k]).abs() / b64[k].abs().clamp_min(1e-300)).max())
                  for k in b64}
    from veris_amd import _native

    shapes = _native.evp_tile_shapes()
    rate32 = cells * nsub / t32
    return {
        "workload": f"{e32.n} regional members, sizes 64..512 x 64..384, {cells} interior cells, "
                    f"{nsub} EVP substeps, one batched launch sequence (hipGraph)",
        "fp32_value": rate32,
        "fp64_value": cells * nsub / t64,
        "unit": "cell-substeps/s",
        "fp32_ms_per_subcycle": t32 * 1e3,
        "fp32_tile": {"index": tile32, "shape": shapes[tile32]},
        "fp64_tile": {"index": tile64, "shape": shapes[tile64]},
        "fp32_tile_us_per_substep": {str(t): round(v * 1e6, 2) for t, v in times32.items()},
        "fp32_over_plain_fp64_evp_1deg": rate32 / fp64_evp_rate,
        "fp32_over_plain_fp64_full_step_1deg": rate32 / fp64_step_rate,
        "target_ratio": 30.0,
        "fp32_roofline": {"bound": "hbm", "achieved": rate32 * B_ALG_WORDS * 4 / 1e9,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": rate32 * B_ALG_WORDS * 4 / 1e9 / HBM_PEAK_GBS},
        "fp32_rel_l2_vs_fp64": err,
        "velocity_error_target": 1e-4,
        "fp64_on_fp32_rounded_inputs_rel_l2_vs_fp64": err_inputs,
        "fp32_rel_l2_vs_fp64_on_fp32_rounded_inputs": err_same,
        "fp64_strength_perturbed_1e-15_rel_l2_vs_fp64": err_pert,
        "fp32_squared_velocity_error_share_of_worst_1pct_cells": top_share,
        "fp32_rel_l2_vs_fp64_on_cells_stable_under_1e-15": err_stable,
        "fp64_with_one_input_rounded_to_fp32": sens,
        "fp32_storage_fp64_arith": {
            "value": cells * nsub / t32a,
            "tile": {"index": tile32a, "shape": shapes[tile32a]},
            "over_plain_fp64_evp_1deg": cells * nsub / t32a / fp64_evp_rate,
            "rel_l2_vs_fp64": err_a},
        "fp32_constants_fp64_state": {
            "value": cells * nsub / t32s,
            "tile": {"index": tile32s, "shape": shapes[tile32s]},
            "over_plain_fp64_evp_1deg": cells * nsub / t32s / fp64_evp_rate,
            "rel_l2_vs_fp64": err_s},
        "fp64_budget_rel_diff_of_fp32_fields": budget_err,
        "setup_s": t_gen,
    }


def _adjoint_run(state, params, target, n, ckpt, tile=0):
    from veris_amd import adjoint as A

    outs = A.evp_subcycle(state, n_substeps=n, ckpt_every=ckpt, tile=tile, **params)
    ut, vt = target
    return ((outs[0] - ut)[2:-2, 2:-2] ** 2).sum() + ((outs[1] - vt)[2:-2, 2:-2] ** 2).sum()


def bench_adjoint(dev, ckpt: int = 10) -> dict:
    import numpy as np
    import torch

    from veris_amd import kernels as K
    from veris_amd import synth

    res = {"workload": "1-degree 360x160, pre-EVP + 120 EVP substeps, checkpoint every "
                       f"{ckpt}; loss = L2 misfit of final uIce/vIce to a seeded target; "
                       "gradients w.r.t. a lognormal P* field, the water and air drag "
                       "coefficients (one scalar each, both hemispheres) and hIceMean"}
    st0 = synth.config3()
    n = int(st0.settings.nEVPsteps)
    pstar_np, ut_np, vt_np = synth.config3_adjoint_inputs(st0)
    for dtype, tag in ((torch.float64, "fp64"), (torch.float32, "fp32")):
        st = st0.to(dev, dtype)
        target = (torch.tensor(ut_np, dtype=dtype, device=dev),
                  torch.tensor(vt_np, dtype=dtype, device=dev))

        def leafs():
            cw = torch.tensor(st0.settings.waterIceDrag, dtype=torch.float64, device=dev,
                              requires_grad=True)
            ca = torch.tensor(st0.settings.airIceDrag, dtype=torch.float64, device=dev,
                              requires_grad=True)
            return {"pStar": torch.tensor(pstar_np, device=dev).requires_grad_(True),
                    "hIceMean": st.variables.hIceMean.detach().clone().requires_grad_(True),
                    "waterIceDrag": cw, "waterIceDrag_south": cw,
                    "airIceDrag": ca, "airIceDrag_south": ca}

        p = leafs()
        pst = torch.tensor(pstar_np, dtype=dtype, device=dev)
        # the inputs are periodic in the x overlap: the forward and the
        # recompute run two substeps per launch (same bits as one)
        from veris_amd.adjoint import pair_tile

        tile = pair_tile(st, pStar=pst)

        def plain_forward():
            vs = st.variables
            vs.SeaIceMassC, vs.SeaIceMassU, vs.SeaIceMassV = K.SeaIceMass(st)
            vs.AreaW, vs.AreaS = K.AreaWS(st)
            vs.WindForcingX, vs.WindForcingY = K.WindForcingXY(st)
            vs.SeaIceStrength = K.SeaIceStrength(st, pStar=pst)
            K.evp_solver(st, n)

        def ckpt_forward():
            with torch.no_grad():
                _adjoint_run(st, p, target, n, ckpt, tile)

        def fwd_bwd():
            for v in p.values():
                v.grad = None
            _adjoint_run(st, p, target, n, ckpt, tile).backward()

        # the plain forward on its fastest launch shape for this state and precision
        from veris_amd.autotune import tune_evp

        plain_forward()
        tune_evp(st, exact=False, n_substeps=n)
        for fn in (plain_forward, ckpt_forward, fwd_bwd):
            fn()
        # the three measured in turn, round after round, so that each call
        # starts from the same condition and a drift of the device's clock
        # touches all three alike (medians of 15); beside them the plain
        # forward repeated in a loop of its own, where its working set stays
        # in the MALL from one call to the next (faster, a stricter divisor)
        t_plain, t_fwd, t_fb = _sync_times_interleaved((plain_forward, ckpt_forward, fwd_bwd), 15)
        t_own = _sync_time(plain_forward, 15)
        r = {"plain_forward_ms": t_plain * 1e3, "checkpointed_forward_ms": t_fwd * 1e3,
             "forward_backward_ms": t_fb * 1e3, "forward_tile": tile,
             "plain_forward_tile": K.EVP_TILE.get(K.evp_tile_key(st), 0),
             "adjoint_over_plain_forward": t_fb / t_plain,
             "adjoint_over_checkpointed_forward": t_fb / t_fwd,
             "plain_forward_ms_own_loop": t_own * 1e3,
             "adjoint_over_plain_forward_own_loop": t_fb / t_own,
             "timing": "plain forward, checkpointed forward and forward + backward called in "
                       "turn, 15 rounds, medians; own_loop: the plain forward 15 times in a row",
             "target": 4.0}
        if tag == "fp64":
            fwd_bwd()
            grads = {k: v.grad.detach().clone() for k, v in p.items()
                     if k not in ("waterIceDrag_south", "airIceDrag_south")}
            base = {k: v.detach() for k, v in p.items()}
            rng = np.random.default_rng(11)
            worst = 0.0
            eps = 1e-6
            for _ in range(20):
                dirs = {k: torch.tensor(rng.standard_normal(tuple(base[k].shape)), device=dev,
                                        dtype=torch.float64) * base[k].double()
                        for k in grads}

                def J(sign):
                    q = {k: (base[k].double() + sign * eps * dirs[k]).to(base[k].dtype)
                         for k in grads}
                    q["waterIceDrag_south"] = q["waterIceDrag"]
                    q["airIceDrag_south"] = q["airIceDrag"]
                    with torch.no_grad():
                        return float(_adjoint_run(st, q, target, n, ckpt))

                fd = (J(1) - J(-1)) / (2 * eps)
                ad = sum(float((grads[k].double() * dirs[k]).sum()) for k in grads)
                worst = max(worst, abs(fd - ad) / max(abs(ad), 1e-300))
            r["fd_directions"] = 20
            r["fd_max_rel_err"] = worst
            r["fd_target"] = 1e-5
        res[tag] = r
    return res


def bench_shard(dev, st, stepped) -> dict:
    """The EVP subcycle on one x-strip of an 8-way split of the 0.25-degree
    grid (180 x 720 interior columns plus the 2-cell overlap), the shape one
    GPU would own, beside the whole grid on one GPU.  The strip is a closed
    domain here: its overlap columns stand for the neighbours' values that a
    split run would exchange each substep (2 columns x 724 rows x u, v)."""
    import torch

    from veris_amd import _native
    from veris_amd import kernels as K
    from veris_amd.autotune import tune_evp
    from veris_amd.state import VerisState, make_state_settings
    from veris_amd.variables import VariablesNamespace

    nsub = int(st.settings.nEVPsteps)
    nx_s, NY = st.settings.nx // 8, st.settings.ny + 4
    res = {"workload": f"{nx_s}x{st.settings.ny} + 2-cell halo (one strip of an 8-way x split of "
                       f"the 0.25-degree grid), {nsub} EVP substeps after one full step",
           "exchange_bytes_per_neighbour_per_substep": 2 * NY * 2 * 8}
    for tag, full in (("fp64", stepped["fp64"]), ("fp32", stepped["fp32"])):
        arrays = {}
        for k, a in full.variables.items():
            arrays[k] = (a[: nx_s + 4].clone() if isinstance(a, torch.Tensor) and a.dim() == 2
                         and a.shape[0] == full.variables.uIce.shape[0] else a)
        sett = make_state_settings(nx_s, st.settings.ny, enable_cyclic_x=False, nEVPsteps=nsub)
        strip = VerisState(sett, VariablesNamespace(arrays))
        tile = tune_evp(strip, exact=False, n_substeps=100)  # a closed strip: every shape exact
        t_strip = _sync_time(lambda: K.evp_solver(strip, nsub), 3)
        t_full = _sync_time(lambda: K.evp_solver(full, nsub), 3)
        res[f"{tag}_strip_us_per_substep"] = t_strip / nsub * 1e6
        res[f"{tag}_full_grid_us_per_substep"] = t_full / nsub * 1e6
        res[f"{tag}_strip_tile"] = list(_native.evp_tile_shapes()[tile])
    return res


def bench_variants(dev, nsub: int) -> dict:
    """The alternative dynamics settings and the forcing producer on the
    1-degree workload, after one full step: microseconds per EVP substep
    (median replay of the nsub-substep subcycle, its argument and packing
    kernels included) with the HBM fraction at 38 words per cell-substep;
    the residual-norm subcycle; free drift and the open-ocean bulk fluxes per
    call of the Python drop-in (argument marshalling included; the bulk
    fluxes move 6 input and 9 output words per cell): the wall time of one
    call between device synchronisations, and the wall time per call of 200
    calls issued back to back with one synchronisation at the end, as a step
    issues them (host work overlapping the kernels)."""
    import numpy as np
    import torch

    from veris_amd import kernels as K
    from veris_amd import synth
    from veris_amd.heat_flux_MITgcm import bulkf_formula_lanl
    from veris_amd.step import seaice_step

    def prepared(**kw):
        st = synth.config1(**kw)
        g = st.to(dev)
        seaice_step(g, torch.from_numpy(st.variables.Qnet.copy()).to(dev),
                    torch.from_numpy(st.variables.SWdown.copy()).to(dev), n_substeps=nsub)
        return st, g

    def back_to_back(fn, reps=200):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    def per_substep(fn, cells):
        fn()
        t = _sync_time(fn) / nsub
        bps = B_ALG_WORDS * 8 * cells / t
        return {"us_per_substep": t * 1e6, "cell_substeps_per_s": cells / t,
                "hbm_frac_at_38_words": bps / 1e9 / HBM_PEAK_GBS}

    from veris_amd import _native
    from veris_amd.autotune import tune_evp

    out = {"workload": f"1-degree 360x160 + halo, FP64, {nsub} EVP substeps after one full step, "
                       "each setting on its autotuned launch shape"}
    for name, kw in (("default", {}), ("adaptive_evp", {"useAdaptiveEVP": True}),
                     ("free_slip_side_drag", {"noSlip": False}),
                     ("second_order_bc", {"secondOrderBC": True})):
        st, g = prepared(**kw)
        cells = st.settings.nx * st.settings.ny
        tile = tune_evp(g, exact=False, n_substeps=100)
        out[name] = per_substep(lambda: K.evp_solver(g, nsub), cells)
        out[name]["tile"] = list(_native.evp_tile_shapes()[tile])
    st, g = prepared()
    cells = st.settings.nx * st.settings.ny
    out["residual_norms"] = per_substep(lambda: K.evp_solver_residuals(g, nsub), cells)
    st, g = prepared(useFreedrift=True, useEVP=False)
    K.freedrift_solver(g)
    t = _sync_time(lambda: K.freedrift_solver(g))
    tb = back_to_back(lambda: K.freedrift_solver(g))
    out["free_drift"] = {"wall_us_per_call": t * 1e6, "cells_per_s": cells / t,
                         "back_to_back_us_per_call": tb * 1e6}
    r = np.random.default_rng(77)
    tsf = r.uniform(271.0, 303.0, st.shape)
    ins = [r.normal(0, 8, st.shape), r.normal(0, 8, st.shape), tsf + r.normal(0, 6, st.shape),
           r.uniform(1e-3, 2e-2, st.shape), tsf, st.variables.iceMask.copy()]
    ins = [torch.from_numpy(a).to(dev) for a in ins]
    bulkf_formula_lanl(g, *ins)
    t = _sync_time(lambda: bulkf_formula_lanl(g, *ins))
    tb = back_to_back(lambda: bulkf_formula_lanl(g, *ins))
    padded = st.shape[0] * st.shape[1]
    out["bulk_fluxes"] = {"wall_us_per_call": t * 1e6, "cells_per_s": padded / t,
                          "hbm_frac_at_15_words": 15 * 8 * padded / t / 1e9 / HBM_PEAK_GBS,
                          "back_to_back_us_per_call": tb * 1e6}
    return out


def bench_hires(dev) -> dict:
    import torch

    from veris_amd import _native
    from veris_amd import adjoint as A
    from veris_amd import kernels as K
    from veris_amd import synth
    from veris_amd.autotune import tune_evp
    from veris_amd.step import seaice_step

    st = synth.config4()
    nsub = int(st.settings.nEVPsteps)
    nx, ny = st.settings.nx, st.settings.ny
    res = {"workload": f"0.25-degree {nx}x{ny} + 2-cell halo, one full step with {nsub} EVP "
                       "substeps, FP32 bulk (Delta/zeta/press and growth in FP64) against FP64"}
    q0 = st.variables.Qnet.copy()
    sw = st.variables.SWdown.copy()
    out = {}
    for dtype, tag in ((torch.float64, "fp64"), (torch.float32, "fp32")):
        g = st.to(dev, dtype)
        q = torch.from_numpy(q0).to(dev, dtype)
        s_ = torch.from_numpy(sw).to(dev, dtype)
        start = g.copy(lambda a: a.clone())
        seaice_step(g, q, s_, n_substeps=nsub)  # first step (graph capture)
        # the EVP launch shape for this grid and precision, among every shape
        # (its inputs are periodic in the x overlap)
        tile = tune_evp(g, exact=False, n_substeps=100)
        res[f"{tag}_evp_tile"] = list(_native.evp_tile_shapes()[tile])
        seaice_step(start.copy(lambda a: a.clone()), q, s_, n_substeps=nsub)  # capture
        out[tag] = g
        g2 = start.copy(lambda a: a.clone())
        t = _sync_time(lambda: seaice_step(g2, q, s_, n_substeps=nsub), 3)
        res[f"{tag}_ms_per_step"] = t * 1e3
        res[f"{tag}_value"] = nx * ny * nsub / t
    res["unit"] = "cell-substeps/s (full step)"
    names = ("uIce", "vIce", "sigma1", "sigma2", "sigma12", "hIceMean", "hSnowMean", "Area",
             "TSurf", "Qnet", "EmPmR", "forc_salt_surface", "OceanStressU", "OceanStressV")
    m = out["fp64"].variables.maskInC
    res["fp32_rel_l2_vs_fp64"] = {
        k: _rel_l2([(out["fp32"].variables[k], out["fp64"].variables[k], m)]) for k in names
        if k in out["fp64"].variables.keys()}
    # the same by latitude band: with a one-day step the polar rows run at
    # advective CFL numbers above 1 and the errors grow poleward
    ny_pad = m.shape[1]
    lat_rows = -90.0 + (torch.arange(ny_pad, device=dev, dtype=torch.float64) - 1.5) * 180.0 / ny
    bands = {}
    for lo, hi in ((0, 60), (60, 75), (75, 85), (85, 90)):
        sel = ((lat_rows.abs() >= lo) & (lat_rows.abs() < hi)).to(m.dtype)[None, :] * m
        bands[f"{lo}-{hi}"] = {k: _rel_l2([(out["fp32"].variables[k], out["fp64"].variables[k],
                                            sel)]) for k in ("uIce", "vIce", "hIceMean", "Area")}
    res["fp32_rel_l2_vs_fp64_by_abs_lat"] = bands
    u = out["fp64"].variables.uIce
    res["max_advective_cfl"] = float((u.abs() * float(st.settings.deltatDyn) /
                                      out["fp64"].variables.dxU)[2:-2, 2:-2].max())
    budgets = {}
    for k in ("hIceMean", "hSnowMean", "Area"):
        b = [float(K.global_sum(out[t], out[t].variables[k], out[t].variables.rA))
             for t in ("fp32", "fp64")]
        budgets[k] = abs(b[0] - b[1]) / max(abs(b[1]), 1e-300)
    res["fp64_budget_rel_diff"] = budgets
    res["shard_8way"] = bench_shard(dev, st, out)

    # batched adjoint: 8 members at 240 x 240, FP32, all members in one launch
    # sequence forward and backward; the forward it is divided by is the
    # batched drop-in (one pre-EVP launch + the batched subcycle), timed the
    # same way on the same stream
    from veris_amd.batched import Ensemble

    ens = Ensemble([synth.config4_adjoint_member(mm) for mm in range(8)], dev, torch.float32)
    n_adj = int(ens.settings.nEVPsteps)
    ps = torch.tensor(float(ens.settings.pStar), dtype=torch.float64, device=dev)

    def batched_adjoint(ckpt):
        h = ens.fields["hIceMean"].detach().clone().requires_grad_(True)
        p_ = ps.clone().requires_grad_(True)
        outs = A.evp_subcycle_batched(ens, pStar=p_, hIceMean=h, n_substeps=n_adj, ckpt_every=ckpt)
        ((outs[0].double() ** 2).sum() + (outs[1].double() ** 2).sum()).backward()

    def batched_ckpt_forward(ckpt):
        with torch.no_grad():
            A.evp_subcycle_batched(ens, n_substeps=n_adj, ckpt_every=ckpt)

    out_b = None
    from veris_amd.autotune import tune_batched

    tile_b, _ = tune_batched(ens, exact=False)  # closed members: every shape is exact

    def batched_forward():
        nonlocal out_b
        ens.pre_evp()
        out_b = ens.evp(n_adj, tile=tile_b, out=out_b)

    # every state kept (one segment: nothing recomputed; 8 x 244 x 244 FP32
    # states take 9.5 MB each, 1.1 GB for the subcycle) and a checkpoint every
    # 10 substeps (each segment recomputed from its checkpoint)
    t = {}
    for ckpt in (n_adj, 10):
        for fn in (lambda: batched_adjoint(ckpt), lambda: batched_ckpt_forward(ckpt)):
            fn()
        t[ckpt] = (_sync_time(lambda: batched_adjoint(ckpt), 5),
                   _sync_time(lambda: batched_ckpt_forward(ckpt), 5))
    batched_forward()
    t_fwd = _sync_time(batched_forward, 5)
    t_adj, t_ck = t[n_adj]
    res["batched_adjoint"] = {
        "workload": f"8 regional members 240x240 (10-cell thickness noise), FP32, {n_adj} EVP "
                    "substeps, gradient of a velocity norm w.r.t. P* and hIceMean; forward "
                    "and reverse each one batched launch sequence; every state kept (1.1 GB "
                    "of HBM, nothing recomputed), and beside it a checkpoint every 10 "
                    "substeps",
        "forward_backward_ms": t_adj * 1e3, "checkpointed_forward_ms": t_ck * 1e3,
        "plain_forward_ms": t_fwd * 1e3, "plain_forward_tile": tile_b,
        "adjoint_over_plain_forward": t_adj / t_fwd,
        "adjoint_over_checkpointed_forward": t_adj / t_ck,
        "ckpt10": {"forward_backward_ms": t[10][0] * 1e3, "checkpointed_forward_ms": t[10][1] * 1e3,
                   "adjoint_over_plain_forward": t[10][0] / t_fwd}}
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--substeps", type=int, default=400)
    ap.add_argument("--tile", type=int, default=0,
                    help="EVP workgroup shape index (default 0: autotune)")
    ap.add_argument("--exact", action="store_true",
                    help="autotune among the single-substep shapes only (every padded cell "
                         "bit for bit, the x-overlap stresses included)")
    ap.add_argument("--general", action="store_true",
                    help="launch the general substep instantiation (settings branches "
                         "evaluated at run time) instead of the default-settings one")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched, adjoint and 0.25-degree measurements")
    ap.add_argument("--only", choices=("batched", "adjoint", "hires", "variants"),
                    help="run one of the extra workloads alone and print its JSON object "
                         "(for profiling)")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="first copy a 64 MiB buffer 5 times (a kernel of known HBM traffic, "
                         "the reference for rocprofv3 FETCH_SIZE / WRITE_SIZE counts)")
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: every box in this setup has a single MI355X; "
              "run with --gpus 1", file=sys.stderr)
        return 2

    if args.general:
        os.environ["VERIS_EVP_PLAIN"] = "0"
    import torch

    from veris_amd import kernels as K
    from veris_amd import synth
    from veris_amd.step import seaice_step

    dev = torch.device("cuda:0")
    if args.only:
        if args.only == "batched":
            res = bench_batched(dev, args.substeps, float("nan"), float("nan"))
        elif args.only == "adjoint":
            res = bench_adjoint(dev)
        elif args.only == "variants":
            res = bench_variants(dev, args.substeps)
        else:
            res = bench_hires(dev)
        print(json.dumps({args.only: res}), flush=True)
        return 0
    if args.calibration_copy:
        src = torch.ones(8 << 20, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    st = synth.config1()
    nx, ny = st.settings.nx, st.settings.ny
    q0 = torch.from_numpy(st.variables.Qnet.copy()).to(dev)
    sw = torch.from_numpy(st.variables.SWdown.copy()).to(dev)
    g = st.to(dev)
    nsub = args.substeps

    # one step first (pre-EVP fields in place), then pick the fastest EVP
    # workgroup shape on this state: among every shape (multi-substep shapes
    # give every velocity and every interior stress bit for bit, and periodic
    # values in the stresses of the four x-overlap columns, where the
    # reference holds numpy.roll wrap-around values nothing else reads), or
    # with --exact among the single-substep shapes only (every padded cell)
    from veris_amd import _native
    from veris_amd.autotune import tune_evp

    seaice_step(g, q0, sw, n_substeps=nsub)
    if args.tile > 0:
        tile = args.tile
        K.EVP_TILE[K.evp_tile_key(g)] = tile
    else:
        tile = tune_evp(g, exact=args.exact, n_substeps=100)
    ri, rj, kk = _native.evp_tile_shapes()[tile]
    n_launch = nsub // kk + nsub % kk
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
    per_substep_s = evp_ms * 1e-3 / (args.steps * nsub)
    per_launch_s = evp_ms * 1e-3 / (args.steps * n_launch)
    cells = nx * ny
    value = cells * nsub * args.steps / wall
    # algorithmic bytes: 38 words per interior cell and substep, kk substeps per launch
    bytes_per_launch = B_ALG_WORDS * 8 * cells * kk
    achieved = B_ALG_WORDS * 8 * cells / per_substep_s / 1e9
    finite = all(bool(torch.isfinite(g.variables[k]).all()) for k in ("uIce", "hIceMean", "Qnet"))
    kernel = evp_kernel_name(ri, rj, kk, plain=not args.general)

    line = {
        # the benchmark's metric; `value` is its cell-substeps/s leg on the
        # workload it is quoted on (the plain FP64 drop-in, 1-degree full step);
        # the two ratios it names are under "ratios"
        "metric": "EVP cell-substeps/s at 1\u00b0 (FP32 batched vs plain FP64); adjoint/forward ratio",
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
                   "nx": nx, "ny": ny, "evp_substeps": nsub, "parallelism": "none (1 GPU)",
                   "evp_launch_shape": {"tile": tile, "threads": [ri, rj],
                                        "substeps_per_launch": kk},
                   "parity": ("bit-exact on every padded cell" if kk == 1 else
                              "bit-exact velocities on every cell and stresses on every cell "
                              "but the 4 x-overlap columns (periodic values there instead of "
                              "the reference's numpy.roll wrap-around values)")},
        "roofline": {
            "bound": "hbm",
            "kernel": kernel,
            "tile": tile,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(kernel),
            "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT),
            "avg_launch_us": per_launch_s * 1e6,
            "substeps_per_launch": kk,
            "frac_basis": "HIP-event time per launch over the timed region on the launch "
                          "stream: kernel plus the gap to the next dependent launch",
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
        "finite": finite,
    }
    kavg = _profile_value(KERNEL_PROFILE_FILE, "kernel_avg_us_rocprof", kernel)
    if kavg:  # the kernel's own duration, from the committed rocprofv3 kernel trace
        line["roofline"]["kernel_only"] = {
            "avg_kernel_us": kavg, "achieved": bytes_per_launch / (kavg * 1e-6) / 1e9,
            "frac": bytes_per_launch / (kavg * 1e-6) / 1e9 / HBM_PEAK_GBS,
            "source": os.path.relpath(KERNEL_PROFILE_FILE, ROOT)}
    def guarded(fn, *a):
        # an extra measurement that fails is reported in its place; the
        # headline figures above still print
        try:
            return fn(*a)
        except Exception as e:  # noqa: BLE001
            print(f"bench: {fn.__name__} failed: {e!r}", file=sys.stderr, flush=True)
            return {"error": repr(e)}

    if not args.no_extras:
        evp_rate = cells / per_substep_s
        # the small adjoint workload first: right after the bandwidth-heavy
        # ensemble runs its FP64 timing read about 10% slower
        adjoint = guarded(bench_adjoint, dev)
        line["batched"] = guarded(bench_batched, dev, nsub, evp_rate, value)
        line["adjoint"] = adjoint
        line["hires"] = guarded(bench_hires, dev)
        line["variants"] = guarded(bench_variants, dev, nsub)
        b, a = line["batched"], line["adjoint"]
        if "error" not in b and "error" not in a:
            line["ratios"] = {
                "fp32_batched_over_plain_fp64_evp": b["fp32_over_plain_fp64_evp_1deg"],
                "fp32_batched_target": b["target_ratio"],
                "adjoint_over_forward_fp64": a["fp64"]["adjoint_over_plain_forward"],
                "adjoint_over_forward_fp32": a["fp32"]["adjoint_over_plain_forward"],
                "adjoint_over_forward_own_loop_fp64": a["fp64"]["adjoint_over_plain_forward_own_loop"],
                "adjoint_over_forward_own_loop_fp32": a["fp32"]["adjoint_over_plain_forward_own_loop"],
                "adjoint_target": a["fp64"]["target"],
            }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = guarded(cpu_baseline, nsub)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
