"""Shared helpers of the GPU parity tests (imported by tests only)."""

from __future__ import annotations

import sys

import numpy as np

from oracle import veris_np as O


def pre_evp(st):
    """Oracle pre-EVP fields of a NumPy state (in place)."""
    vs = st.variables
    vs.SeaIceMassC, vs.SeaIceMassU, vs.SeaIceMassV = O.SeaIceMass(st)
    vs.AreaW, vs.AreaS = O.AreaWS(st)
    vs.WindForcingX, vs.WindForcingY = O.WindForcingXY(st)
    vs.SeaIceStrength = O.SeaIceStrength(st)
    return st


def g2n(t):
    return t.detach().to("cpu").double().numpy()


def report(name, got, want):
    d = np.abs(got - want)
    scale = max(float(np.max(np.abs(want))), 1e-300)
    nbad = int(np.count_nonzero(got != want))
    msg = (f"{name}: max|diff|={float(d.max()):.3e} rel={float(d.max()) / scale:.3e} "
           f"unequal={nbad}/{got.size}")
    print(msg, file=sys.stderr, flush=True)
    return float(d.max()) / scale, nbad


def assert_exact(name, got, want):
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, name
    assert np.isfinite(got).all() == np.isfinite(want).all(), f"{name}: finiteness differs"
    _, nbad = report(name, got, want)
    assert nbad == 0, f"{name}: {nbad} cells differ from the oracle"


def assert_close(name, got, want, rtol):
    """max|got - want| <= rtol * max|want| over the array."""
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, name
    assert np.isfinite(got).all(), f"{name}: non-finite values"
    rel, _ = report(name, got, want)
    assert rel <= rtol, f"{name}: relative error {rel:.3e} > {rtol:.1e}"


def round_to_fp32(st):
    """The state with every FP64 field rounded to the nearest FP32 value
    (kept in FP64), in place."""
    vs = st.variables
    for k in list(vs.keys()):
        a = vs[k]
        if isinstance(a, np.ndarray) and a.dtype == np.float64:
            vs[k] = a.astype(np.float32).astype(np.float64)
    return st


def _ensemble_member_oracle(args):
    """Worker: pre-EVP fields and the EVP subcycle of one member of the
    ragged ensemble through the oracle (run in a child process)."""
    from veris_amd import synth

    m, n, rounded = args
    st = synth.config2_member(m)
    if rounded:
        round_to_fp32(st)
    pre_evp(st)
    return O.evp_solver(st, n_substeps=n)


def ensemble_oracle(members, n, rounded=False, workers=12):
    """Oracle EVP outputs of the ragged-ensemble members ``members`` after
    ``n`` substeps from rest, computed in parallel CPU processes (spawned, so
    no GPU state is inherited).  Returns one 5-tuple per member."""
    import multiprocessing as mp
    import os

    from veris_amd import synth

    sizes = synth.config2_sizes()
    order = sorted(range(len(members)), key=lambda q: -sizes[members[q]][0] * sizes[members[q]][1])
    workers = max(1, min(workers, len(members), len(os.sched_getaffinity(0))))
    ctx = mp.get_context("spawn")
    out = [None] * len(members)
    with ctx.Pool(workers) as pool:  # largest members first
        res = pool.map(_ensemble_member_oracle, [(members[q], n, rounded) for q in order],
                       chunksize=1)
    for q, r in zip(order, res):
        out[q] = r
    return out


def evp_shapes_from_source():
    """{tile index: (ri, rj, k)} of the EVP launch-shape list in
    veris_amd/csrc/evp_kernel.h (the list the library's table and launch
    switches are generated from), read at collection time without loading
    the library."""
    import re
    from pathlib import Path

    src = (Path(__file__).resolve().parent.parent / "veris_amd" / "csrc" / "evp_kernel.h").read_text()
    body = re.search(r"#define VH_EVP_SHAPES\(X\)(.*?)\n\s*\n", src, re.S).group(1)
    rows = re.findall(r"X\((\d+),\s*(\d+),\s*(\d+),\s*(\d+)\)", body)
    return {int(a): (int(b), int(c), int(d)) for a, b, c, d in rows}
