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
