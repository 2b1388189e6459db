"""Small hand-built states for the oracle pin tests (imported by tests only)."""

from __future__ import annotations

import importlib.util
import os

import numpy as np

from veris_amd.state import VerisState, make_state_settings
from veris_amd.variables import VARIABLE_NAMES, VariablesNamespace


def uniform_state(nx=16, ny=12, dx=1.0e4, dy=1.5e4, cyclic=True, **overrides):
    """All-ocean uniform Cartesian grid: every mask 1 (halos included), spacings
    dx, dy, areas dx*dy (rAz follows the plugin's two-term formula, which is
    dx*dy/2 on a uniform grid), f = 0, everything else zero."""
    sett = make_state_settings(nx, ny, enable_cyclic_x=cyclic, **overrides)
    shape = (nx + 4, ny + 4)
    vs = VariablesNamespace({k: np.zeros(shape) for k in VARIABLE_NAMES})
    one = np.ones(shape)
    for k in ("iceMask", "iceMaskU", "iceMaskV", "maskInC", "maskInU", "maskInV"):
        vs[k] = one.copy()
    for k in ("dxC", "dxG", "dxU", "dxV"):
        vs[k] = one * dx
    for k in ("dyC", "dyG", "dyU", "dyV"):
        vs[k] = one * dy
    for k in ("rA", "rAu", "rAv"):
        vs[k] = one * dx * dy
    vs.rAz = one * (0.5 * dx * dy)
    for k in ("dxC", "dyC", "dxG", "dyG", "dxU", "dyU", "dxV", "dyV", "rA", "rAu", "rAv", "rAz"):
        vs["recip_" + k] = 1 / vs[k]
    vs.TSurf = one * 273.0
    vs.ocSalt = one * 34.7
    vs.surfPress = one * 1e5
    vs.wSpeed = one * 5.0
    vs.LWdown = one * 250.0
    vs.ATemp = one * 260.0
    vs.aqh = one * 2e-3
    vs.theta = one * 271.0
    return VerisState(sett, vs)


def xy(state):
    """Coordinates of u-points (x_i, y_j+1/2), v-points (x_i+1/2, y_j) and
    c-points for a uniform grid with spacing from dxU/dyV."""
    dx = state.variables.dxU[0, 0]
    dy = state.variables.dyV[0, 0]
    NX, NY = state.shape
    i = np.arange(NX)[:, None] * np.ones((1, NY))
    j = np.ones((NX, 1)) * np.arange(NY)[None, :]
    return i * dx, j * dy, dx, dy


def load_oracle():
    """The NumPy oracle module, or a planted-mistake copy of it when
    ORACLE_NP_PATH names one (used by test_oracle_mutations on the CPU)."""
    path = os.environ.get("ORACLE_NP_PATH")
    if not path:
        from oracle import veris_np

        return veris_np
    spec = importlib.util.spec_from_file_location("veris_np_under_test", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod
