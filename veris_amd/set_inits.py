"""Derivation of sea-ice masks and C-grid metrics from a host ocean grid.

Follows ``veris/set_inits.py:5-52``: masks are the surface level of the ocean
masks (``:11-16``), ``dxC``/``dxU`` vary with ``i`` only and ``dyC``/``dyU``
with ``j`` only (``:22-25``), the staggered spacings are two-point averages
(``:26-29``) and ``rAz`` keeps the reference's two-term weighting
``0.25*(rA + rA[i-1])`` shifted by one in ``j`` (``:33-34``), which is half of a
four-point average.  Reciprocals follow ``:36-47``; ``TSurf`` starts at 273 K
(``:49``).

Runs once per setup on the host (NumPy); the result is moved to the device by
the caller.
"""

from __future__ import annotations

import numpy as np


def set_inits(state) -> None:
    """Plugin setup entry point: derive the sea-ice grid from the Veros state.

    Reads the host-model variables the reference reads (``maskT``, ``maskU``,
    ``maskV`` as 3-D arrays whose last level is the surface, ``dxt``, ``dxu``,
    ``dyt``, ``dyu``, ``area_t``, ``area_u``, ``area_v``, ``coriolis_t``,
    ``ht``) from ``state.variables`` and fills the sea-ice variables in place.
    """
    vs = state.variables
    grid = {k: np.asarray(vs[k]) for k in ("dxt", "dxu", "dyt", "dyu", "area_t", "area_u",
                                          "area_v", "coriolis_t", "ht")}
    for k in ("maskT", "maskU", "maskV"):
        grid[k] = np.asarray(vs[k])[:, :, -1]
    derive_grid(vs, grid)


def derive_grid(vs, ocean_grid: dict) -> None:
    """Fill ``vs`` (a variables namespace) from surface-level ``ocean_grid``.

    ``ocean_grid`` holds the host-model arrays the reference reads from Veros:
    ``maskT``, ``maskU``, ``maskV`` (surface level, 2-D), ``dxt``, ``dxu``
    (length nx+4), ``dyt``, ``dyu`` (length ny+4), ``area_t``, ``area_u``,
    ``area_v``, ``coriolis_t`` and ``ht`` (2-D).
    """
    g = ocean_grid
    vs.iceMask = np.array(g["maskT"], dtype=np.float64)
    vs.iceMaskU = np.array(g["maskU"], dtype=np.float64)
    vs.iceMaskV = np.array(g["maskV"], dtype=np.float64)
    vs.maskInC = vs.iceMask
    vs.maskInU = vs.iceMaskU
    vs.maskInV = vs.iceMaskV

    vs.R_low = np.array(g["ht"], dtype=np.float64)
    vs.fCori = np.array(g["coriolis_t"], dtype=np.float64)
    ones2d = np.ones_like(vs.maskInC)
    vs.dxC = ones2d * np.asarray(g["dxt"], dtype=np.float64)[:, None]
    vs.dyC = ones2d * np.asarray(g["dyt"], dtype=np.float64)
    vs.dxU = ones2d * np.asarray(g["dxu"], dtype=np.float64)[:, None]
    vs.dyU = ones2d * np.asarray(g["dyu"], dtype=np.float64)
    vs.dxG = 0.5 * (vs.dxU + np.roll(vs.dxU, 1, 1))
    vs.dyG = 0.5 * (vs.dyU + np.roll(vs.dyU, 1, 0))
    vs.dxV = 0.5 * (vs.dxC + np.roll(vs.dxC, 1, 1))
    vs.dyV = 0.5 * (vs.dyC + np.roll(vs.dyC, 1, 0))
    vs.rA = np.array(g["area_t"], dtype=np.float64)
    vs.rAu = np.array(g["area_u"], dtype=np.float64)
    vs.rAv = np.array(g["area_v"], dtype=np.float64)
    rAz = vs.rA + np.roll(vs.rA, 1, 0)
    vs.rAz = 0.25 * np.roll(rAz, 1, 1)

    for name in ("dxC", "dyC", "dxG", "dyG", "dxU", "dyU", "dxV", "dyV"):
        vs["recip_" + name] = 1 / vs[name]
    for name in ("rA", "rAu", "rAv", "rAz"):
        vs["recip_" + name] = 1 / vs[name]

    vs.TSurf = np.ones_like(vs.maskInC) * 273
