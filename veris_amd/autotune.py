"""Tile-shape autotuner of the EVP substep kernel.

The best workgroup shape depends on the domain (1 degree is latency bound at
about one wave per SIMD; ensembles of large members are bandwidth bound), so
it is measured rather than guessed: each candidate shape runs a short
subcycle on the caller's own state, timed with HIP events on the current
stream, and the fastest one is remembered per (shape, dtype) in
:data:`veris_amd.kernels.EVP_TILE` (per grid, dtype and settings class).

Exact shapes (one substep per launch) reproduce the reference on every cell
and are the default candidates; multi-substep shapes (``exact=False``) also
reproduce every velocity but return periodic values instead of the
reference's wrap-around artefacts in the stresses of the filled overlap
columns (rows).  They read each filled overlap through the interior cells the
fill copies from, so they are candidates only for states whose inputs hold
those values there (:func:`veris_amd.kernels.overlaps_periodic`; always true
without a filled overlap).
"""

from __future__ import annotations

import torch

from . import _native as N
from . import kernels as K


def candidate_tiles(exact: bool = True, settings=None, periodic: bool = True) -> list[int]:
    """Tile indices to measure: single-substep shapes, plus the multi-substep
    ones when ``exact`` is False, ``settings`` allow them and the inputs are
    ``periodic`` in their filled overlaps."""
    shapes = N.evp_tile_shapes()
    multi = not exact and periodic and (
        settings is None or not (settings.noSlip and settings.secondOrderBC))
    return [t for t, (_, _, k) in shapes.items() if (k == 1 or multi)]


def time_evp(state, tile: int, n_substeps: int = 100, repeats: int = 3) -> float:
    """Seconds per substep of ``tile`` on ``state`` (best of ``repeats``)."""
    stream = torch.cuda.current_stream()
    best = float("inf")
    K.evp_solver(state, n_substeps, tile=tile)  # warm-up (and graph capture)
    for _ in range(repeats):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        K.evp_solver(state, n_substeps, tile=tile)
        b.record(stream)
        b.synchronize()
        best = min(best, a.elapsed_time(b) * 1e-3 / n_substeps)
    return best


def tune_evp(state, exact: bool = True, n_substeps: int = 100, verbose: bool = False) -> int:
    """Measure the candidate shapes on ``state``, remember and return the best."""
    times = {}
    periodic = exact or K.overlaps_periodic(state)
    for t in candidate_tiles(exact, state.settings, periodic):
        times[t] = time_evp(state, t, n_substeps)
        if verbose:
            print(f"tile {t} {N.evp_tile_shapes()[t]}: {times[t] * 1e6:.2f} us/substep")
    best = min(times, key=times.get)
    K.EVP_TILE[K.evp_tile_key(state)] = best
    return best


def time_batched(ens, tile: int, n_substeps: int = 40, repeats: int = 3) -> float:
    """Seconds per substep of ``tile`` on the ensemble ``ens`` (best of ``repeats``)."""
    stream = torch.cuda.current_stream()
    out = ens.evp(n_substeps, tile=tile)  # warm-up (and graph capture)
    best = float("inf")
    for _ in range(repeats):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        ens.evp(n_substeps, tile=tile, out=out)
        b.record(stream)
        b.synchronize()
        best = min(best, a.elapsed_time(b) * 1e-3 / n_substeps)
    return best


def tune_batched(ens, exact: bool = True, n_substeps: int = 40, verbose: bool = False):
    """Measure the candidate shapes on an :class:`~veris_amd.batched.Ensemble`
    and return (best tile, {tile: seconds per substep}).  Multi-substep shapes
    are exact for members without a filled overlap, so an ensemble of closed
    regional members may pass ``exact=False`` without changing any result;
    with ``exact=False`` an ensemble with a filled member measures only
    single-substep shapes."""
    times = {}
    fill_free = ens.fill_mode == 0 and not any(ens.cyclic)
    for t in candidate_tiles(exact, ens.settings, fill_free):
        times[t] = time_batched(ens, t, n_substeps)
        if verbose:
            print(f"tile {t} {N.evp_tile_shapes()[t]}: {times[t] * 1e6:.2f} us/substep")
    best = min(times, key=times.get)
    return best, times
