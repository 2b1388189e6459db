"""GPU drop-in for ``veris/heat_flux_MITgcm.py``: the open-ocean bulk fluxes
that feed the ocean's surface forcing and the evaporation used by Growth."""

from __future__ import annotations

import ctypes

from . import _native as N
from .kernels import Ctx, _p


def bulkf_formula_lanl(state, uw, vw, ta, qa, tsf, ocn_mask):
    """veris/heat_flux_MITgcm.py:5-123 ->
    (flwupa, flha, fsha, df0dT, ust, vst, evp, ssq, dEvdT).

    Inputs are device tensors of the state's padded shape and dtype; gravity
    is the host model's ``settings.grav`` (Veros), or the plugin's
    ``gravity`` when the settings carry no ``grav``."""
    c = Ctx.of(state)
    for name, t in (("uw", uw), ("vw", vw), ("ta", ta), ("qa", qa), ("tsf", tsf),
                    ("ocn_mask", ocn_mask)):
        c._check(name, t)
    sett = state.settings
    grav = float(sett["grav"]) if "grav" in sett.keys() else float(sett.gravity)
    out = c.empty(len(N.BULK_FLUX_NAMES))
    o = N.BulkFluxOut(*[t.data_ptr() for t in out])
    N.call("vh_bulkf_formula_lanl", ctypes.byref(c.dom), ctypes.byref(c.sett), grav, _p(uw),
           _p(vw), _p(ta), _p(qa), _p(tsf), _p(ocn_mask), ctypes.byref(o), c.stream)
    return tuple(out)
