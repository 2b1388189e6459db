"""GPU drop-in for ``veris/heat_flux_MITgcm.py``: the open-ocean bulk fluxes
that feed the ocean's surface forcing and the evaporation used by Growth."""

from __future__ import annotations

import ctypes

from . import _native as N
from .kernels import Ctx


def _gravity(c: Ctx, sett) -> float:
    """The host model's ``settings.grav`` (Veros), or the plugin's
    ``gravity`` when the settings carry no ``grav``; kept on the call
    context, which is rebuilt whenever a setting is assigned."""
    g = getattr(c, "_grav", None)
    if g is None:
        g = c._grav = float(sett["grav"]) if "grav" in sett.keys() else float(sett.gravity)
    return g


def bulkf_formula_lanl(state, uw, vw, ta, qa, tsf, ocn_mask):
    """veris/heat_flux_MITgcm.py:5-123 ->
    (flwupa, flha, fsha, df0dT, ust, vst, evp, ssq, dEvdT).

    Inputs are device tensors of the state's padded shape and dtype."""
    c = Ctx.of(state)
    chk = c._check
    chk("uw", uw)
    chk("vw", vw)
    chk("ta", ta)
    chk("qa", qa)
    chk("tsf", tsf)
    chk("ocn_mask", ocn_mask)
    out, ptrs = c.empty_addressed(9)
    N.call("vh_bulkf_formula_lanl", ctypes.byref(c.dom), ctypes.byref(c.sett),
           _gravity(c, state.settings), uw.data_ptr(), vw.data_ptr(), ta.data_ptr(),
           qa.data_ptr(), tsf.data_ptr(), ocn_mask.data_ptr(), ctypes.byref(N.BulkFluxOut(*ptrs)),
           c.stream)
    return out
