"""Exchange with the host ocean around a sea-ice step, on the device.

The reference setup does these updates inline in its forcing kernel
(veris/setup/seaice_global_4deg/seaice_global_4deg.py:847-876 before the
sea-ice kernels, :933-940 after them); here each side is one fused launch,
so a whole coupled step stays resident on the GPU.
"""

from __future__ import annotations

import ctypes

from . import _native as N
from .kernels import Ctx, _p


def ocean_to_ice(state, u_surface, v_surface, temp_surface, salt_surface, forc_temp_surface):
    """Surface ocean fields -> (uOcean, vOcean, theta, ocSalt, Qnet, Qsw, wSpeed).

    ``u_surface`` ... ``salt_surface`` are the ocean's top level at the current
    time level (``u[:, :, -1, tau]`` etc., made contiguous), temperature in
    degrees Celsius; ``forc_temp_surface`` is the ocean's surface heat
    forcing.  ``Qsw`` and ``wSpeed`` come from ``state.variables.SWdown`` and
    the wind (taken before its overlap fill, as in the reference)."""
    c = Ctx.of(state)
    for name, t in (("u_surface", u_surface), ("v_surface", v_surface),
                    ("temp_surface", temp_surface), ("salt_surface", salt_surface),
                    ("forc_temp_surface", forc_temp_surface)):
        c._check(name, t)
    out = c.empty(len(N.OCEAN_TO_ICE_NAMES))
    o = N.OceanToIceOut(*[t.data_ptr() for t in out])
    N.call("vh_ocean_to_ice", *c.args(), _p(u_surface), _p(v_surface), _p(temp_surface),
           _p(salt_surface), _p(forc_temp_surface), ctypes.byref(o), c.stream)
    return tuple(out)


def ice_to_ocean(state, surface_taux, surface_tauy, forc_salt_surface_ice, forc_salt_surface_res):
    """-> (surface_taux, surface_tauy, forc_temp_surface, forc_salt_surface):
    the wind stress blended with OceanStressU/V by AreaW/AreaS, the heat
    forcing from Qnet and the salt forcing (ice flux + restoring)."""
    c = Ctx.of(state)
    for name, t in (("surface_taux", surface_taux), ("surface_tauy", surface_tauy),
                    ("forc_salt_surface_ice", forc_salt_surface_ice),
                    ("forc_salt_surface_res", forc_salt_surface_res)):
        c._check(name, t)
    out = c.empty(len(N.ICE_TO_OCEAN_NAMES))
    o = N.IceToOceanOut(*[t.data_ptr() for t in out])
    N.call("vh_ice_to_ocean", *c.args(), _p(surface_taux), _p(surface_tauy),
           _p(forc_salt_surface_ice), _p(forc_salt_surface_res), ctypes.byref(o), c.stream)
    return tuple(out)
