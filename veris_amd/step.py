"""The sea-ice step on the GPU, in the order of the reference driver.

``seaice_global_4deg.py:874-932`` (inside ``set_forcing_kernel``) runs, after
copying the ocean heat flux and short-wave forcing into ``Qnet``/``Qsw``:
SeaIceMass, AreaWS, WindForcingXY, SeaIceStrength, IceVelocities,
OceanStressUV, Advection, clean_up_advection, ridging, Growth and five
overlap fills.  :func:`seaice_step` does the same with the GPU drop-ins and
keeps the stresses returned by IceVelocities in the state (the reference's
setup unpacks only two of the five values, so its stresses restart from zero
every step).
"""

from __future__ import annotations

from . import kernels as K

STEP_FILLS = ("hIceMean", "hSnowMean", "Area", "forc_salt_surface_ice", "Qnet")


def seaice_step(state, qnet_forcing=None, swdown=None, n_substeps=None):
    """Advance ``state`` by one model step in place; returns ``state``."""
    vs = state.variables
    if qnet_forcing is not None:
        vs.Qnet = qnet_forcing.clone()
    if swdown is not None:
        vs.Qsw = -swdown
    # SeaIceMass, AreaWS, WindForcingXY, SeaIceStrength in one fused launch
    (vs.SeaIceMassC, vs.SeaIceMassU, vs.SeaIceMassV, vs.AreaW, vs.AreaS, vs.WindForcingX,
     vs.WindForcingY, vs.SeaIceStrength) = K.pre_evp(state)
    vs.uIce, vs.vIce, vs.sigma1, vs.sigma2, vs.sigma12 = K.IceVelocities(state, n_substeps)
    vs.OceanStressU, vs.OceanStressV = K.OceanStressUV(state)
    vs.hIceMean, vs.hSnowMean, vs.Area = K.Advection(state)
    (vs.hIceMean, vs.hSnowMean, vs.Area, vs.TSurf,
     vs.os_hIceMean, vs.os_hSnowMean) = K.clean_up_advection(state)
    vs.Area = K.ridging(state)
    (vs.hIceMean, vs.hSnowMean, vs.Area, vs.TSurf, vs.EmPmR, vs.forc_salt_surface_ice,
     vs.Qsw, vs.Qnet, vs.SeaIceLoad, vs.IcePenetSW, vs.recip_hIceMean) = K.Growth(state)
    # the reference's fill_overlap of each (filled copies); the fields are the
    # fresh outputs of Growth above, so they are filled in place, in one launch
    K.fill_overlaps_(state, [vs[name] for name in STEP_FILLS])
    return state
