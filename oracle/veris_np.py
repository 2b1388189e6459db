"""NumPy FP64 oracle of the Veris sea-ice step -- TEST INFRASTRUCTURE ONLY.

This module restates the reference algorithm (``veris/*.py``) on the CPU so
that the HIP product path can be checked against it.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker; the product never routes through it.

Conventions (identical to the reference):
* arrays are padded ``(nx+2*olx, ny+2*oly)``; axis 0 = x (``i``), axis 1 = y
  (``j``);
* neighbour access is ``numpy.roll`` over the *padded* array, so halo cells
  hold the same wrapped values the reference computes;
* every expression keeps the reference's association order, so an IEEE
  evaluation that follows the same order gives the same bits.

Pins (see ``tests/test_oracle_*.py``): closed forms for strain rates,
viscosities, the stress relaxation, drag (water and lateral) and the
free-drift fixed point;
the telescoping invariant of ``stressdiv``; exact one-cell upwind shift,
conservation and the TVD bound of the advection; the scipy root of the
surface energy balance for ``solve4temp``; budget identities of ``Growth``;
and, for the EVP subcycle and free drift, an independent C scalar-loop
restatement with explicit wrapped indices (``oracle/c/``, loaded by
``oracle/c_ref.py``) that must agree with this module bit for bit wherever
only +, -, *, /, sqrt and comparisons enter.
"""

from __future__ import annotations

import numpy as np

# --------------------------------------------------------------------------
# neighbour shifts over the padded array (np.roll semantics)
# --------------------------------------------------------------------------


def im1(a):  # a[i-1, j]
    return np.roll(a, 1, 0)


def ip1(a):  # a[i+1, j]
    return np.roll(a, -1, 0)


def im2(a):  # a[i-2, j]
    return np.roll(a, 2, 0)


def jm2(a):  # a[i, j-2]
    return np.roll(a, 2, 1)


def jm1(a):  # a[i, j-1]
    return np.roll(a, 1, 1)


def jp1(a):  # a[i, j+1]
    return np.roll(a, -1, 1)


def _vs(state):
    return state.variables, state.settings


# --------------------------------------------------------------------------
# halo fill and global sum
# --------------------------------------------------------------------------


def fill_overlap(state, a):
    """``veris/fill_overlap.py:6-18``.

    ``veros_fill`` (default) reproduces Veros' single-process
    ``enforce_boundaries``: x-periodic copy of two columns when the host grid is
    x-cyclic, otherwise nothing.  ``veros_fill=False`` is the plugin's own
    doubly periodic copy (``:12-17``), x first, then y.
    """
    s = state.settings
    a = np.array(a, copy=True)
    if s.veros_fill:
        if s.enable_cyclic_x:
            a[-2:] = a[2:4]
            a[:2] = a[-4:-2]
        return a
    ox, oy = s.olx, s.oly
    a[:ox, :] = a[-2 * ox: -ox, :]
    a[-ox:, :] = a[ox: 2 * ox, :]
    a[:, :oy] = a[:, -2 * oy: -oy]
    a[:, -oy:] = a[:, oy: 2 * oy]
    return a


def fill_overlap_uv(state, u, v):
    """``veris/fill_overlap.py:21-23``."""
    return fill_overlap(state, u), fill_overlap(state, v)


def global_sum(a):
    """``veris/global_sum.py:1-4`` is an identity stub; the oracle's budget sum is
    an exactly rounded FP64 sum (``math.fsum``) for checking device reductions."""
    import math

    return math.fsum(np.asarray(a, dtype=np.float64).ravel().tolist())


# --------------------------------------------------------------------------
# pre-EVP: mass, area, strength, wind forcing
# --------------------------------------------------------------------------


def SeaIceMass(state):
    """``veris/area_mass.py:17-28``."""
    vs, s = _vs(state)
    mC = s.rhoIce * vs.hIceMean + s.rhoSnow * vs.hSnowMean
    mU = 0.5 * (mC + im1(mC))
    mV = 0.5 * (mC + jm1(mC))
    return mC, mU, mV


def AreaWS(state):
    """``veris/area_mass.py:5-14``."""
    vs, _ = _vs(state)
    return 0.5 * (vs.Area + im1(vs.Area)), 0.5 * (vs.Area + jm1(vs.Area))


def SeaIceStrength(state, pStar=None):
    """``veris/dynamics_routines.py:5-18``; ``pStar`` may be a per-cell field."""
    vs, s = _vs(state)
    ps = s.pStar if pStar is None else pStar
    return ps * vs.hIceMean * np.exp(-s.cStar * (1 - vs.Area)) * vs.iceMask


def tauXY(state, airIceDrag=None, airIceDrag_south=None):
    """``veris/dynsolver.py:7-47``; the drag scalars may be overridden."""
    vs, s = _vs(state)
    ca_n = s.airIceDrag if airIceDrag is None else airIceDrag
    ca_s = s.airIceDrag_south if airIceDrag_south is None else airIceDrag_south
    sinWin = np.sin(np.deg2rad(s.airTurnAngle))
    cosWin = np.cos(np.deg2rad(s.airTurnAngle))
    if s.useRelativeWind:
        urel = vs.uWind - 0.5 * (vs.uIce + ip1(vs.uIce))
        vrel = vs.vWind - 0.5 * (vs.vIce + jp1(vs.vIce))
    else:
        urel = vs.uWind
        vrel = vs.vWind
    wsq = urel**2 + vrel**2
    ws = np.where(wsq < s.wSpeedMin**2, s.wSpeedMin, np.sqrt(wsq))
    cda = np.where(vs.fCori < 0, ca_s, ca_n) * s.rhoAir * ws
    sg = np.sign(vs.fCori)
    tx = cda * (cosWin * urel - sg * sinWin * vrel)
    ty = cda * (cosWin * vrel + sg * sinWin * urel)
    tx = 0.5 * (tx + im1(tx)) * vs.iceMaskU
    ty = 0.5 * (ty + jm1(ty)) * vs.iceMaskV
    return tx, ty


def WindForcingXY(state, **drag):
    """``veris/dynsolver.py:50-84``."""
    vs, s = _vs(state)
    tx, ty = tauXY(state, **drag)
    fx = tx * vs.AreaW
    fy = ty * vs.AreaS
    phi = s.gravity * vs.ssh_an
    if s.useRealFreshWaterFlux:
        phi = phi + (vs.surfPress + vs.SeaIceLoad * s.gravity * s.seaIceLoadFac) * s.recip_rhoSea
    else:
        phi = phi + vs.surfPress * s.recip_rhoSea
    fx = fx - vs.SeaIceMassU * vs.recip_dxC * (phi - im1(phi))
    fy = fy - vs.SeaIceMassV * vs.recip_dyC * (phi - jm1(phi))
    return fx, fy


# --------------------------------------------------------------------------
# EVP building blocks
# --------------------------------------------------------------------------


def ocean_drag_coeffs(state, u, v, waterIceDrag=None, waterIceDrag_south=None):
    """``veris/dynamics_routines.py:21-52``."""
    vs, s = _vs(state)
    cw_n = s.waterIceDrag if waterIceDrag is None else waterIceDrag
    cw_s = s.waterIceDrag_south if waterIceDrag_south is None else waterIceDrag_south
    D = np.where(vs.fCori < 0, cw_s, cw_n) * s.rhoSea
    du = (u - vs.uOcean) * vs.maskInU
    dv = (v - vs.vOcean) * vs.maskInV
    w = 0.25 * ((du + ip1(du)) ** 2 + (dv + jp1(dv)) ** 2)
    with np.errstate(invalid="ignore"):
        cd = np.where(D**2 * w > s.cDragMin**2, D * np.sqrt(w), s.cDragMin)
    return cd * vs.iceMask


def basal_drag_coeffs(state, u, v):
    """``veris/dynamics_routines.py:55-96``."""
    vs, s = _vs(state)
    um = u * vs.maskInU
    vm = v * vs.maskInV
    t = 0.25 * (um**2 + ip1(um) ** 2 + vm**2 + jp1(vm) ** 2)
    t = s.basalDragK2 / np.sqrt(t + s.basalDragU0**2)
    hcrit = np.abs(vs.R_low) * vs.Area / s.basalDragK1
    fac = 10.0
    rfac = 1.0 / fac
    with np.errstate(over="ignore", invalid="ignore"):
        val = (t * np.log(np.exp(fac * (vs.hIceMean - hcrit)) + 1.0) * rfac
               * np.exp(-s.cBasalStar * (1.0 - vs.Area)))
    return np.where(vs.Area > 0.01, val, 0.0)


def strainrates(state, u, v):
    """``veris/dynamics_routines.py:143-217``.

    The second-order boundary branch (``:177-215``) is restated as the
    reference writes it: its one-sided differences shift u along x and v
    along y (axis 0 for u, axis 1 for v) and scale by recip_dyU/recip_dxV."""
    vs, s = _vs(state)
    dudx = (ip1(u) - u) * vs.recip_dxU
    uave = (ip1(u) + u) * 0.5
    dvdy = (jp1(v) - v) * vs.recip_dyV
    vave = (jp1(v) + v) * 0.5
    e11 = (dudx + vave * vs.k2AtC) * vs.maskInC
    e22 = (dvdy + uave * vs.k1AtC) * vs.maskInC
    dudy = (u - jm1(u)) * vs.recip_dyU
    uavz = (u + jm1(u)) * 0.5
    dvdx = (v - im1(v)) * vs.recip_dxV
    vavz = (v + im1(v)) * 0.5
    mz = vs.iceMask * im1(vs.iceMask)
    mz = mz * jm1(mz)
    e12 = 0.5 * (dudy + dvdx - vs.k1AtZ * vavz - vs.k2AtZ * uavz) * mz
    if s.noSlip:
        hu = vs.iceMaskU - jm1(vs.iceMaskU)
        hv = vs.iceMaskV - im1(vs.iceMaskV)
        e12 = e12 + (2.0 * uavz * vs.recip_dyU * hu + 2.0 * vavz * vs.recip_dxV * hv)
    if s.noSlip and s.secondOrderBC:
        MU, MV = vs.iceMaskU, vs.iceMaskV
        hu2 = (MU - im1(MU)) / 3.0
        hv2 = (MV - jm1(MV)) / 3.0
        hu2 = hu2 * (im2(MU) * im1(MU) + ip1(MU) * MU)
        hv2 = hv2 * (jm2(MV) * jm1(MV) + jp1(MV) * MV)
        e12 = e12 + 0.5 * (
            vs.recip_dyU * (6.0 * uavz - im2(u) * im1(MU) - ip1(u) * MU) * hu2
            + vs.recip_dxV * (6.0 * vavz - jm2(v) * jm1(MV) - jp1(v) * MV) * hv2)
    return e11, e22, e12


def side_drag(state, u, v):
    """``veris/dynamics_routines.py:98-140``: lateral drag coefficients at
    u- and v-points (land neighbours along the face normal, or the coastline
    form factors Fu/Fv)."""
    vs, s = _vs(state)
    speed = 0.5 * np.sqrt((u + ip1(u)) ** 2 + (v + jp1(v)) ** 2)
    speedU = np.where(vs.AreaW > 0, 0.5 * (speed + im1(speed)), 0)
    speedV = np.where(vs.AreaS > 0, 0.5 * (speed + jm1(speed)), 0)
    maskU = (2 - im1(vs.iceMaskU) - ip1(vs.iceMaskU)) * vs.iceMaskU
    maskV = (2 - jm1(vs.iceMaskV) - jp1(vs.iceMaskV)) * vs.iceMaskV
    if s.use_coastline:
        maskU = vs.Fu
        maskV = vs.Fv
    sdU = vs.SeaIceMassU * s.sideDragCoeff * maskU / (speedU + s.sideDragU0)
    sdV = vs.SeaIceMassV * s.sideDragCoeff * maskV / (speedV + s.sideDragU0)
    return sdU, sdV


def viscosities(state, e11, e22, e12, P=None):
    """``veris/dynamics_routines.py:220-268``; returns (zeta, eta, press, deltaC)."""
    vs, s = _vs(state)
    P = vs.SeaIceStrength if P is None else P
    rpd = 1.0 / s.PlasDefCoeff**2
    q = vs.rAz * e12**2
    q = q + ip1(q)
    e12c = 0.25 * vs.recip_rA * (q + jp1(q))
    dsq = (e11 + e22) ** 2 + rpd * ((e11 - e22) ** 2 + 4.0 * e12c)
    dC = np.sqrt(dsq)
    dreg = dC + s.deltaMin
    zeta = 0.5 * (P * (1 + s.tensileStrFac)) / dreg
    eta = zeta * rpd
    press = (1 * (P * (1 - s.pressReplFac)
                  + 2.0 * zeta * dC * s.pressReplFac / (1 + s.tensileStrFac))
             * (1 - s.tensileStrFac))
    return zeta, eta, press, dC


def c_point_to_z_point(state, c):
    """``veris/averaging.py:4-19`` with the default (always used) normaliser."""
    vs, _ = _vs(state)
    n = vs.iceMask + im1(vs.iceMask)
    n = n + jm1(n)
    with np.errstate(divide="ignore"):
        n = np.where(n > 0, 1.0 / n, 0.0)
    z = c + im1(c)
    return n * (z + jm1(z))


def stressdiv(state, s11, s22, s12):
    """``veris/dynamics_routines.py:284-302``."""
    vs, _ = _vs(state)
    a = s11 * vs.dyV
    b = s12 * vs.dxV
    dx = (a - im1(a) - b + jp1(b)) * vs.recip_rAu
    c = s22 * vs.dxU
    d = s12 * vs.dyU
    dy = (c - jm1(c) - d + ip1(d)) * vs.recip_rAv
    return dx, dy


# --------------------------------------------------------------------------
# EVP subcycle
# --------------------------------------------------------------------------


def evp_cfac(state):
    """``veris/evp_solver.py:266-271``: factor of the adaptive relaxation parameter."""
    _, s = _vs(state)
    if not s.useAdaptiveEVP:
        return 0
    return s.deltatDyn * 4 * (np.pi * s.aEvpCoeff) ** 2


def evp_substep(state, u, v, uNm1, vNm1, s1, s2, s12, P=None, drag=None, trace=None):
    """One pass of ``veris/evp_solver.py:22-254``.

    Returns the new (u, v, sigma1, sigma2, sigma12).  ``drag`` overrides the
    water drag scalars; ``trace`` (a dict) receives intermediates for tests.
    With useAdaptiveEVP the relaxation parameters follow ``:65-73`` (c-points),
    ``:97-100`` (z-points) and ``:148-150`` (u/v points).
    """
    vs, s = _vs(state)
    drag = drag or {}
    e11, e22, e12 = strainrates(state, u, v)
    zeta, _eta, press, dC = viscosities(state, e11, e22, e12, P=P)
    if s.useAdaptiveEVP:
        alphaC = np.sqrt(zeta * evp_cfac(state) / np.maximum(vs.SeaIceMassC, 1e-4)
                         * vs.recip_rA) * vs.iceMask
        alphaC = np.maximum(alphaC, s.aEVPalphaMin)
        den1 = 1.0 / alphaC
        den2 = den1
    else:
        alphaC = s.evpAlpha
        den1 = 1 / s.evpAlpha
        den2 = den1
    rev = 1
    rrev = 0.25
    ep = e11 + e22
    em = e11 - e22
    divergence = 2 * zeta * ep - press
    tension = 2 * zeta * em
    shear = 2 * c_point_to_z_point(state, zeta) * e12
    s1 = (s1 * (alphaC - rev) + divergence) * den1 * vs.iceMask
    s2 = (s2 * (alphaC - rev) + tension * rrev) * den2 * vs.iceMask
    s11 = 0.5 * (s1 + s2)
    s22 = 0.5 * (s1 - s2)
    if s.useAdaptiveEVP:
        alphaZ = 0.5 * (alphaC + jm1(alphaC))
        alphaZ = 0.5 * (alphaZ + im1(alphaZ))
        den2 = 1.0 / alphaZ
    else:
        alphaZ = s.evpAlpha
    s12 = (s12 * (alphaZ - rev) + shear * rrev) * den2
    divx, divy = stressdiv(state, s11, s22, s12)
    cd = ocean_drag_coeffs(state, u, v, **drag)
    cb = basal_drag_coeffs(state, u, v)
    lmU = np.where(vs.SeaIceMassU != 0, 1, vs.SeaIceMassU)
    lmV = np.where(vs.SeaIceMassV != 0, 1, vs.SeaIceMassV)
    sg = np.sign(vs.fCori)
    du0 = vs.uOcean - u
    dv0 = vs.vOcean - v
    fx = vs.WindForcingX + (
        0.5 * (cd + im1(cd)) * s.cosWat * vs.uOcean
        - sg * s.sinWat * 0.5 * (
            cd * 0.5 * (du0 + jp1(du0))
            + im1(cd) * 0.5 * (im1(du0) + jp1(im1(du0)))
        ) * lmU) * vs.AreaW
    fy = vs.WindForcingY + (
        0.5 * (cd + jm1(cd)) * s.cosWat * vs.vOcean
        + sg * s.sinWat * 0.5 * (
            cd * 0.5 * (dv0 + ip1(dv0))
            + jm1(cd) * 0.5 * (jm1(dv0) + ip1(jm1(dv0)))
        ) * lmV) * vs.AreaS
    fv = vs.SeaIceMassC * vs.fCori * 0.5 * (v + jp1(v))
    fu = vs.SeaIceMassC * vs.fCori * 0.5 * (u + ip1(u))
    fx = fx + 0.5 * (fv + im1(fv))
    fy = fy - 0.5 * (fu + jm1(fu))
    if s.useAdaptiveEVP:
        evpBetaU = 0.5 * (alphaC + im1(alphaC))
        evpBetaV = 0.5 * (alphaC + jm1(alphaC))
        betaU = evpBetaU * s.recip_deltatDyn
        betaV = evpBetaV * s.recip_deltatDyn
    else:
        evpBetaU = evpBetaV = s.evpBeta
        betaU = s.evpBeta * s.recip_deltatDyn
        betaV = betaU
    bp1U = betaU + s.recip_deltatDyn
    bp1V = betaV + s.recip_deltatDyn
    denU = vs.SeaIceMassU * bp1U + vs.AreaW * (
        0.5 * (cd + im1(cd)) * s.cosWat + 0.5 * (cb + im1(cb)))
    denV = vs.SeaIceMassV * bp1V + vs.AreaS * (
        0.5 * (cd + jm1(cd)) * s.cosWat + 0.5 * (cb + jm1(cb)))
    denU = np.where(denU == 0, 1, denU)
    denV = np.where(denV == 0, 1, denV)
    if not s.noSlip:  # veris/evp_solver.py:172-179
        sdU, sdV = side_drag(state, u, v)
        denU = denU + sdU
        denV = denV + sdV
    un = vs.iceMaskU * (betaU * vs.SeaIceMassU * u + vs.SeaIceMassU * s.recip_deltatDyn * uNm1
                        + fx + divx) / denU
    vn = vs.iceMaskV * (betaV * vs.SeaIceMassV * v + vs.SeaIceMassV * s.recip_deltatDyn * vNm1
                        + fy + divy) / denV
    un, vn = fill_overlap_uv(state, un, vn)
    if trace is not None:
        trace.update(e11=e11, e22=e22, e12=e12, zeta=zeta, press=press, deltaC=dC,
                     cDrag=cd, cBot=cb, divX=divx, divY=divy, ForcingX=fx, ForcingY=fy,
                     alphaC=alphaC, alphaZ=alphaZ, betaU=betaU, betaV=betaV,
                     evpBetaU=evpBetaU, evpBetaV=evpBetaV)
    return un, vn, s1, s2, s12


def evp_residuals(state, old, new, trace):
    """Residual norms of one substep, ``veris/evp_solver.py:192-230``: the
    relaxation-weighted changes of sig11 = (sigma1 + sigma2)/2, sig22 =
    (sigma1 - sigma2)/2, sigma12 and of u, v, squared and summed over the
    interior.  The reference reads the previous sig11/sig22 from names it
    never defines; they are taken here as the previous iterate's
    (sigma1 +- sigma2)/2, the quantities its stress divergence uses
    (``:93-94``).  Returns (resSig, resU)."""
    vs, s = _vs(state)
    uo, vo, s1o, s2o, s12o = old
    un, vn, s1n, s2n, s12n = new
    aC, aZ = trace["alphaC"], trace["alphaZ"]
    r11 = (0.5 * (s1n + s2n) - 0.5 * (s1o + s2o)) * aC * vs.iceMask
    r22 = (0.5 * (s1n - s2n) - 0.5 * (s1o - s2o)) * aC * vs.iceMask
    r12 = (s12n - s12o) * aZ
    ru = vs.iceMaskU * (un - uo) * trace["evpBetaU"]
    rv = vs.iceMaskV * (vn - vo) * trace["evpBetaV"]
    inner = (slice(s.olx, -s.olx), slice(s.oly, -s.oly))
    return float((r11**2 + r22**2 + r12**2)[inner].sum()), float((ru**2 + rv**2)[inner].sum())


def evp_solver_residuals(state, n_substeps=None):
    """``veris/evp_solver.py:257-319`` with computeEvpResidual: returns
    (uIce, vIce, sigma1, sigma2, sigma12, resU, resSig), resU and resSig one
    value per substep."""
    vs, s = _vs(state)
    n = s.nEVPsteps if n_substeps is None else n_substeps
    cur = (vs.uIce, vs.vIce, vs.sigma1, vs.sigma2, vs.sigma12)
    resU = np.zeros(n)
    resSig = np.zeros(n)
    for k in range(n):
        tr = {}
        new = evp_substep(state, *cur[:2], vs.uIce, vs.vIce, *cur[2:], trace=tr)
        resSig[k], resU[k] = evp_residuals(state, cur, new, tr)
        cur = new
    return (*cur, resU, resSig)


def evp_solver(state, n_substeps=None, P=None, drag=None):
    """``veris/evp_solver.py:257-319``: returns (uIce, vIce, sigma1, sigma2, sigma12)."""
    vs, s = _vs(state)
    n = s.nEVPsteps if n_substeps is None else n_substeps
    u, v = vs.uIce, vs.vIce
    uNm1, vNm1 = vs.uIce, vs.vIce
    s1, s2, s12 = vs.sigma1, vs.sigma2, vs.sigma12
    for _ in range(n):
        u, v, s1, s2, s12 = evp_substep(state, u, v, uNm1, vNm1, s1, s2, s12, P=P, drag=drag)
    return u, v, s1, s2, s12


def freedrift_solver(state):
    """``veris/freedrift_solver.py:5-62``: steady free-drift velocities (no
    internal stress) solved in closed form at c-points, averaged to u/v-points
    and masked.  Returns (uIce, vIce)."""
    vs, s = _vs(state)
    tx = 0.5 * (vs.WindForcingX + ip1(vs.WindForcingX))
    ty = 0.5 * (vs.WindForcingY + jp1(vs.WindForcingY))
    mf = s.rhoIce * vs.hIceMean * vs.fCori
    uoc = 0.5 * (vs.uOcean + ip1(vs.uOcean))
    voc = 0.5 * (vs.vOcean + jp1(vs.vOcean))
    rx = -tx - mf * voc
    ry = -ty + mf * uoc
    t1 = rx**2 + ry**2
    nz = t1 > 0
    rN = np.where(nz, np.sqrt(rx**2 + ry**2), 0)
    rA = np.where(nz, np.arctan2(ry, rx), 0)
    t1 = 1 / (np.where(vs.fCori < 0, s.waterIceDrag_south, s.waterIceDrag) * s.rhoSea)
    t2 = t1**2 * mf**2
    t3 = t1**2 * rN**2
    t4 = t2**2 + 4 * t3
    sN = np.where(t3 > 0, np.sqrt(0.5 * (np.sqrt(t4) - t2)), 0)
    t1 = 1 / t1
    t2 = t1 * sN**2
    t3 = mf * sN
    t4 = t2**2 + t3**2
    ang = np.where(t4 > 0, rA - np.arctan2(t3, t2), 0)
    uc = uoc - sN * np.cos(ang)
    vc = voc - sN * np.sin(ang)
    u = 0.5 * (im1(uc) + uc) * vs.iceMaskU
    v = 0.5 * (jm1(vc) + vc) * vs.iceMaskV
    return u, v


def IceVelocities(state, **kw):
    """``veris/dynsolver.py:87-102``.  The reference's free-drift branch reads
    the stresses from an undefined name; here it keeps the state's stresses,
    and EVP, when also on, overrides it as in the reference."""
    vs, s = _vs(state)
    if s.useEVP:
        return evp_solver(state, **kw)
    if s.useFreedrift:
        u, v = freedrift_solver(state)
        return u, v, vs.sigma1, vs.sigma2, vs.sigma12
    raise ValueError("IceVelocities: neither useEVP nor useFreedrift is set")


# --------------------------------------------------------------------------
# ocean stress
# --------------------------------------------------------------------------


def OceanStressUV(state):
    """``veris/ocean_stress.py:8-40``."""
    vs, s = _vs(state)
    cd = ocean_drag_coeffs(state, vs.uIce, vs.vIce)
    sinW = np.sin(np.deg2rad(s.waterTurnAngle))
    cosW = np.cos(np.deg2rad(s.waterTurnAngle))
    du = vs.uIce - vs.uOcean
    dv = vs.vIce - vs.vOcean
    duC = 0.5 * (du + ip1(du))
    dvC = 0.5 * (dv + jp1(dv))
    sg = np.sign(vs.fCori)
    tu = 0.5 * (cd + im1(cd)) * cosW * du - sg * sinW * 0.5 * (cd * dvC + jm1(cd * dvC))
    tv = 0.5 * (cd + jm1(cd)) * cosW * dv + sg * sinW * 0.5 * (cd * duC + im1(cd * duC))
    return fill_overlap_uv(state, tu, tv)


# --------------------------------------------------------------------------
# advection
# --------------------------------------------------------------------------


def limiter(cr):
    """``veris/advection.py:170-175``: superbee."""
    return np.maximum(0, np.maximum(np.minimum(1, 2 * cr), np.minimum(2, cr)))


def _slope_ratio(rjm, rj, rjp, trans, crmax):
    cr = np.where(trans > 0, rjm, rjp)
    with np.errstate(divide="ignore", invalid="ignore"):
        q = cr / rj
    cr = np.where(np.abs(rj) * crmax > np.abs(cr), q, np.sign(cr) * crmax * np.sign(rj))
    return limiter(cr)


def calc_ZonalFlux(state, f, uT):
    """``veris/advection.py:94-129``."""
    vs, s = _vs(state)
    mW = vs.iceMaskU * vs.maskInU
    cfl = np.abs(vs.uIce * s.deltatTherm * vs.recip_dxC)
    rjp = (f[3:, :] - f[2:-1, :]) * mW[3:, :]
    rj = (f[2:-1, :] - f[1:-2, :]) * mW[2:-1, :]
    rjm = (f[1:-2, :] - f[:-3, :]) * mW[1:-2, :]
    cr = _slope_ratio(rjm, rj, rjp, uT[2:-1, :], s.CrMax)
    F = np.zeros(vs.iceMask.shape)
    F[2:-1, :] = (uT[2:-1, :] * (f[2:-1, :] + f[1:-2, :]) * 0.5
                  - np.abs(uT[2:-1, :]) * ((1 - cr) + cfl[2:-1, :] * cr) * rj * 0.5)
    return fill_overlap(state, F)


def calc_MeridionalFlux(state, f, vT):
    """``veris/advection.py:132-167``."""
    vs, s = _vs(state)
    mS = vs.iceMaskV * vs.maskInV
    cfl = np.abs(vs.vIce * s.deltatTherm * vs.recip_dyC)
    rjp = (f[:, 3:] - f[:, 2:-1]) * mS[:, 3:]
    rj = (f[:, 2:-1] - f[:, 1:-2]) * mS[:, 2:-1]
    rjm = (f[:, 1:-2] - f[:, :-3]) * mS[:, 1:-2]
    cr = _slope_ratio(rjm, rj, rjp, vT[:, 2:-1], s.CrMax)
    F = np.zeros(vs.iceMask.shape)
    F[:, 2:-1] = (vT[:, 2:-1] * (f[:, 2:-1] + f[:, 1:-2]) * 0.5
                  - np.abs(vT[:, 2:-1]) * ((1 - cr) + cfl[:, 2:-1] * cr) * rj * 0.5)
    return fill_overlap(state, F)


def calc_Advection(state, field):
    """``veris/advection.py:28-91`` (extensive fields; the other branch is broken
    in the reference and unsupported here)."""
    vs, s = _vs(state)
    if not s.extensiveFld:
        raise NotImplementedError("extensiveFld=False")
    uT = vs.uIce * (vs.dyG * vs.iceMaskU)
    vT = vs.vIce * (vs.dxG * vs.iceMaskV)
    f = field
    F = calc_ZonalFlux(state, f, uT)
    f = f - s.deltatTherm * vs.maskInC * vs.recip_rA * (ip1(F) - F)
    G = calc_MeridionalFlux(state, f, vT)
    f = f - s.deltatTherm * vs.maskInC * vs.recip_rA * (jp1(G) - G)
    return f * vs.iceMask


def Advection(state):
    """``veris/advection.py:15-25``."""
    vs, _ = _vs(state)
    return (calc_Advection(state, vs.hIceMean), calc_Advection(state, vs.hSnowMean),
            calc_Advection(state, vs.Area))


# --------------------------------------------------------------------------
# clean-up and ridging
# --------------------------------------------------------------------------


def clean_up_advection(state):
    """``veris/clean_up.py:5-39``."""
    vs, s = _vs(state)
    os_h = np.maximum(-vs.hIceMean, 0)
    os_hs = np.maximum(-vs.hSnowMean, 0)
    h = np.maximum(vs.hIceMean, 0)
    hs = np.maximum(vs.hSnowMean, 0)
    A = np.maximum(vs.Area, 0)
    thin = h <= s.hIce_min
    h = h * ~thin
    hs = hs * ~thin
    T = np.where(thin, s.celsius2K, vs.TSurf)
    A = np.where((h == 0) & (hs == 0), 0, A)
    A = np.where((h > 0) | (hs > 0), np.maximum(A, s.Area_min), A)
    return h, hs, A, T, os_h, os_hs


def ridging(state):
    """``veris/clean_up.py:42-47``."""
    return np.minimum(state.variables.Area, 1)


# --------------------------------------------------------------------------
# thermodynamics
# --------------------------------------------------------------------------

AA1 = 2663.5
AA2 = 12.537
BB1 = 0.622
BB2 = 1 - BB1
PPASCALS = 100000
CC0 = 10**AA2
CC1 = CC0 * AA1 * BB1 * PPASCALS * np.log(10)
CC2 = CC0 * BB2


def solve4temp(state, hI, hS, Tin, Tfrz, n_newton=6):
    """``veris/solve4temp.py:5-174``."""
    vs, s = _vs(state)
    d1 = s.dalton * s.cpAir * s.rhoAir
    d1i = s.dalton * s.lhSublim * s.rhoAir
    Tmelt = s.celsius2K
    Tmeltw = Tmelt + s.wetAlbTemp
    T = Tin
    lwc = np.maximum(s.minLWdown, vs.LWdown)
    ta = np.maximum(s.celsius2K + s.minTAir, vs.ATemp)
    ug = np.maximum(s.wSpeedMin, vs.wSpeed)
    ice = hI > 0
    snow = hS > 0
    emis = np.where(snow, s.snowEmiss, s.iceEmiss)
    d3 = emis * s.stefBoltz
    lw = np.where(snow, s.snowEmiss, s.iceEmiss) * lwc
    aI = np.where(ice, s.dryIceAlb, 0)
    aS = np.where(ice, s.drySnowAlb, 0)
    wet = (hI > 0) & (T >= Tmeltw)
    aI = np.where(wet, s.wetIceAlb, aI)
    aS = np.where(wet, s.wetSnowAlb, aS)
    south = (hI > 0) & (vs.fCori < 0)
    aI = np.where(south, s.dryIceAlb_south, aI)
    aS = np.where(south, s.drySnowAlb_south, aS)
    wets = (hI > 0) & (vs.fCori < 0) & (T >= Tmeltw)
    aI = np.where(wets, s.wetIceAlb_south, aI)
    aS = np.where(wets, s.wetSnowAlb_south, aS)
    alb = np.where(ice, aI + hS / s.hCut * (aS - aI), 0)
    alb = np.where(hS > s.hCut, aS, alb)
    alb = np.where(hS == 0, aI, alb)
    pen = np.where(ice, s.shortwave * np.exp(-1.5 * hI), 0)
    pen = np.where(snow, 0, pen)
    penSW = np.where(ice, -(1 - alb) * pen * vs.SWdown, 0)
    absSW = np.where(ice, (1 - alb) * (1 - pen) * vs.SWdown, 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        kEff = np.where(ice, s.iceConduct * s.snowConduct
                        / (s.snowConduct * hI + s.iceConduct * hS), 0)

    def fluxes(t1):
        t2 = t1 * t1
        t3 = t2 * t1
        t4 = t2 * t2
        with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
            svp = 10 ** (-AA1 / t1 + AA2)
            qs = np.where(ice, BB1 * svp / (PPASCALS - (1 - BB1) * svp), 0)
            c3 = 10 ** (AA1 / t1)
            dq = np.where(ice, CC1 * c3 / ((CC2 - c3 * PPASCALS) ** 2 * t2), 0)
        Fc = np.where(ice, kEff * (Tfrz - t1), 0)
        Flh = np.where(ice, d1i * ug * (qs - vs.aqh), 0)
        Flw = np.where(ice, t4 * d3, 0)
        Fs = np.where(ice, d1 * ug * (t1 - ta), 0)
        Fia = np.where(ice, (-lw - absSW + Flw + Fs + Flh), 0)
        dF = np.where(ice, 4 * d3 * t3 + d1 * ug + d1i * ug * dq, 0)
        return Fc, Flh, Fia, dF

    for _ in range(n_newton):
        Fc, _Flh, Fia, dF = fluxes(T)
        with np.errstate(divide="ignore", invalid="ignore"):
            T = np.where(ice, T + (Fc - Fia) / (kEff + dF), 0)
        T = np.minimum(T, Tmelt)
        T = np.maximum(T, s.celsius2K + s.minTIce)
    Fc, Flh, Fia, _dF = fluxes(T)
    up = Fc > 0
    Fio = np.where(up, Fc, 0)
    Fia_net = np.where(up, 0, Fia)
    Tout = np.where(ice, T, Tin)
    fws = np.where(ice, Flh / s.lhSublim, 0)
    return Tout, Fio, Fia_net, penSW, fws


def Growth(state):
    """``veris/growth.py:8-397``: returns the reference's 11-tuple."""
    vs, s = _vs(state)
    qi = 1 / (s.rhoIce * s.lhFusion)
    qs = 1 / (s.rhoSnow * s.lhFusion)
    noIce = (vs.hIceMean == 0) | (vs.Area == 0)
    hP = vs.hIceMean * ~noIce
    hsP = vs.hSnowMean * ~noIce
    AP = vs.Area * ~noIce
    isIce = hP > 0
    regA = np.sqrt(AP**2 + s.Area_reg)
    rregA = 1 / regA
    hIA = np.where(isIce, hP * rregA, 0)
    rhIA = AP / np.sqrt(hP**2 + s.hIce_reg)
    hSA = np.where(isIce, hsP * rregA, 0)
    hIA = np.maximum(hIA, 0.05)
    Tfrz = s.tempFrz + s.dtempFrz_dS * vs.ocSalt + s.celsius2K
    Tc, Fio_c, Fia_c, pen_c = [], [], [], []
    for l in range(s.nITC):
        p = (2 * (l + 1) - 1) * s.recip_nITC
        out = solve4temp(state, hIA * p, hSA * p, vs.TSurf, Tfrz)
        Tc.append(out[0])
        Fio_c.append(out[1])
        Fia_c.append(out[2])
        pen_c.append(out[3])

    def csum(lst):
        acc = lst[0]
        for x in lst[1:]:
            acc = acc + x
        return acc

    Tsum = csum(Tc)
    TSurf = Tsum * s.recip_nITC
    Fio = csum(Fio_c) * s.recip_nITC * AP
    Fia = csum(Fia_c) * s.recip_nITC * AP
    pen = csum(pen_c) * s.recip_nITC * AP
    cold = (AP > 0) & (Tsum / s.nITC < s.celsius2K)
    snowRate = np.where(cold, vs.snowfall + vs.precip, vs.snowfall) * s.rhoFresh2rhoSnow
    rainToSea = np.where(cold, 0, vs.precip)
    snowAcc = snowRate * AP * s.deltatTherm
    gUnder = Fio * qi
    gUnder = np.where(AP == 0, 0, gUnder)
    potMeltRate = -Fia * qs
    potMelt = potMeltRate * s.deltatTherm
    allMelt = potMelt >= hsP
    melt = np.where(allMelt, hsP, potMelt)
    meltRate = np.where(allMelt, melt * s.recip_deltatTherm, potMeltRate)
    toMelt = np.where(allMelt, -hsP * s.recip_deltatTherm / qs, Fia)
    Fia = Fia - toMelt
    gSurf = Fia * qi
    gNet = (gUnder + gSurf) * rregA
    t0 = 0.4
    t1 = 7 / t0
    t2 = s.stantonNr * s.uStarBase * s.rhoSea * s.cpWater
    sth = np.maximum(vs.theta, Tfrz)
    mltf = 1 + (s.McPheeTaperFac - 1) / (1 + np.exp((AP - t0) * t1))
    Foi = -t2 * (sth - Tfrz) * mltf
    gML = Foi * qi
    gOW = qi * vs.Qnet
    dh = gNet * AP + gOW * (1 - AP) + gML
    dhs = (snowRate - meltRate) * AP
    half = 0.5 * rhIA
    dAoa = half * gOW * (1 - AP)
    grow = (gOW > 0) & ((AP > 0) | (dh > 0))
    dAoa = np.where(grow, gOW * (1 - AP), dAoa)
    dAoa = np.where(grow & (vs.fCori < 0), dAoa * s.recip_h0_south, dAoa * s.recip_h0)
    dAoi = np.where(gML <= 0, half * gML, 0)
    dAia = np.where((gNet <= 0) & (hP > 0), half * gNet * AP, 0)
    dA = dAoa + dAoi + dAia
    A = AP + dA * vs.iceMask * s.deltatTherm
    h = hP + dh * vs.iceMask * s.deltatTherm
    hs = hsP + dhs * vs.iceMask * s.deltatTherm
    A = np.clip(A, 0, 1)
    h = np.maximum(h, 0)
    hs = np.maximum(hs, 0)
    noIce = (h == 0) | (A == 0)
    A = A * ~noIce
    hs = hs * ~noIce
    h = h * ~noIce
    hsub = (hs * s.rhoSnow + h * s.rhoIce) * s.recip_rhoSea
    flood = np.maximum(0, hsub - h)
    h = h + flood
    hs = hs - flood * s.rhoIce2rhoSnow
    rh = 1 / np.sqrt(h**2 + s.hIce_reg)
    Qsw = pen * AP + vs.Qsw * (1 - AP)
    dH = h - hP
    meltedSnow = hsP + snowAcc - hs
    eIce = dH / qi
    eOut = (AP * (Fia + Fio + pen) + (1 - AP) * vs.Qnet) * s.deltatTherm
    Qnet = (eOut - eIce) * s.recip_deltatTherm
    with np.errstate(divide="ignore", invalid="ignore"):
        sal = np.where((vs.ocSalt > 0) & (vs.ocSalt > s.saltIce_ref),
                       (1 - s.saltIce_ref / vs.ocSalt), 1)
    fwIce = -dH * s.rhoIce2rhoFresh * sal
    t0s = np.minimum(s.saltIce_ref, vs.ocSalt)
    saltflux = (dH + vs.os_hIceMean) * t0s * vs.iceMask * s.rhoIce * s.recip_deltatTherm
    fwSnow = meltedSnow / s.rhoFresh2rhoSnow
    EmPmR = (vs.iceMask * ((-vs.evap - vs.precip) * (1 - AP) - rainToSea * AP - vs.runoff
                           - (fwIce + fwSnow) / s.deltatTherm) * s.rhoFresh
             + vs.iceMask * (vs.os_hIceMean * s.rhoIce + vs.os_hSnowMean * s.rhoSnow)
             * s.recip_deltatTherm)
    fss = EmPmR * vs.ocSalt / s.rhoFresh + saltflux
    load = h * s.rhoIce + hs * s.rhoSnow
    return h, hs, A, TSurf, EmPmR, fss, Qsw, Qnet, load, pen, rh


# --------------------------------------------------------------------------
# whole step
# --------------------------------------------------------------------------

STEP_FILLS = ("hIceMean", "hSnowMean", "Area", "forc_salt_surface_ice", "Qnet")


def seaice_step(state, qnet_forcing=None, swdown=None, n_substeps=None):
    """The sea-ice part of ``seaice_global_4deg.py:set_forcing_kernel``.

    Order of ``:874-932``: ocean heat flux and short-wave copies, SeaIceMass,
    AreaWS, WindForcingXY, SeaIceStrength, IceVelocities, OceanStressUV,
    Advection, clean_up_advection, ridging, Growth, five overlap fills.
    The stresses returned by IceVelocities are kept in the state (the
    reference's setup discards them).  Updates ``state.variables`` in place.
    """
    vs, s = _vs(state)
    if qnet_forcing is not None:
        vs.Qnet = qnet_forcing.copy()
    if swdown is not None:
        vs.Qsw = -swdown
    vs.SeaIceMassC, vs.SeaIceMassU, vs.SeaIceMassV = SeaIceMass(state)
    vs.AreaW, vs.AreaS = AreaWS(state)
    vs.WindForcingX, vs.WindForcingY = WindForcingXY(state)
    vs.SeaIceStrength = SeaIceStrength(state)
    vs.uIce, vs.vIce, vs.sigma1, vs.sigma2, vs.sigma12 = IceVelocities(state, n_substeps=n_substeps)
    vs.OceanStressU, vs.OceanStressV = OceanStressUV(state)
    vs.hIceMean, vs.hSnowMean, vs.Area = Advection(state)
    (vs.hIceMean, vs.hSnowMean, vs.Area, vs.TSurf,
     vs.os_hIceMean, vs.os_hSnowMean) = clean_up_advection(state)
    vs.Area = ridging(state)
    (vs.hIceMean, vs.hSnowMean, vs.Area, vs.TSurf, vs.EmPmR, vs.forc_salt_surface_ice,
     vs.Qsw, vs.Qnet, vs.SeaIceLoad, vs.IcePenetSW, vs.recip_hIceMean) = Growth(state)
    for name in STEP_FILLS:
        vs[name] = fill_overlap(state, vs[name])
    return state


# --------------------------------------------------------------------------
# exchange with the host ocean (the setup's inline updates)
# --------------------------------------------------------------------------


def ocean_to_ice(state, u_surface, v_surface, temp_surface, salt_surface, forc_temp_surface):
    """``seaice_global_4deg.py:847, 868-876`` -> (uOcean, vOcean, theta,
    ocSalt, Qnet, Qsw, wSpeed); wSpeed from the wind before its overlap fill."""
    vs, s = _vs(state)
    theta = temp_surface + s.celsius2K
    Qnet = -forc_temp_surface * s.cpWater * s.rhoSea
    Qsw = -vs.SWdown
    wSpeed = np.sqrt(vs.uWind**2 + vs.vWind**2)
    return (np.array(u_surface, dtype=np.float64), np.array(v_surface, dtype=np.float64), theta,
            np.array(salt_surface, dtype=np.float64), Qnet, Qsw, wSpeed)


def ice_to_ocean(state, surface_taux, surface_tauy, forc_salt_surface_ice, forc_salt_surface_res):
    """``seaice_global_4deg.py:935-940`` -> (surface_taux, surface_tauy,
    forc_temp_surface, forc_salt_surface)."""
    vs, s = _vs(state)
    taux = surface_taux * (1 - vs.AreaW) + vs.OceanStressU * vs.AreaW
    tauy = surface_tauy * (1 - vs.AreaS) + vs.OceanStressV * vs.AreaS
    forc_temp = -vs.Qnet / (s.cpWater * s.rhoSea)
    forc_salt = forc_salt_surface_ice + forc_salt_surface_res
    return taux, tauy, forc_temp, forc_salt


# --------------------------------------------------------------------------
# forcing producer: open-ocean bulk fluxes
# --------------------------------------------------------------------------


def bulkf_formula_lanl(state, uw, vw, ta, qa, tsf, ocn_mask, grav=9.81):
    """``veris/heat_flux_MITgcm.py:5-123``: open-ocean bulk fluxes with
    stability-dependent transfer coefficients from five Monin-Obukhov
    iterations.  ``grav`` is the host model's gravity (Veros' ``grav``).
    Returns (flwupa, flha, fsha, df0dT, ust, vst, evp, ssq, dEvdT)."""
    s = state.settings
    ht, zref, zice = 2.0, 10.0, 0.0005
    aln = np.log(ht / zref)
    czol = zref * s.karman * grav
    lath = np.ones_like(ocn_mask) * s.latvap
    us = np.sqrt(uw * uw + vw * vw)
    usm = np.maximum(us, 1.0)
    t0 = ta * (1.0 + s.zvir * qa)
    ssq = 3.797915 * np.exp(lath * (7.93252e-6 - 2.166847e-3 / tsf)) / 1013.0
    deltap = ta - tsf + s.gamma_blk * ht
    delq = qa - ssq
    rdn = s.karman / np.log(zref / zice)
    rhn = rdn
    ren = rdn
    ustar = rdn * usm
    tstar = rhn * deltap
    qstar = ren * delq
    for _ in range(5):
        huol = czol / ustar**2 * (tstar / t0 + qstar / (1.0 / s.zvir + qa))
        huol = np.minimum(np.abs(huol), 10.0) * np.sign(huol)
        stable = 0.5 + 0.5 * np.sign(huol)
        xsq = np.maximum(np.sqrt(np.abs(1.0 - 16.0 * huol)), 1.0)
        x = np.sqrt(xsq)
        psimh = -5.0 * huol * stable + (1.0 - stable) * (
            2.0 * np.log(0.5 * (1.0 + x)) + 2.0 * np.log(0.5 * (1.0 + xsq))
            - 2.0 * np.arctan(x) + np.pi * 0.5)
        psixh = -5.0 * huol * stable + (1.0 - stable) * (2.0 * np.log(0.5 * (1.0 + xsq)))
        rd = rdn / (1.0 + rdn * (aln - psimh) / s.karman)
        rh = rhn / (1.0 + rhn * (aln - psixh) / s.karman)
        re = rh
        ustar = rd * usm
        qstar = re * delq
        tstar = rh * deltap
    csha = s.rhoAir * s.cpdair * us * rh * rd
    clha = s.rhoAir * lath * us * re * rd
    fsha = csha * deltap
    flha = clha * delq
    evp = -flha / lath
    flwupa = s.ocean_emissivity * s.stefBoltz * tsf**4
    dflwupdt = 4.0 * s.ocean_emissivity * s.stefBoltz * tsf**3
    devdt = clha * ssq * 2.166847e-3 / (tsf * tsf)
    dflhdt = -lath * devdt
    dfshdt = -csha
    df0dt = -dflwupdt + dfshdt + dflhdt
    cdn = 2.7e-3 / usm + 0.142e-3 + 0.0764e-3 * usm
    ust = s.rhoAir * cdn * us * uw
    vst = s.rhoAir * cdn * us * vw
    return flwupa, flha, fsha, df0dt, ust, vst, evp, ssq, devdt
