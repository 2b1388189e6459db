"""Torch (CPU, FP64) restatement of the pre-EVP forcing and the EVP subcycle,
differentiated by torch.autograd -- TEST INFRASTRUCTURE ONLY.

It stands in for ``jax.grad`` of the reference (JAX is absent here) when the
hand-written HIP adjoint is checked.  It follows the reference's expressions
(veris/evp_solver.py:22-319, veris/dynamics_routines.py:5-302,
veris/averaging.py:4-19, veris/area_mass.py:5-28, veris/dynsolver.py:7-84)
optionally with adaptive EVP, the second-order boundary term, the
water-stress turning term, the metric terms, free slip with lateral drag
and basal drag, written with torch.roll, and defines the derivative of sqrt at 0 as 0 (the reference's
Delta = sqrt(Delta^2) and drag sqrt are not differentiable there; the HIP
adjoint uses the same convention).
"""

from __future__ import annotations

import math

import torch


class _SafeSqrt(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        y = torch.sqrt(x)
        ctx.save_for_backward(y)
        return y

    @staticmethod
    def backward(ctx, g):
        (y,) = ctx.saved_tensors
        return torch.where(y > 0, g * 0.5 / torch.where(y > 0, y, torch.ones_like(y)), torch.zeros_like(y))


ssqrt = _SafeSqrt.apply


def im1(a):
    return torch.roll(a, 1, 0)


def ip1(a):
    return torch.roll(a, -1, 0)


def jm1(a):
    return torch.roll(a, 1, 1)


def jp1(a):
    return torch.roll(a, -1, 1)


def im2(a):
    return torch.roll(a, 2, 0)


def jm2(a):
    return torch.roll(a, 2, 1)


def fill(sett, a):
    if sett.veros_fill:
        if not sett.enable_cyclic_x:
            return a
        a = a.clone()
        a[-2:] = a[2:4]
        a[:2] = a[-4:-2]
        return a
    a = a.clone()
    a[:2, :] = a[-4:-2, :]
    a[-2:, :] = a[2:4, :]
    a[:, :2] = a[:, -4:-2]
    a[:, -2:] = a[:, 2:4]
    return a


def pre_evp(vs, sett, pstar, h, ca):
    """SeaIceMass, AreaWS, WindForcingXY, SeaIceStrength with a per-cell P*,
    thickness h and air drag ca (both hemispheres)."""
    mC = sett.rhoIce * h + sett.rhoSnow * vs["hSnowMean"]
    mU = 0.5 * (mC + im1(mC))
    mV = 0.5 * (mC + jm1(mC))
    A = vs["Area"]
    AW = 0.5 * (A + im1(A))
    AS = 0.5 * (A + jm1(A))
    u, v = vs["uIce"], vs["vIce"]
    urel = vs["uWind"] - 0.5 * (u + ip1(u))
    vrel = vs["vWind"] - 0.5 * (v + jp1(v))
    wsq = urel**2 + vrel**2
    ws = torch.where(wsq < sett.wSpeedMin**2, torch.full_like(wsq, sett.wSpeedMin), ssqrt(wsq))
    cda = ca * sett.rhoAir * ws
    tx = cda * urel
    ty = cda * vrel
    tx = 0.5 * (tx + im1(tx)) * vs["iceMaskU"]
    ty = 0.5 * (ty + jm1(ty)) * vs["iceMaskV"]
    if sett.useRealFreshWaterFlux:  # sea-ice load in the tilt (veris/dynsolver.py:65-72)
        phi = sett.gravity * vs["ssh_an"] + (vs["surfPress"] + vs["SeaIceLoad"] * sett.gravity
                                             * sett.seaIceLoadFac) * sett.recip_rhoSea
    else:
        phi = sett.gravity * vs["ssh_an"] + vs["surfPress"] * sett.recip_rhoSea
    fx = tx * AW - mU * vs["recip_dxC"] * (phi - im1(phi))
    fy = ty * AS - mV * vs["recip_dyC"] * (phi - jm1(phi))
    P = pstar * h * torch.exp(-sett.cStar * (1 - A)) * vs["iceMask"]
    return dict(SeaIceMassC=mC, SeaIceMassU=mU, SeaIceMassV=mV, AreaW=AW, AreaS=AS,
                WindForcingX=fx, WindForcingY=fy, SeaIceStrength=P, hIceMean=h)


def evp_substep(vs, sett, p, u, v, u0, v0, s1, s2, s12, cw):
    """One EVP substep; ``p`` holds the pre-EVP fields, ``cw`` the water drag."""
    M, MU, MV = vs["iceMask"], vs["iceMaskU"], vs["iceMaskV"]
    dudx = (ip1(u) - u) * vs["recip_dxU"]
    dvdy = (jp1(v) - v) * vs["recip_dyV"]
    kt = {k: vs.get(k) for k in ("k1AtC", "k2AtC", "k1AtZ", "k2AtZ")}
    if kt["k1AtC"] is not None:  # metric terms (veris/dynamics_routines.py:151-169)
        dudx = dudx + 0.5 * (jp1(v) + v) * kt["k2AtC"]
        dvdy = dvdy + 0.5 * (ip1(u) + u) * kt["k1AtC"]
    e11 = dudx * vs["maskInC"]
    e22 = dvdy * vs["maskInC"]
    dudy = (u - jm1(u)) * vs["recip_dyU"]
    uavz = (u + jm1(u)) * 0.5
    dvdx = (v - im1(v)) * vs["recip_dxV"]
    vavz = (v + im1(v)) * 0.5
    mz = M * im1(M)
    mz = mz * jm1(mz)
    inner = dudy + dvdx
    if kt["k1AtZ"] is not None:
        inner = inner - kt["k1AtZ"] * vavz - kt["k2AtZ"] * uavz
    e12 = 0.5 * inner * mz
    if sett.noSlip:
        e12 = e12 + (2.0 * uavz * vs["recip_dyU"] * (MU - jm1(MU))
                     + 2.0 * vavz * vs["recip_dxV"] * (MV - im1(MV)))
    if sett.noSlip and sett.secondOrderBC:  # veris/dynamics_routines.py:177-215
        hu2 = (MU - im1(MU)) / 3.0 * (im2(MU) * im1(MU) + ip1(MU) * MU)
        hv2 = (MV - jm1(MV)) / 3.0 * (jm2(MV) * jm1(MV) + jp1(MV) * MV)
        e12 = e12 + 0.5 * (vs["recip_dyU"] * (6.0 * uavz - im2(u) * im1(MU) - ip1(u) * MU) * hu2
                           + vs["recip_dxV"] * (6.0 * vavz - jm2(v) * jm1(MV) - jp1(v) * MV) * hv2)
    rpd = 1.0 / sett.PlasDefCoeff**2
    q = vs["rAz"] * e12**2
    q = q + ip1(q)
    e12c = 0.25 * vs["recip_rA"] * (q + jp1(q))
    dsq = (e11 + e22) ** 2 + rpd * ((e11 - e22) ** 2 + 4.0 * e12c)
    dC = ssqrt(dsq)
    P = p["SeaIceStrength"]
    zeta = 0.5 * (P * (1 + sett.tensileStrFac)) / (dC + sett.deltaMin)
    press = (P * (1 - sett.pressReplFac) + 2.0 * zeta * dC * sett.pressReplFac
             / (1 + sett.tensileStrFac)) * (1 - sett.tensileStrFac)
    alpha = sett.evpAlpha
    ep, em = e11 + e22, e11 - e22
    n = M + im1(M)
    n = n + jm1(n)
    nrm = torch.where(n > 0, 1.0 / torch.where(n > 0, n, torch.ones_like(n)), torch.zeros_like(n))
    zz = zeta + im1(zeta)
    zz = nrm * (zz + jm1(zz))
    aC = aZ = alpha
    if sett.useAdaptiveEVP:  # veris/evp_solver.py:65-73, 97-100, 266-271
        cfac = sett.deltatDyn * 4 * (math.pi * sett.aEvpCoeff) ** 2
        mC0 = torch.clamp(p["SeaIceMassC"], min=1e-4)
        aC = ssqrt(zeta * cfac / mC0 * vs["recip_rA"]) * M
        aC = torch.maximum(aC, torch.full_like(aC, sett.aEVPalphaMin))
        aZ = 0.5 * (aC + jm1(aC))
        aZ = 0.5 * (aZ + im1(aZ))
    s1 = (s1 * (aC - 1) + (2 * zeta * ep - press)) / aC * M
    s2 = (s2 * (aC - 1) + 2 * zeta * em * 0.25) / aC * M
    s12 = (s12 * (aZ - 1) + 2 * zz * e12 * 0.25) / aZ
    s11 = 0.5 * (s1 + s2)
    s22 = 0.5 * (s1 - s2)
    a = s11 * vs["dyV"]
    b = s12 * vs["dxV"]
    divx = (a - im1(a) - b + jp1(b)) * vs["recip_rAu"]
    c = s22 * vs["dxU"]
    d = s12 * vs["dyU"]
    divy = (c - jm1(c) - d + ip1(d)) * vs["recip_rAv"]
    D = cw * sett.rhoSea
    du = (u - vs["uOcean"]) * vs["maskInU"]
    dv = (v - vs["vOcean"]) * vs["maskInV"]
    w = 0.25 * ((du + ip1(du)) ** 2 + (dv + jp1(dv)) ** 2)
    cd = torch.where(D**2 * w > sett.cDragMin**2, D * ssqrt(w), torch.full_like(w, sett.cDragMin)) * M
    f = vs["fCori"]
    mC, mU, mV = p["SeaIceMassC"], p["SeaIceMassU"], p["SeaIceMassV"]
    fv = mC * f * 0.5 * (v + jp1(v))
    fu = mC * f * 0.5 * (u + ip1(u))
    cdU = 0.5 * (cd + im1(cd))
    cdV = 0.5 * (cd + jm1(cd))
    inx = cdU * sett.cosWat * vs["uOcean"]
    iny = cdV * sett.cosWat * vs["vOcean"]
    if sett.sinWat != 0:  # turning term (veris/evp_solver.py:119-139)
        sg = torch.sign(f)
        lmU = (mU != 0).to(mU.dtype)
        lmV = (mV != 0).to(mV.dtype)
        du0 = vs["uOcean"] - u
        dv0 = vs["vOcean"] - v
        X = cd * 0.5 * (du0 + jp1(du0)) + im1(cd) * 0.5 * (im1(du0) + jp1(im1(du0)))
        Y = cd * 0.5 * (dv0 + ip1(dv0)) + jm1(cd) * 0.5 * (jm1(dv0) + ip1(jm1(dv0)))
        inx = inx - sg * sett.sinWat * 0.5 * X * lmU
        iny = iny + sg * sett.sinWat * 0.5 * Y * lmV
    fx = p["WindForcingX"] + inx * p["AreaW"] + 0.5 * (fv + im1(fv))
    fy = p["WindForcingY"] + iny * p["AreaS"] - 0.5 * (fu + jm1(fu))
    bU = bV = sett.evpBeta * sett.recip_deltatDyn
    if sett.useAdaptiveEVP:  # veris/evp_solver.py:148-155
        bU = 0.5 * (aC + im1(aC)) * sett.recip_deltatDyn
        bV = 0.5 * (aC + jm1(aC)) * sett.recip_deltatDyn
    bP1 = bU + sett.recip_deltatDyn
    bP1V = bV + sett.recip_deltatDyn
    cbU = cbV = 0.0
    if sett.basalDragK2 != 0:  # basal drag (veris/dynamics_routines.py:55-96)
        A, h = vs["Area"], p.get("hIceMean", vs["hIceMean"])
        um, vm = u * vs["maskInU"], v * vs["maskInV"]
        tb = 0.25 * (um**2 + ip1(um) ** 2 + vm**2 + jp1(vm) ** 2)
        tb = sett.basalDragK2 / torch.sqrt(tb + sett.basalDragU0**2)
        hc = torch.abs(vs["R_low"]) * A / sett.basalDragK1
        val = (tb * torch.log(torch.exp(10.0 * (h - hc)) + 1.0) * 0.1
               * torch.exp(-sett.cBasalStar * (1.0 - A)))
        cb = torch.where(A > 0.01, val, torch.zeros_like(val))
        cbU, cbV = 0.5 * (cb + im1(cb)), 0.5 * (cb + jm1(cb))
    denU = mU * bP1 + p["AreaW"] * (cdU * sett.cosWat + cbU)
    denV = mV * bP1V + p["AreaS"] * (cdV * sett.cosWat + cbV)
    denU = torch.where(denU == 0, torch.ones_like(denU), denU)
    denV = torch.where(denV == 0, torch.ones_like(denV), denV)
    if not sett.noSlip:  # lateral drag (veris/dynamics_routines.py:98-140, evp_solver.py:169-176)
        speed = 0.5 * ssqrt((u + ip1(u)) ** 2 + (v + jp1(v)) ** 2)
        z = torch.zeros_like(speed)
        spU = torch.where(p["AreaW"] > 0, 0.5 * (speed + im1(speed)), z)
        spV = torch.where(p["AreaS"] > 0, 0.5 * (speed + jm1(speed)), z)
        if sett.use_coastline:
            mskU, mskV = vs["Fu"], vs["Fv"]
        else:
            mskU = (2 - im1(MU) - ip1(MU)) * MU
            mskV = (2 - jm1(MV) - jp1(MV)) * MV
        denU = denU + mU * sett.sideDragCoeff * mskU / (spU + sett.sideDragU0)
        denV = denV + mV * sett.sideDragCoeff * mskV / (spV + sett.sideDragU0)
    un = MU * (bU * mU * u + mU * sett.recip_deltatDyn * u0 + fx + divx) / denU
    vn = MV * (bV * mV * v + mV * sett.recip_deltatDyn * v0 + fy + divy) / denV
    return fill(sett, un), fill(sett, vn), s1, s2, s12


def evp(vs, sett, p, n, cw):
    u, v = vs["uIce"], vs["vIce"]
    s1, s2, s12 = vs["sigma1"], vs["sigma2"], vs["sigma12"]
    for _ in range(n):
        u, v, s1, s2, s12 = evp_substep(vs, sett, p, u, v, vs["uIce"], vs["vIce"], s1, s2, s12, cw)
    return u, v, s1, s2, s12


def misfit(vs, sett, pstar, h, cw, ca, n, ut, vt):
    """L2 misfit of the final velocities to (ut, vt) over the interior."""
    p = pre_evp(vs, sett, pstar, h, ca)
    u, v, *_ = evp(vs, sett, p, n, cw)
    return ((u - ut)[2:-2, 2:-2] ** 2).sum() + ((v - vt)[2:-2, 2:-2] ** 2).sum()


def _superbee(cr):
    """Superbee as a cascade of selections (same values as
    max(0, max(min(1, 2 cr), min(2, cr)))), so autograd takes the slope of
    the segment the value comes from: 2 on (0, 1/2], 1 on (1, 2], else 0."""
    z = torch.zeros_like(cr)
    return torch.where(cr <= 0, z, torch.where(cr <= 0.5, 2 * cr, torch.where(
        cr <= 1, z + 1, torch.where(cr <= 2, cr, z + 2))))


def _face_flux(trans, cfl, fp, f0, fm, fmm, mWp, mW0, mWm, crmax):
    rjp = (fp - f0) * mWp
    rj = (f0 - fm) * mW0
    rjm = (fm - fmm) * mWm
    c0 = torch.where(trans > 0, rjm, rjp)
    ratio = torch.abs(rj) * crmax > torch.abs(c0)
    safe = torch.where(ratio, rj, torch.ones_like(rj))
    cr = torch.where(ratio, c0 / safe, torch.sign(c0) * crmax * torch.sign(rj))
    psi = _superbee(cr)
    return trans * (f0 + fm) * 0.5 - torch.abs(trans) * ((1 - psi) + cfl * psi) * rj * 0.5


def advection(vs, sett, u, v, f):
    """One field through the flux-limited, dimensionally split advection
    (veris/advection.py:28-175): zonal fluxes on faces i in [2, NX-1), the x
    update, meridional fluxes of the updated field on faces j in [2, NY-1),
    the y update, the land mask."""
    uT = u * (vs["dyG"] * vs["iceMaskU"])
    vT = v * (vs["dxG"] * vs["iceMaskV"])
    coef = sett.deltatTherm * vs["maskInC"] * vs["recip_rA"]
    mW = vs["iceMaskU"] * vs["maskInU"]
    cfl = torch.abs(u * sett.deltatTherm * vs["recip_dxC"])
    F = torch.zeros_like(f)
    F = torch.cat([F[:2], _face_flux(uT[2:-1], cfl[2:-1], f[3:], f[2:-1], f[1:-2], f[:-3],
                                     mW[3:], mW[2:-1], mW[1:-2], sett.CrMax), F[-1:]], 0)
    F = fill(sett, F)
    f1 = f - coef * (ip1(F) - F)
    mS = vs["iceMaskV"] * vs["maskInV"]
    cfl = torch.abs(v * sett.deltatTherm * vs["recip_dyC"])
    G = torch.zeros_like(f)
    G = torch.cat([G[:, :2], _face_flux(vT[:, 2:-1], cfl[:, 2:-1], f1[:, 3:], f1[:, 2:-1],
                                        f1[:, 1:-2], f1[:, :-3], mS[:, 3:], mS[:, 2:-1],
                                        mS[:, 1:-2], sett.CrMax), G[:, -1:]], 1)
    G = fill(sett, G)
    f2 = f1 - coef * (jp1(G) - G)
    return f2 * vs["iceMask"]
