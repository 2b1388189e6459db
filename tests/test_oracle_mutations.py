"""The oracle pins must fail on plausible restatement mistakes.

Each case plants one mistake (a flipped roll direction, a swapped metric, a
dropped factor, a wrong sign) into a temporary copy of the NumPy oracle, on the
CPU, and runs the pin tests against that copy; the pins must report a failure
for every planted mistake.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
ORACLE = ROOT / "oracle" / "veris_np.py"

MUTATIONS = {
    "slope ratio upwind side": ("cr = np.where(trans > 0, rjm, rjp)", "cr = np.where(trans > 0, rjp, rjm)"),
    "shear corner direction": ("e12c = 0.25 * vs.recip_rA * (q + jp1(q))", "e12c = 0.25 * vs.recip_rA * (q + jm1(q))"),
    "stressdiv x shear term": ("dx = (a - im1(a) - b + jp1(b)) * vs.recip_rAu", "dx = (a - im1(a) - b + jm1(b)) * vs.recip_rAu"),
    "stressdiv y shear term": ("dy = (c - jm1(c) - d + ip1(d)) * vs.recip_rAv", "dy = (c - jm1(c) - d + im1(d)) * vs.recip_rAv"),
    "u denominator drag": ("0.5 * (cd + im1(cd)) * s.cosWat + 0.5 * (cb + im1(cb)))", "0.5 * (cd + ip1(cd)) * s.cosWat + 0.5 * (cb + im1(cb)))"),
    "v denominator drag": ("0.5 * (cd + jm1(cd)) * s.cosWat + 0.5 * (cb + jm1(cb)))", "0.5 * (cd + jp1(cd)) * s.cosWat + 0.5 * (cb + jm1(cb)))"),
    "no-slip v term": ("hv = vs.iceMaskV - im1(vs.iceMaskV)", "hv = vs.iceMaskV - ip1(vs.iceMaskV)"),
    "z-mask corners": ("mz = mz * jm1(mz)", "mz = mz * jp1(mz)"),
    "zeta factor": ("zeta = 0.5 * (P * (1 + s.tensileStrFac)) / dreg", "zeta = (P * (1 + s.tensileStrFac)) / dreg"),
    "replacement pressure": ("+ 2.0 * zeta * dC * s.pressReplFac / (1 + s.tensileStrFac))", "+ zeta * dC * s.pressReplFac / (1 + s.tensileStrFac))"),
    "alpha relaxation": ("s1 = (s1 * (alphaC - rev) + divergence) * den1 * vs.iceMask", "s1 = (s1 * alphaC + divergence) * den1 * vs.iceMask"),
    "tension 1/e^2": ("s2 = (s2 * (alphaC - rev) + tension * rrev) * den2 * vs.iceMask", "s2 = (s2 * (alphaC - rev) + tension) * den2 * vs.iceMask"),
    "shear 1/e^2": ("s12 = (s12 * (alphaZ - rev) + shear * rrev) * den2", "s12 = (s12 * (alphaZ - rev) + shear) * den2"),
    "adaptive alpha z average": ("alphaZ = 0.5 * (alphaZ + im1(alphaZ))", "alphaZ = 0.5 * (alphaZ + ip1(alphaZ))"),
    "adaptive z denominator": ("        den2 = 1.0 / alphaZ", "        den2 = 1.0 / alphaC"),
    "adaptive beta u": ("evpBetaU = 0.5 * (alphaC + im1(alphaC))", "evpBetaU = 0.5 * (alphaC + jm1(alphaC))"),
    "adaptive alpha mass": ("/ np.maximum(vs.SeaIceMassC, 1e-4)", "/ np.maximum(vs.SeaIceMassU, 1e-4)"),
    "adaptive alpha floor": ("alphaC = np.maximum(alphaC, s.aEVPalphaMin)", "alphaC = np.maximum(alphaC, 1.0)"),
    "adaptive c-factor": ("return s.deltatDyn * 4 * (np.pi * s.aEvpCoeff) ** 2", "return s.deltatDyn * 4 * (np.pi * s.aEvpCoeff)"),
    "side drag speed v faces": ("(u + ip1(u)) ** 2 + (v + jp1(v)) ** 2)", "(u + ip1(u)) ** 2 + (v + jm1(v)) ** 2)"),
    "side drag u-point speed": ("0.5 * (speed + im1(speed)), 0)", "0.5 * (speed + ip1(speed)), 0)"),
    "side drag v area gate": ("speedV = np.where(vs.AreaS > 0", "speedV = np.where(vs.AreaW > 0"),
    "side drag land count": ("maskU = (2 - im1(vs.iceMaskU) - ip1(vs.iceMaskU))", "maskU = (2 - im1(vs.iceMaskU) - jp1(vs.iceMaskU))"),
    "side drag mass": ("sdU = vs.SeaIceMassU * s.sideDragCoeff", "sdU = vs.SeaIceMassC * s.sideDragCoeff"),
    "coastline form factor": ("        maskU = vs.Fu", "        maskU = vs.Fv"),
    "side drag u denominator": ("denU = denU + sdU", "denU = denU + sdV"),
    "free-slip boundary term": ("    if s.noSlip:\n", "    if True:\n"),
    "second-order u two back": ("- im2(u) * im1(MU)", "- im1(u) * im1(MU)"),
    "second-order v scale": ("hv2 = (MV - jm1(MV)) / 3.0", "hv2 = (MV - jm1(MV)) / 2.0"),
    "second-order v axis": ("hv2 = (MV - jm1(MV)) / 3.0", "hv2 = (MV - im1(MV)) / 3.0"),
    "second-order v metric": ("+ vs.recip_dxV * (6.0 * vavz", "+ vs.recip_dyV * (6.0 * vavz"),
    "second-order six": ("vs.recip_dyU * (6.0 * uavz", "vs.recip_dyU * (4.0 * uavz"),
    "free drift wind centre": ("tx = 0.5 * (vs.WindForcingX + ip1(vs.WindForcingX))", "tx = 0.5 * (vs.WindForcingX + im1(vs.WindForcingX))"),
    "free drift Coriolis sign": ("rx = -tx - mf * voc", "rx = -tx + mf * voc"),
    "free drift turning": ("ang = np.where(t4 > 0, rA - np.arctan2(t3, t2), 0)", "ang = np.where(t4 > 0, rA + np.arctan2(t3, t2), 0)"),
    "free drift quadratic": ("t4 = t2**2 + 4 * t3", "t4 = t2**2 + 2 * t3"),
    "free drift hemisphere": ("np.where(vs.fCori < 0, s.waterIceDrag_south, s.waterIceDrag)", "np.where(vs.fCori > 0, s.waterIceDrag_south, s.waterIceDrag)"),
    "free drift u average": ("u = 0.5 * (im1(uc) + uc) * vs.iceMaskU", "u = 0.5 * (ip1(uc) + uc) * vs.iceMaskU"),
    "free drift v ocean centre": ("voc = 0.5 * (vs.vOcean + jp1(vs.vOcean))", "voc = 0.5 * (vs.vOcean + jm1(vs.vOcean))"),
    "residual sig22": ("r22 = (0.5 * (s1n - s2n) - 0.5 * (s1o - s2o))", "r22 = (0.5 * (s1n + s2n) - 0.5 * (s1o + s2o))"),
    "residual z alpha": ("r12 = (s12n - s12o) * aZ", "r12 = (s12n - s12o) * aC"),
    "residual beta factor": ('ru = vs.iceMaskU * (un - uo) * trace["evpBetaU"]', 'ru = vs.iceMaskU * (un - uo) * trace["betaU"]'),
    "residual interior": ("inner = (slice(s.olx, -s.olx), slice(s.oly, -s.oly))", "inner = (slice(None), slice(None))"),
    "bulk stable psi": ("psimh = -5.0 * huol * stable", "psimh = -6.0 * huol * stable"),
    "bulk stability clip": ("huol = np.minimum(np.abs(huol), 10.0)", "huol = np.minimum(np.abs(huol), 5.0)"),
    "bulk log-law sign": ("rd = rdn / (1.0 + rdn * (aln - psimh) / s.karman)", "rd = rdn / (1.0 + rdn * (aln + psimh) / s.karman)"),
    "bulk heat psi": ("(1.0 - stable) * (2.0 * np.log(0.5 * (1.0 + xsq)))", "(1.0 - stable) * (np.log(0.5 * (1.0 + xsq)))"),
    "bulk humidity buoyancy": ("qstar / (1.0 / s.zvir + qa)", "qstar / (s.zvir + qa)"),
    "bulk lapse rate": ("deltap = ta - tsf + s.gamma_blk * ht", "deltap = ta - tsf - s.gamma_blk * ht"),
    "bulk evaporation derivative": ("devdt = clha * ssq * 2.166847e-3 / (tsf * tsf)", "devdt = clha * ssq * 2.166847e-3 / tsf"),
    "bulk long-wave derivative": ("dflwupdt = 4.0 * s.ocean_emissivity", "dflwupdt = 3.0 * s.ocean_emissivity"),
    "bulk unstable pi/2": ("- 2.0 * np.arctan(x) + np.pi * 0.5)", "- 2.0 * np.arctan(x) + np.pi * 0.25)"),
    "z-point average": ("z = c + im1(c)", "z = c + ip1(c)"),
    "Coriolis sign": ("fx = fx + 0.5 * (fv + im1(fv))", "fx = fx - 0.5 * (fv + im1(fv))"),
    "Coriolis u average": ("fu = vs.SeaIceMassC * vs.fCori * 0.5 * (u + ip1(u))", "fu = vs.SeaIceMassC * vs.fCori * 0.5 * (u + im1(u))"),
    "Coriolis v-face average": ("fy = fy - 0.5 * (fu + jm1(fu))", "fy = fy - 0.5 * (fu + jp1(fu))"),
    "stress divergence sign": ("+ fx + divx) / denU", "+ fx - divx) / denU"),
    "beta + 1": ("bp1U = betaU + s.recip_deltatDyn", "bp1U = betaU"),
    "drag v faces": ("w = 0.25 * ((du + ip1(du)) ** 2 + (dv + jp1(dv)) ** 2)", "w = 0.25 * ((du + ip1(du)) ** 2 + (dv + jm1(dv)) ** 2)"),
    "metric term k1C": ("e22 = (dvdy + uave * vs.k1AtC) * vs.maskInC", "e22 = (dvdy + vave * vs.k1AtC) * vs.maskInC"),
    "mass at u-points": ("mU = 0.5 * (mC + im1(mC))", "mU = 0.5 * (mC + ip1(mC))"),
    "area at v-points": ("return 0.5 * (vs.Area + im1(vs.Area)), 0.5 * (vs.Area + jm1(vs.Area))", "return 0.5 * (vs.Area + im1(vs.Area)), 0.5 * (vs.Area + jp1(vs.Area))"),
    "tilt direction": ("fx = fx - vs.SeaIceMassU * vs.recip_dxC * (phi - im1(phi))", "fx = fx - vs.SeaIceMassU * vs.recip_dxC * (phi - ip1(phi))"),
    "wind stress v average": ("ty = 0.5 * (ty + jm1(ty)) * vs.iceMaskV", "ty = 0.5 * (ty + jp1(ty)) * vs.iceMaskV"),
    "relative wind": ("vrel = vs.vWind - 0.5 * (vs.vIce + jp1(vs.vIce))", "vrel = vs.vWind - 0.5 * (vs.vIce + jm1(vs.vIce))"),
    "ocean stress drag": ("tu = 0.5 * (cd + im1(cd)) * cosW * du", "tu = 0.5 * (cd + ip1(cd)) * cosW * du"),
    "x CFL metric": ("cfl = np.abs(vs.uIce * s.deltatTherm * vs.recip_dxC)", "cfl = np.abs(vs.uIce * s.deltatTherm * vs.recip_dyC)"),
    "y slope mask": ("rj = (f[:, 2:-1] - f[:, 1:-2]) * mS[:, 2:-1]", "rj = (f[:, 2:-1] - f[:, 1:-2]) * mS[:, 1:-2]"),
    "x update faces": ("f = f - s.deltatTherm * vs.maskInC * vs.recip_rA * (ip1(F) - F)", "f = f - s.deltatTherm * vs.maskInC * vs.recip_rA * (F - im1(F))"),
    "thin ice boundary": ("thin = h <= s.hIce_min", "thin = h < s.hIce_min"),
    "Newton denominator": ("T = np.where(ice, T + (Fc - Fia) / (kEff + dF), 0)", "T = np.where(ice, T + (Fc - Fia) / (kEff - dF), 0)"),
    "sensible heat sign": ("Fs = np.where(ice, d1 * ug * (t1 - ta), 0)", "Fs = np.where(ice, d1 * ug * (ta - t1), 0)"),
    "latent heat sign": ("Flh = np.where(ice, d1i * ug * (qs - vs.aqh), 0)", "Flh = np.where(ice, d1i * ug * (vs.aqh - qs), 0)"),
    "snow albedo": ("alb = np.where(hS > s.hCut, aS, alb)", "alb = np.where(hS > s.hCut, aI, alb)"),
    "penetrating short wave": ("pen = np.where(ice, s.shortwave * np.exp(-1.5 * hI), 0)", "pen = np.where(ice, s.shortwave * np.exp(-1.5 * hS), 0)"),
    "McPhee taper": ("mltf = 1 + (s.McPheeTaperFac - 1) / (1 + np.exp((AP - t0) * t1))", "mltf = 1 + (s.McPheeTaperFac - 1) / (1 + np.exp((AP + t0) * t1))"),
    "snow melt energy": ("toMelt = np.where(allMelt, -hsP * s.recip_deltatTherm / qs, Fia)", "toMelt = np.where(allMelt, -hsP * s.recip_deltatTherm / qi, Fia)"),
    "lead closing": ("dAoa = np.where(grow, gOW * (1 - AP), dAoa)", "dAoa = np.where(grow, gOW * AP, dAoa)"),
    "flooding density ratio": ("hs = hs - flood * s.rhoIce2rhoSnow", "hs = hs - flood"),
    "fresh water sign": ("- (fwIce + fwSnow) / s.deltatTherm) * s.rhoFresh", "+ (fwIce + fwSnow) / s.deltatTherm) * s.rhoFresh"),
    "basal east-face speed": ("t = 0.25 * (um**2 + ip1(um) ** 2 + vm**2 + jp1(vm) ** 2)", "t = 0.25 * (um**2 + im1(um) ** 2 + vm**2 + jp1(vm) ** 2)"),
    "basal north-face speed": ("t = 0.25 * (um**2 + ip1(um) ** 2 + vm**2 + jp1(vm) ** 2)", "t = 0.25 * (um**2 + ip1(um) ** 2 + vm**2 + jm1(vm) ** 2)"),
    "basal velocity scale": ("np.sqrt(t + s.basalDragU0**2)", "np.sqrt(t + s.basalDragU0)"),
    "basal critical height": ("hcrit = np.abs(vs.R_low) * vs.Area / s.basalDragK1", "hcrit = np.abs(vs.R_low) * vs.Area * s.basalDragK1"),
    "basal depth sign": ("hcrit = np.abs(vs.R_low) * vs.Area / s.basalDragK1", "hcrit = vs.R_low * vs.Area / s.basalDragK1"),
    "basal soft maximum": ("rfac = 1.0 / fac", "rfac = 1.0"),
    "basal concentration decay": ("np.exp(-s.cBasalStar * (1.0 - vs.Area))", "np.exp(-s.cBasalStar * vs.Area)"),
    "basal area gate": ("np.where(vs.Area > 0.01, val, 0.0)", "np.where(vs.Area >= 0.01, val, 0.0)"),
    "basal drag explicit": ("denU = vs.SeaIceMassU * bp1U + vs.AreaW * (\n        0.5 * (cd + im1(cd)) * s.cosWat + 0.5 * (cb + im1(cb)))", "denU = vs.SeaIceMassU * bp1U + vs.AreaW * (\n        0.5 * (cd + im1(cd)) * s.cosWat)"),
    "turning x sign": ("        - sg * s.sinWat * 0.5 * (", "        + sg * s.sinWat * 0.5 * ("),
    "turning y sign": ("        + sg * s.sinWat * 0.5 * (", "        - sg * s.sinWat * 0.5 * ("),
    "turning x north corner": ("cd * 0.5 * (du0 + jp1(du0))", "cd * 0.5 * (du0 + jm1(du0))"),
    "turning x west cell": ("+ im1(cd) * 0.5 * (im1(du0) + jp1(im1(du0)))", "+ im1(cd) * 0.5 * (ip1(du0) + jp1(ip1(du0)))"),
    "turning y east corner": ("cd * 0.5 * (dv0 + ip1(dv0))", "cd * 0.5 * (dv0 + im1(dv0))"),
    "turning y south cell": ("+ jm1(cd) * 0.5 * (jm1(dv0) + ip1(jm1(dv0)))", "+ jm1(cd) * 0.5 * (jp1(dv0) + ip1(jp1(dv0)))"),
    "turning ice difference": ("du0 = vs.uOcean - u", "du0 = vs.uOcean + u"),
    "turning mass gate": ("lmU = np.where(vs.SeaIceMassU != 0, 1, vs.SeaIceMassU)", "lmU = np.where(vs.SeaIceMassU != 0, 1, 1)"),
    "ocean stress turning sign": ("tv = 0.5 * (cd + jm1(cd)) * cosW * dv + sg * sinW", "tv = 0.5 * (cd + jm1(cd)) * cosW * dv - sg * sinW"),
    "ocean stress turning average": ("- sg * sinW * 0.5 * (cd * dvC + jm1(cd * dvC))", "- sg * sinW * 0.5 * (cd * dvC + jp1(cd * dvC))"),
    "ocean stress c-point difference": ("duC = 0.5 * (du + ip1(du))", "duC = 0.5 * (du + im1(du))"),
    "ocean stress angle units": ("sinW = np.sin(np.deg2rad(s.waterTurnAngle))", "sinW = np.sin(s.waterTurnAngle)"),
    "air turning sign": ("tx = cda * (cosWin * urel - sg * sinWin * vrel)", "tx = cda * (cosWin * urel + sg * sinWin * vrel)"),
    "air turning angle units": ("cosWin = np.cos(np.deg2rad(s.airTurnAngle))", "cosWin = np.cos(s.airTurnAngle)"),
    "load gravity": ("(vs.surfPress + vs.SeaIceLoad * s.gravity * s.seaIceLoadFac)", "(vs.surfPress + vs.SeaIceLoad * s.seaIceLoadFac)"),
    "load density": ("+ vs.SeaIceLoad * s.gravity * s.seaIceLoadFac) * s.recip_rhoSea", "+ vs.SeaIceLoad * s.gravity * s.seaIceLoadFac) * s.rhoSea"),
    "wet albedo threshold": ("wet = (hI > 0) & (T >= Tmeltw)", "wet = (hI > 0) & (T > Tmeltw)"),
    "wet ice albedo": ("aI = np.where(wet, s.wetIceAlb, aI)", "aI = np.where(wet, s.dryIceAlb, aI)"),
    "wet snow albedo": ("aS = np.where(wet, s.wetSnowAlb, aS)", "aS = np.where(wet, s.drySnowAlb, aS)"),
    "south dry albedo": ("aI = np.where(south, s.dryIceAlb_south, aI)", "aI = np.where(south, s.dryIceAlb, aI)"),
    "south wet snow albedo": ("aS = np.where(wets, s.wetSnowAlb_south, aS)", "aS = np.where(wets, s.wetSnowAlb, aS)"),
    "south wet threshold": ("wets = (hI > 0) & (vs.fCori < 0) & (T >= Tmeltw)", "wets = (hI > 0) & (vs.fCori > 0) & (T >= Tmeltw)"),
}


def _run(name, tmp_path):
    src = ORACLE.read_text()
    a, b = MUTATIONS[name]
    assert src.count(a) == 1, f"mutation anchor not unique or missing: {name}"
    path = tmp_path / (name.replace(" ", "_").replace("/", "_") + ".py")
    path.write_text(src.replace(a, b))
    env = dict(os.environ, ORACLE_NP_PATH=str(path))
    r = subprocess.run([sys.executable, "-m", "pytest", str(ROOT / "tests" / "test_oracle_pins.py"),
                        "-x", "-q", "-p", "no:cacheprovider"], cwd=ROOT, env=env,
                       capture_output=True, text=True)
    return name, r.returncode


def test_every_planted_mistake_is_caught(tmp_path):
    with cf.ThreadPoolExecutor(max_workers=6) as ex:
        results = list(ex.map(lambda n: _run(n, tmp_path), MUTATIONS))
    missed = [n for n, rc in results if rc == 0]
    assert not missed, f"pins did not catch: {missed}"
    assert all(rc == 1 for _, rc in results), results


def test_unmutated_oracle_passes_the_pins():
    r = subprocess.run([sys.executable, "-m", "pytest", str(ROOT / "tests" / "test_oracle_pins.py"),
                        "-q", "-p", "no:cacheprovider"], cwd=ROOT, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
