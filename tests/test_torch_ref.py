"""CPU checks of tests/torch_ref.py, the torch restatement whose autograd is
the reference gradient for the HIP adjoint (the stand-in for jax.grad):

* its forward equals the NumPy oracle: masses and areas exactly, wind
  forcing and strength to an ulp (pow/exp of the two libraries), the EVP
  subcycle to rounding;
* its autograd gradient agrees with central finite differences of its own
  forward along random directions in (P* field, hIceMean, water and air drag
  per hemisphere), which also shows that the directions used by the GPU
  finite-difference test flip no branch.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import torch_ref as TR
from gpu_helpers import pre_evp
from oracle import veris_np as O
from veris_amd import synth


TURN = dict(sinWat=0.42261826174069944, cosWat=0.9063077870366499)


def _state(kind):
    if kind == "global":
        st = synth.global_latlon(31, 36, 20, -75.0, 80.0)
    elif kind == "turning_metric":
        # water-stress turning angle of 25 degrees and random metric terms
        st = synth.global_latlon(31, 36, 20, -75.0, 80.0, **TURN)
        r = np.random.default_rng(4)
        for k in ("k1AtC", "k2AtC", "k1AtZ", "k2AtZ"):
            st.variables[k] = r.normal(0, 1e-6, st.shape)
    elif kind == "basal":  # grounded keels: water 8 m deep, given as a negative depth
        st = synth.global_latlon(31, 36, 20, -75.0, 80.0, basalDragK2=15.0)
        st.variables.R_low = -np.asarray(st.variables.R_low) * 0.002
    elif kind == "sea_ice_load":  # the load in the surface tilt, with a tilted surface
        st = synth.global_latlon(31, 36, 20, -75.0, 80.0, useRealFreshWaterFlux=True)
        vs0, s0 = st.variables, st.settings
        r = np.random.default_rng(6)
        vs0.SeaIceLoad = (s0.rhoIce * np.asarray(vs0.hIceMean) + s0.rhoSnow * np.asarray(vs0.hSnowMean)
                          ) * r.uniform(0.8, 1.2, st.shape)
        vs0.ssh_an = r.normal(0.0, 0.05, st.shape)
    elif kind == "second_order":
        st = synth.global_latlon(31, 36, 20, -75.0, 80.0, secondOrderBC=True)
    elif kind == "adaptive":
        # about a tenth of the ice cells above alpha's floor
        st = synth.regional(32, 0, 30, 22, useAdaptiveEVP=True, aEvpCoeff=0.5, aEVPalphaMin=20.0)
    elif kind == "side_drag":
        st = synth.regional(32, 0, 30, 22, noSlip=False, sideDragCoeff=0.004, sideDragU0=0.02)
    elif kind == "coastline":
        st = synth.global_latlon(31, 36, 20, -75.0, 80.0, noSlip=False, use_coastline=True)
        r = np.random.default_rng(8)
        st.variables.Fu = r.uniform(0.0, 2.0, st.shape) * st.variables.iceMaskU
        st.variables.Fv = r.uniform(0.0, 2.0, st.shape) * st.variables.iceMaskV
    else:
        st = synth.regional(32, 0, 30, 22)
    pre_evp(st)
    vs = st.variables
    vs.uIce, vs.vIce, vs.sigma1, vs.sigma2, vs.sigma12 = O.evp_solver(st, n_substeps=6)
    return pre_evp(st)


def _torch_vars(st):
    return {k: torch.from_numpy(np.asarray(v, dtype=np.float64)) for k, v in st.variables.items()}


@pytest.mark.parametrize("kind", ["global", "regional", "turning_metric", "side_drag",
                                  "coastline", "basal", "adaptive", "second_order",
                                  "sea_ice_load"])
def test_forward_matches_oracle(kind):
    st = _state(kind)
    vs, s = _torch_vars(st), st.settings
    f = vs["fCori"]
    ca = torch.where(f < 0, torch.tensor(s.airIceDrag_south, dtype=torch.float64), torch.tensor(s.airIceDrag, dtype=torch.float64))
    cw = torch.where(f < 0, torch.tensor(s.waterIceDrag_south, dtype=torch.float64), torch.tensor(s.waterIceDrag, dtype=torch.float64))
    p = TR.pre_evp(vs, s, torch.full(st.shape, float(s.pStar), dtype=torch.float64),
                   vs["hIceMean"], ca)
    for name in ("SeaIceMassC", "SeaIceMassU", "SeaIceMassV", "AreaW", "AreaS"):
        np.testing.assert_array_equal(p[name].numpy(), st.variables[name], err_msg=name)
    # torch's pow/exp and NumPy's differ by an ulp here and there
    for name in ("WindForcingX", "WindForcingY", "SeaIceStrength"):
        ref = np.asarray(st.variables[name])
        np.testing.assert_allclose(p[name].numpy(), ref, rtol=2e-15,
                                   atol=1e-15 * np.abs(ref).max(), err_msg=name)
    out = TR.evp(vs, s, p, 30, cw)
    want = O.evp_solver(st, n_substeps=30)
    for q in range(5):
        scale = np.abs(want[q]).max()
        assert np.abs(out[q].detach().numpy() - want[q]).max() <= 1e-13 * scale, q


@pytest.mark.parametrize("kind", ["global", "regional", "turning_metric", "side_drag",
                                  "coastline", "basal", "adaptive", "second_order",
                                  "sea_ice_load"])
def test_autograd_matches_finite_differences(kind):
    st = _state(kind)
    vs, s = _torch_vars(st), st.settings
    n = 20
    rng = np.random.default_rng(5)
    w = [torch.from_numpy(rng.standard_normal(st.shape)) for _ in range(5)]
    base = {"pStar": torch.from_numpy(s.pStar * (1 + 0.2 * rng.random(st.shape))),
            "h": vs["hIceMean"].clone(),
            "cwN": torch.tensor(float(s.waterIceDrag), dtype=torch.float64),
            "cwS": torch.tensor(float(s.waterIceDrag_south) * 1.1, dtype=torch.float64),
            "caN": torch.tensor(float(s.airIceDrag), dtype=torch.float64),
            "caS": torch.tensor(float(s.airIceDrag_south) * 0.9, dtype=torch.float64)}
    f = vs["fCori"]

    # the scale factors are data dependent: freeze them at the base point
    with torch.no_grad():
        ca = torch.where(f < 0, base["caS"], base["caN"])
        cw = torch.where(f < 0, base["cwS"], base["cwN"])
        outs0 = TR.evp(vs, s, TR.pre_evp(vs, s, base["pStar"], base["h"], ca), n, cw)
        frozen = [float(o.abs().max()) or 1.0 for o in outs0]

    def Jf(p):
        ca = torch.where(f < 0, p["caS"], p["caN"])
        cw = torch.where(f < 0, p["cwS"], p["cwN"])
        outs = TR.evp(vs, s, TR.pre_evp(vs, s, p["pStar"], p["h"], ca), n, cw)
        return float(sum(((o / sc) * wq)[2:-2, 2:-2].sum() for o, sc, wq in zip(outs, frozen, w)))

    leaves2 = {k: v.clone().requires_grad_(True) for k, v in base.items()}
    ca = torch.where(f < 0, leaves2["caS"], leaves2["caN"])
    cw = torch.where(f < 0, leaves2["cwS"], leaves2["cwN"])
    outs = TR.evp(vs, s, TR.pre_evp(vs, s, leaves2["pStar"], leaves2["h"], ca), n, cw)
    sum(((o / sc) * wq)[2:-2, 2:-2].sum() for o, sc, wq in zip(outs, frozen, w)).backward()
    grads = {k: v.grad for k, v in leaves2.items()}
    south = bool((f < 0).any())
    for d in range(4):
        dirs = {k: torch.from_numpy(rng.standard_normal(tuple(v.shape))) * v
                for k, v in base.items()}
        # adaptive EVP: alpha's floor and the mass clamp put kinks within
        # relative distance 1e-6 of this state (the FD error grows linearly
        # with the step there); a smaller step stays clear of them
        eps = 1e-7 if kind == "adaptive" else 1e-6
        fd = (Jf({k: base[k] + eps * dirs[k] for k in base}) -
              Jf({k: base[k] - eps * dirs[k] for k in base})) / (2 * eps)
        ad = sum(float((grads[k] * dirs[k]).sum()) for k in base)
        assert abs(fd - ad) <= 1e-6 * abs(ad), (d, fd, ad)
    assert float(grads["cwS"].abs()) > 0 if south else float(grads["cwS"].abs()) == 0


def test_torch_advection_equals_the_oracle():
    """The torch restatement of the advection (the stand-in for jax.grad of
    the reference's) reproduces the oracle's values on moving ice, on a
    global, a regional and a doubly periodic domain."""
    from gpu_helpers import pre_evp

    for st in (synth.config0(), synth.regional(8, 1, 40, 33), synth.tiny(16, 12, veros_fill=False)):
        st = pre_evp(st)
        vs = st.variables
        vs.uIce, vs.vIce, vs.sigma1, vs.sigma2, vs.sigma12 = O.evp_solver(st, n_substeps=60)
        want = O.Advection(st)
        tv = {k: torch.from_numpy(np.asarray(a, dtype=np.float64)) for k, a in vs.items()}
        for name, w in zip(("hIceMean", "hSnowMean", "Area"), want):
            got = TR.advection(tv, st.settings, tv["uIce"], tv["vIce"], tv[name]).numpy()
            scale = max(float(np.abs(w).max()), 1e-300)
            assert float(np.abs(got - w).max()) <= 1e-14 * scale, name


def test_torch_advection_gradient_against_central_differences():
    """The derivative convention of the advection adjoint (the branch the
    forward takes at every selection) agrees with central differences along
    relative perturbations of the velocities and the fields, which keep zero
    cells zero and so flip no branch."""
    st = pre_evp(synth.tiny(16, 12))
    vs = st.variables
    vs.uIce, vs.vIce, vs.sigma1, vs.sigma2, vs.sigma12 = O.evp_solver(st, n_substeps=60)
    tv = {k: torch.from_numpy(np.asarray(a, dtype=np.float64)) for k, a in vs.items()}
    rng = np.random.default_rng(31)
    w = torch.from_numpy(rng.standard_normal(st.shape))
    names = ("uIce", "vIce", "hIceMean")

    def J(x):
        return (TR.advection(tv, st.settings, x["uIce"], x["vIce"], x["hIceMean"]) * w).sum()

    leaves = {n: tv[n].clone().requires_grad_(True) for n in names}
    J(leaves).backward()
    for d in range(5):
        dirs = {n: torch.from_numpy(rng.standard_normal(st.shape)) * tv[n] for n in names}
        eps = 1e-7
        with torch.no_grad():
            fd = (float(J({n: tv[n] + eps * dirs[n] for n in names}))
                  - float(J({n: tv[n] - eps * dirs[n] for n in names}))) / (2 * eps)
        ad = sum(float((leaves[n].grad * dirs[n]).sum()) for n in names)
        assert abs(fd - ad) <= 1e-6 * abs(ad), (d, fd, ad)
