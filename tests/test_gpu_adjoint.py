"""GPU checks of the hand-written EVP adjoint (vh_evp_forward_checkpointed,
vh_evp_adjoint, vh_pre_evp_adjoint) through veris_amd.adjoint.evp_subcycle.

* The checkpointed forward equals the drop-in path (pre-EVP kernels +
  vh_evp_solver) bit for bit.
* Gradients of a scalar objective of all five outputs w.r.t. P* (per cell),
  hIceMean (per cell) and the four drag coefficients, at default settings,
  with the water-stress turning angle and the metric terms, and under free
  slip with lateral drag (land-neighbour and coastline form factors), equal
  torch.autograd of
  an independent torch FP64 restatement (tests/torch_ref.py, the stand-in for
  the reference's jax.grad) to 1e-9 relative: both are FP64 and differ only
  in rounding order.
* Central finite differences of the GPU forward along 20 random directions
  match the directional derivative to 1e-5 (the 20 directions perturb all
  parameters at once, with a step small enough not to flip a branch).
* The checkpoint interval changes nothing: gradients are bit-identical.
* FP32 state (Delta and zeta in FP64 in the forward) against the FP64
  adjoint, at the tolerance stated in the test.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import torch_ref as TR
from gpu_helpers import assert_close, assert_exact, g2n, pre_evp, report
from oracle import veris_np as O
from veris_amd import _native as N
from veris_amd import adjoint as A
from veris_amd import kernels as K
from veris_amd import synth

pytestmark = pytest.mark.gpu

NAMES = ("uIce", "vIce", "sigma1", "sigma2", "sigma12")


def _state(kind):
    if kind == "global":
        st = synth.global_latlon(31, 36, 20, -75.0, 80.0)
    elif kind == "doubly_periodic":
        st = synth.global_latlon(31, 36, 20, -75.0, 80.0, veros_fill=False)
    elif kind == "regional":
        st = synth.regional(32, 0, 30, 22)
    elif kind == "sea_ice_load":  # the load in the surface tilt, with a tilted surface
        st = synth.global_latlon(31, 36, 20, -75.0, 80.0, useRealFreshWaterFlux=True)
        vs0, s0 = st.variables, st.settings
        r = np.random.default_rng(6)
        vs0.SeaIceLoad = (s0.rhoIce * np.asarray(vs0.hIceMean) + s0.rhoSnow * np.asarray(vs0.hSnowMean)
                          ) * r.uniform(0.8, 1.2, st.shape)
        vs0.ssh_an = r.normal(0.0, 0.05, st.shape)
    elif kind == "second_order":  # the second-order no-slip boundary term at the coasts
        st = synth.global_latlon(31, 36, 20, -75.0, 80.0, secondOrderBC=True)
    elif kind == "adaptive":  # about a tenth of the ice cells above alpha's floor
        st = synth.regional(32, 0, 30, 22, useAdaptiveEVP=True, aEvpCoeff=0.5, aEVPalphaMin=20.0)
    elif kind == "basal":  # grounded keels: water 8 m deep, given as a negative depth
        st = synth.global_latlon(31, 36, 20, -75.0, 80.0, basalDragK2=15.0)
        st.variables.R_low = -np.asarray(st.variables.R_low) * 0.002
    elif kind == "side_drag":  # free slip with lateral drag from land neighbours
        st = synth.regional(32, 0, 30, 22, noSlip=False, sideDragCoeff=0.004, sideDragU0=0.02)
    elif kind == "coastline":  # free slip with lateral drag from coastline form factors
        st = synth.global_latlon(31, 36, 20, -75.0, 80.0, noSlip=False, use_coastline=True)
        r = np.random.default_rng(8)
        st.variables.Fu = r.uniform(0.0, 2.0, st.shape) * st.variables.iceMaskU
        st.variables.Fv = r.uniform(0.0, 2.0, st.shape) * st.variables.iceMaskV
    elif kind == "unpacked":
        # metrics varying along x: the kernels read the full constant fields
        # instead of the packed mask bytes and metric rows
        st = synth.global_latlon(31, 36, 20, -75.0, 80.0)
        ramp = 1.0 + 1e-3 * np.arange(st.shape[0])[:, None]
        for k in ("recip_dxU", "recip_dyV", "rAz", "dxV"):
            st.variables[k] = np.asarray(st.variables[k]) * ramp
    elif kind == "turning_metric":
        # water-stress turning angle of 25 degrees and random metric terms
        st = synth.global_latlon(31, 36, 20, -75.0, 80.0, sinWat=0.42261826174069944,
                                 cosWat=0.9063077870366499)
        r = np.random.default_rng(4)
        for k in ("k1AtC", "k2AtC", "k1AtZ", "k2AtZ"):
            st.variables[k] = r.normal(0, 1e-6, st.shape)
    else:
        raise ValueError(kind)
    pre_evp(st)
    vs = st.variables  # spin up: non-zero velocities and stresses
    vs.uIce, vs.vIce, vs.sigma1, vs.sigma2, vs.sigma12 = O.evp_solver(st, n_substeps=6)
    return pre_evp(st)


def _weights(st, seed):
    """Cotangent weights of the objective, one per output, scaled so that
    every output contributes at the same order."""
    rng = np.random.default_rng(seed)
    out = O.evp_solver(st, n_substeps=1)
    ws = []
    for q in range(5):
        w = rng.standard_normal(st.shape)
        w[:2] = w[-2:] = 0
        w[:, :2] = w[:, -2:] = 0
        ws.append(w / max(float(np.abs(out[q]).max()), 1e-30))
    return ws


def _objective(outs, ws):
    return sum((o * w).sum() for o, w in zip(outs, ws)) + 0.5 * sum(
        ((o * w) ** 2).sum() for o, w in zip(outs[:2], ws[:2]))


def _params(st, device, dtype=torch.float64, requires_grad=True):
    s = st.settings
    rng = np.random.default_rng(7)
    pstar = s.pStar * (1 + 0.2 * rng.random(st.shape))
    mk = lambda a: torch.tensor(a, dtype=dtype, device=device).requires_grad_(requires_grad)
    return dict(pStar=mk(pstar), hIceMean=mk(np.asarray(st.variables.hIceMean)),
                waterIceDrag=mk(s.waterIceDrag), waterIceDrag_south=mk(s.waterIceDrag_south * 1.1),
                airIceDrag=mk(s.airIceDrag), airIceDrag_south=mk(s.airIceDrag_south * 0.9))


def _gpu_grads(st, n, ckpt, dtype=torch.float64, seed=3):
    dev = torch.device("cuda")
    gst = st.to(dev, dtype)
    p = _params(st, dev, dtype=torch.float64)
    p["hIceMean"] = p["hIceMean"].detach().to(dtype).requires_grad_(True)
    outs = A.evp_subcycle(gst, n_substeps=n, ckpt_every=ckpt, **p)
    ws = [torch.tensor(w, dtype=dtype, device=dev) for w in _weights(st, seed)]
    J = _objective(outs, ws)
    J.backward()
    return {k: g2n(v.grad) for k, v in p.items()}, [g2n(o) for o in outs], float(J.detach())


def _ref_grads(st, n, seed=3):
    vs = {k: torch.from_numpy(np.asarray(v, dtype=np.float64)) for k, v in st.variables.items()}
    s = st.settings
    p = _params(st, "cpu")
    f = vs["fCori"]
    ca = torch.where(f < 0, p["airIceDrag_south"], p["airIceDrag"])
    cw = torch.where(f < 0, p["waterIceDrag_south"], p["waterIceDrag"])
    pre = TR.pre_evp(vs, s, p["pStar"], p["hIceMean"], ca)
    outs = TR.evp(vs, s, pre, n, cw)
    ws = [torch.from_numpy(w) for w in _weights(st, seed)]
    J = _objective(outs, ws)
    J.backward()
    return {k: v.grad.numpy() for k, v in p.items()}, [o.detach().numpy() for o in outs], float(J.detach())


@pytest.mark.parametrize("kind", ["global", "doubly_periodic", "regional", "turning_metric",
                                  "side_drag", "coastline", "basal", "adaptive", "second_order"])
def test_checkpointed_forward_is_the_drop_in(cuda, kind):
    st = _state(kind)
    gst = st.to(cuda)
    vs = gst.variables
    vs.SeaIceMassC, vs.SeaIceMassU, vs.SeaIceMassV = K.SeaIceMass(gst)
    vs.AreaW, vs.AreaS = K.AreaWS(gst)
    vs.WindForcingX, vs.WindForcingY = K.WindForcingXY(gst)
    vs.SeaIceStrength = K.SeaIceStrength(gst)
    for name in ("SeaIceMassC", "SeaIceMassU", "SeaIceMassV", "AreaW", "AreaS", "WindForcingX",
                 "WindForcingY"):
        assert_exact(f"{kind} pre-EVP {name}", g2n(vs[name]), st.variables[name])
    # exp() of the device math library differs from NumPy's by an ulp here and there
    assert_close(f"{kind} pre-EVP SeaIceStrength", g2n(vs.SeaIceStrength),
                 st.variables.SeaIceStrength, 1e-15)
    for n, ckpt in ((23, 10), (7, 3), (5, 10)):
        outs = A.evp_subcycle(gst, n_substeps=n, ckpt_every=ckpt)
        want = K.evp_solver(gst, n_substeps=n)
        for q in range(5):
            assert_exact(f"{kind} {NAMES[q]} n={n} ckpt={ckpt}", g2n(outs[q]), g2n(want[q]))


def test_adjoint_leaves_the_drop_in_call_context_alone(cuda):
    """The adjoint's forward starts from a copy of the drop-ins' cached call
    context and overrides drags, thickness, forcing and launch shape in the
    copy only: drop-in calls after it give the bits they gave before it."""
    st = _state("global")
    gst = st.to(cuda)
    vs = gst.variables
    vs.SeaIceMassC, vs.SeaIceMassU, vs.SeaIceMassV = K.SeaIceMass(gst)
    vs.AreaW, vs.AreaS = K.AreaWS(gst)
    vs.WindForcingX, vs.WindForcingY = K.WindForcingXY(gst)
    vs.SeaIceStrength = K.SeaIceStrength(gst)
    before = [t.clone() for t in K.evp_solver(gst, n_substeps=7)]
    s = st.settings
    outs = A.evp_subcycle(gst, n_substeps=7, ckpt_every=3, waterIceDrag=2.0 * s.waterIceDrag,
                          airIceDrag=0.5 * s.airIceDrag, pStar=torch.tensor(1.5 * s.pStar),
                          hIceMean=vs.hIceMean * 1.1)
    assert not torch.equal(outs[0], before[0])
    after = K.evp_solver(gst, n_substeps=7)
    for q in range(5):
        assert torch.equal(after[q], before[q]), NAMES[q]


@pytest.mark.parametrize("kind", ["global", "doubly_periodic", "regional", "turning_metric",
                                  "unpacked", "side_drag", "coastline", "basal", "adaptive",
                                  "second_order", "sea_ice_load"])
def test_adjoint_matches_torch_autograd(cuda, kind):
    st = _state(kind)
    n = 25
    got, outs, J = _gpu_grads(st, n, ckpt=10)
    want, routs, Jr = _ref_grads(st, n)
    for q in range(5):
        report(f"{kind} forward {NAMES[q]}", outs[q], routs[q])
    assert abs(J - Jr) <= 1e-11 * abs(Jr)
    for k in want:
        g, w = np.asarray(got[k]), np.asarray(want[k])
        err = float(np.linalg.norm(g - w) / max(np.linalg.norm(w), 1e-300))
        print(f"{kind} d/d{k}: rel L2 err {err:.3e} |grad| {np.linalg.norm(w):.3e}", flush=True)
        if k.endswith("_south") and not (np.asarray(st.variables.fCori) < 0).any():
            # a domain without southern cells
            assert np.linalg.norm(w) == 0 and np.linalg.norm(g) == 0
            continue
        assert np.linalg.norm(w) > 0, k
        assert err <= 1e-9, f"{kind} d/d{k}: {err:.3e}"


@pytest.mark.parametrize("kind", ["turning_metric", "coastline", "basal", "adaptive",
                                  "second_order"])
def test_every_reverse_output_through_the_c_abi(cuda, kind):
    """vh_evp_adjoint called directly with every output requested, including
    the optional AreaW/AreaS and uNm1/vNm1 adjoints that the autograd wrapper
    leaves out, against torch autograd with the pre-EVP fields, the initial
    state and the velocities of the previous time level as separate leaves;
    turning angle and metric terms, free slip with coastline side drag, or
    basal drag (with the thickness adjoint of its coefficient).  Bound: 1e-9
    relative L2 (both FP64)."""
    st = _state(kind)
    n, ce = 12, 5
    gst = st.to(cuda)
    vs = gst.variables
    vs.SeaIceMassC, vs.SeaIceMassU, vs.SeaIceMassV = K.SeaIceMass(gst)
    vs.AreaW, vs.AreaS = K.AreaWS(gst)
    vs.WindForcingX, vs.WindForcingY = K.WindForcingXY(gst)
    vs.SeaIceStrength = K.SeaIceStrength(gst)
    c = K.Ctx(gst)
    nbytes = ctypes.c_size_t(0)
    N.call("vh_evp_adjoint_workspace_bytes", ctypes.byref(c.dom), n, ce, ctypes.byref(nbytes))
    ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device=cuda)
    out = torch.empty((5,) + c.shape, dtype=torch.float64, device=cuda).unbind(0)
    N.call("vh_evp_forward_checkpointed", *c.args(), n, ce, *(t.data_ptr() for t in out),
           ws.data_ptr(), ws.numel(), c.stream)
    w = [torch.tensor(x, device=cuda) for x in _weights(st, 5)]
    ebuf = torch.empty((len(N.EVP_ADJOINT_NAMES),) + c.shape, dtype=torch.float64, device=cuda)
    evp = dict(zip(N.EVP_ADJOINT_NAMES, ebuf.unbind(0)))
    eo = N.EvpAdjointOut(**{"adj_" + k: v.data_ptr() for k, v in evp.items()})
    N.call("vh_evp_adjoint", *c.args(), n, ce, *(x.data_ptr() for x in w), ctypes.byref(eo),
           ws.data_ptr(), ws.numel(), c.stream)
    torch.cuda.synchronize()

    # torch: J = sum of the outputs weighted by w, every input a leaf
    tv = {k: torch.from_numpy(np.asarray(v, dtype=np.float64)) for k, v in st.variables.items()}
    s = st.settings
    f = tv["fCori"]
    pre = {k: torch.from_numpy(np.asarray(st.variables[k], dtype=np.float64)).requires_grad_(True)
           for k in ("SeaIceMassC", "SeaIceMassU", "SeaIceMassV", "AreaW", "AreaS",
                     "WindForcingX", "WindForcingY", "SeaIceStrength", "hIceMean")}
    cwN = torch.tensor(float(s.waterIceDrag), dtype=torch.float64, requires_grad=True)
    cwS = torch.tensor(float(s.waterIceDrag_south), dtype=torch.float64, requires_grad=True)
    cw = torch.where(f < 0, cwS, cwN)
    st0 = {k: tv[k].clone().requires_grad_(True) for k in NAMES}
    u0 = tv["uIce"].clone().requires_grad_(True)
    v0 = tv["vIce"].clone().requires_grad_(True)
    q = [st0[k] for k in NAMES]
    for _ in range(n):
        q = list(TR.evp_substep(tv, s, pre, *q[:2], u0, v0, *q[2:], cw))
    J = sum((o * torch.from_numpy(g2n(x))).sum() for o, x in zip(q, w))
    J.backward()
    want = {**{k: st0[k].grad for k in NAMES}, "uNm1": u0.grad, "vNm1": v0.grad,
            **{k: pre[k].grad for k in pre}}
    if not st.settings.basalDragK2:  # the thickness enters through the basal drag only
        assert want.pop("hIceMean") is None
    for k, ref in want.items():
        got = g2n(evp[k])
        err = float(np.linalg.norm(got - ref.numpy()) / max(float(ref.norm()), 1e-300))
        print(f"adj_{k}: rel L2 err {err:.3e} |grad| {float(ref.norm()):.3e}", flush=True)
        assert float(ref.norm()) > 0, k
        assert err <= 1e-9, k
    for k, leaf in (("waterDrag_north", cwN), ("waterDrag_south", cwS)):
        got = float(evp[k].double().sum())
        ref = float(leaf.grad)
        print(f"adj_{k} summed: {got:.12e} torch {ref:.12e}", flush=True)
        assert abs(got - ref) <= 1e-9 * abs(ref), k


def _periodic_params(st, p):
    """The per-cell parameters with their x overlap filled periodically (as the
    fill leaves a model field), so that two-substep launches apply."""
    for k in ("pStar", "hIceMean"):
        a = p[k].detach().cpu().numpy().copy()
        a[:2] = a[-4:-2]
        a[-2:] = a[2:4]
        p[k] = torch.tensor(a, dtype=p[k].dtype, device=p[k].device).requires_grad_(True)
    return p


@pytest.mark.parametrize("kind", ["global", "regional"])
def test_two_substep_forward_launches_change_no_gradient(cuda, kind, monkeypatch):
    """Under the default settings with an even subcycle and checkpoint
    interval, the checkpointed forward and the recompute may run two
    substeps per launch (the recompute keeps the intermediate state): on a
    domain without a filled overlap by themselves, on a filled one when the
    caller names a two-substep tile that pair_tile confirms.  Every gradient
    equals the one-substep-per-launch path bit for bit, and so do the
    returned velocities; the returned stresses too, except in the filled
    overlap columns (periodic values there instead of the reference's
    wrap-around values)."""
    dev = torch.device("cuda")
    st = _state(kind)
    gst = st.to(dev)
    p = _periodic_params(st, _params(st, dev))
    tile = A.pair_tile(gst, pStar=p["pStar"], hIceMean=p["hIceMean"])
    assert tile > 0
    ws = [torch.tensor(w, device=dev) for w in _weights(st, 3)]

    def run(t):
        q = {k: v.detach().clone().requires_grad_(True) for k, v in p.items()}
        outs = A.evp_subcycle(gst, n_substeps=24, ckpt_every=10, tile=t, **q)
        _objective(outs, ws).backward()
        return {k: g2n(v.grad) for k, v in q.items()}, [g2n(o) for o in outs]

    monkeypatch.setenv("VERIS_ADJ_PAIRS", "0")
    g1, o1 = run(0)
    monkeypatch.setenv("VERIS_ADJ_PAIRS", "1")
    g2, o2 = run(tile if kind == "global" else 0)
    for k in g1:
        assert np.array_equal(g1[k], g2[k]), k
    cyclic = st.settings.enable_cyclic_x or not st.settings.veros_fill
    keep = np.arange(2, st.shape[0] - 2) if cyclic else np.arange(st.shape[0])
    for q in range(5):
        a, b = (o1[q], o2[q]) if q < 2 else (o1[q][keep], o2[q][keep])
        assert_exact(f"{kind} {NAMES[q]} pairs", b, a)


def test_pair_tile_refuses_inputs_without_periodic_overlaps(cuda):
    """A per-cell P* that is not periodic in the x overlap (or settings with
    an optional branch) gives no two-substep tile; an explicit two-substep
    tile under such settings is refused."""
    dev = torch.device("cuda")
    st = _state("global")
    gst = st.to(dev)
    p = _params(st, dev)
    assert A.pair_tile(gst, pStar=p["pStar"], hIceMean=p["hIceMean"]) == 0
    p = _periodic_params(st, p)
    tile = A.pair_tile(gst, pStar=p["pStar"], hIceMean=p["hIceMean"])
    assert tile > 0
    st2 = _state("adaptive")
    g2 = st2.to(dev)
    assert A.pair_tile(g2) == 0
    with pytest.raises(N.VerisHipError, match="VH_EUNSUPPORTED"):
        A.evp_subcycle(g2, n_substeps=24, ckpt_every=10, tile=tile)


def test_checkpoint_interval_is_irrelevant(cuda):
    st = _state("global")
    g1, _, _ = _gpu_grads(st, 24, ckpt=1)
    g2, _, _ = _gpu_grads(st, 24, ckpt=7)
    g3, _, _ = _gpu_grads(st, 24, ckpt=24)
    for k in g1:
        assert np.array_equal(g1[k], g2[k]) and np.array_equal(g1[k], g3[k]), k


@pytest.mark.parametrize("kind", ["global", "turning_metric", "coastline", "basal", "adaptive",
                                  "second_order"])
def test_adjoint_finite_differences(cuda, kind):
    """20 random directions in (P*, h, Cw_N, Cw_S, Ca_N, Ca_S) space; central
    differences of the GPU forward with relative step 1e-6."""
    st = _state(kind)
    n = 40
    got, _, _ = _gpu_grads(st, n, ckpt=10)
    gst = st.to(cuda)
    base = _params(st, cuda, requires_grad=False)
    ws = [torch.tensor(w, device=cuda) for w in _weights(st, 3)]
    rng = np.random.default_rng(11)
    worst = 0.0
    for d in range(20):
        # relative perturbations: h stays 0 where it is 0 (no mass appears
        # on an ice-free cell, which would flip the denominator guard)
        dirs = {k: torch.tensor(rng.standard_normal(tuple(v.shape)), device=cuda,
                                dtype=torch.float64) * v for k, v in base.items()}
        # adaptive EVP: alpha's floor puts kinks close to the state; a
        # smaller step stays clear of them
        eps = 1e-7 if kind == "adaptive" else 1e-6

        def J(sign):
            p = {k: base[k] + sign * eps * dirs[k] for k in base}
            with torch.no_grad():
                return float(_objective(A.evp_subcycle(gst, n_substeps=n, ckpt_every=10, **p), ws))

        fd = (J(1) - J(-1)) / (2 * eps)
        ad = sum(float(np.sum(np.asarray(got[k]) * g2n(dirs[k]))) for k in base)
        rel = abs(fd - ad) / max(abs(ad), 1e-300)
        worst = max(worst, rel)
        print(f"direction {d}: fd={fd:.10e} adjoint={ad:.10e} rel={rel:.2e}", flush=True)
    assert worst <= 1e-5


def test_fp32_adjoint_against_fp64(cuda):
    """FP32 state and adjoint against the FP64 adjoint, 25 substeps.
    Bound: 2e-2 relative L2 per gradient."""
    st = _state("global")
    g64, _, _ = _gpu_grads(st, 25, ckpt=10)
    g32, _, _ = _gpu_grads(st, 25, ckpt=10, dtype=torch.float32)
    for k in g64:
        w, g = np.asarray(g64[k]), np.asarray(g32[k], dtype=np.float64)
        err = float(np.linalg.norm(g - w) / max(np.linalg.norm(w), 1e-300))
        print(f"fp32 d/d{k}: rel L2 err {err:.3e}", flush=True)
        assert err <= 2e-2, k


def test_adjoint_rejects_unsupported(cuda):
    from veris_amd._native import VerisHipError

    st = _state("global")
    st.settings.useEVP = False  # free drift has no subcycle to differentiate
    with pytest.raises(VerisHipError, match="VH_EUNSUPPORTED"):
        A.evp_subcycle(st.to(cuda), n_substeps=3)


def test_backward_after_the_state_is_reassigned(cuda):
    """The forward saves every tensor whose device pointer went into the ABI
    structs: reassigning the state's variables between forward and backward
    (as a step does) and filling new allocations of the same size must not
    change the gradients; a second backward without retain_graph raises
    autograd's own error, and one with retain_graph gives the same result."""
    st = _state("global")
    n = 17
    want, _, _ = _gpu_grads(st, n, ckpt=5)

    gst = st.to(cuda)
    p = _params(st, cuda)
    outs = A.evp_subcycle(gst, n_substeps=n, ckpt_every=5, **p)
    ws = [torch.tensor(w, device=cuda) for w in _weights(st, 3)]
    J = _objective(outs, ws)
    vs = gst.variables
    for name in list(vs.keys()):  # the pattern of a step: every field replaced
        vs[name] = torch.full_like(vs[name], 7.0)
    junk = [torch.full((1 << 20,), 3.0, dtype=torch.float64, device=cuda) for _ in range(64)]
    J.backward(retain_graph=True)
    got = {k: g2n(v.grad) for k, v in p.items()}
    for k in want:
        assert np.array_equal(got[k], want[k]), k
    for v in p.values():
        v.grad = None
    J.backward()
    for k in want:
        assert np.array_equal(g2n(p[k].grad), want[k]), k
    with pytest.raises(RuntimeError, match="backward through the graph a second time"):
        J.backward()
    del junk


def _batched_vs_single(states, dtype, n, ckpt, members_to_check):
    """Gradients of sum over members of |u|^2 + |v|^2 w.r.t. a packed P* field,
    the packed hIceMean and the shared drags: the batched adjoint against the
    single-domain one member by member."""
    from veris_amd.batched import Ensemble

    ens = Ensemble(states, device="cuda", dtype=dtype)
    rng = np.random.default_rng(12)
    s = ens.settings
    pst = torch.from_numpy(s.pStar * (1 + 0.2 * rng.random(ens.total))).cuda()
    drag = dict(waterIceDrag=s.waterIceDrag, waterIceDrag_south=s.waterIceDrag_south * 1.1,
                airIceDrag=s.airIceDrag, airIceDrag_south=s.airIceDrag_south * 0.9)

    def leaves(p0, h0):
        d = {k: torch.tensor(v, dtype=torch.float64, device="cuda", requires_grad=True)
             for k, v in drag.items()}
        return p0.clone().requires_grad_(True), h0.clone().requires_grad_(True), d

    def loss(outs):
        return (outs[0].double() ** 2).sum() + (outs[1].double() ** 2).sum()

    pb, hb, db = leaves(pst, ens.fields["hIceMean"].double())
    outs = A.evp_subcycle_batched(ens, pStar=pb, hIceMean=hb, n_substeps=n, ckpt_every=ckpt, **db)
    loss(outs).backward()
    drag_sum = {k: 0.0 for k in drag}
    for m in range(ens.n):
        st = ens.member_state(m)
        ps, hs, ds = leaves(ens.view(pst, m), ens.view(ens.fields["hIceMean"].double(), m))
        o1 = A.evp_subcycle(st, pStar=ps, hIceMean=hs, n_substeps=n, ckpt_every=ckpt, **ds)
        loss(o1).backward()
        for k in drag:
            drag_sum[k] += float(ds[k].grad)
        if m not in members_to_check:
            continue
        for q in range(5):
            assert torch.equal(ens.view(outs[q], m), o1[q]), f"member {m} output {q}"
        assert torch.equal(ens.view(pb.grad, m), ps.grad), f"member {m} d/dP*"
        assert torch.equal(ens.view(hb.grad, m), hs.grad), f"member {m} d/dh"
    for k in drag:
        got = float(db[k].grad)
        print(f"batched d/d{k}: {got:.12e}, sum of members {drag_sum[k]:.12e}", flush=True)
        assert abs(got - drag_sum[k]) <= 1e-10 * max(abs(drag_sum[k]), 1e-300), k


def test_batched_adjoint_equals_single_domain(cuda):
    """The batched adjoint of equal members (one launch sequence for all)
    gives every member's outputs and per-cell gradients bit for bit as the
    single-domain adjoint does, and the shared-drag gradients as the sum over
    the members; x-cyclic and closed members, FP64 and FP32."""
    for dtype in (torch.float64, torch.float32):
        closed = [pre_evp(synth.regional(33, m, 30, 22)) for m in range(3)]
        _batched_vs_single(closed, dtype, 25, 10, members_to_check=(0, 1, 2))
        cyclic = [pre_evp(synth.global_latlon(34, 36, 20, -75.0, 80.0, member=m)) for m in range(2)]
        _batched_vs_single(cyclic, dtype, 23, 7, members_to_check=(0, 1))


def _moving(kind):
    """A state whose ice has moved for 60 substeps (the advecting velocities)."""
    st = _state(kind)
    vs = st.variables
    vs.uIce, vs.vIce, vs.sigma1, vs.sigma2, vs.sigma12 = O.evp_solver(st, n_substeps=60)
    return st


@pytest.mark.parametrize("kind", ["global", "doubly_periodic", "regional"])
def test_advection_adjoint_matches_torch_autograd(cuda, kind):
    """vh_advection_adjoint (through veris_amd.adjoint.advect) against
    torch.autograd of the torch restatement of the advection: gradients of a
    weighted sum of the three advected fields w.r.t. the ice velocities and
    the three input fields, FP64, 1e-12 relative L2 each; the forward equals
    the drop-in bit for bit."""
    st = _moving(kind)
    gst = st.to(cuda)
    rng = np.random.default_rng(19)
    wts = [rng.standard_normal(st.shape) for _ in range(3)]
    names = ("uIce", "vIce", "hIceMean", "hSnowMean", "Area")
    leaves = {n: gst.variables[n].detach().clone().requires_grad_(True) for n in names}
    outs = A.advect(gst, **leaves)
    for a, b in zip(outs, K.Advection(gst)):
        assert torch.equal(a, b)
    sum((o * torch.from_numpy(w).to(cuda)).sum() for o, w in zip(outs, wts)).backward()
    tv = {k: torch.from_numpy(np.asarray(v, dtype=np.float64)) for k, v in st.variables.items()}
    tl = {n: tv[n].clone().requires_grad_(True) for n in names}
    J = sum((TR.advection(tv, st.settings, tl["uIce"], tl["vIce"], tl[f]) * torch.from_numpy(w)).sum()
            for f, w in zip(("hIceMean", "hSnowMean", "Area"), wts))
    J.backward()
    for n in names:
        g, w = g2n(leaves[n].grad), tl[n].grad.numpy()
        err = float(np.linalg.norm(g - w) / max(np.linalg.norm(w), 1e-300))
        print(f"{kind} advection d/d{n}: rel L2 err {err:.3e} |grad| {np.linalg.norm(w):.3e}",
              flush=True)
        assert np.linalg.norm(w) > 0 and err <= 1e-12, (n, err)


def test_thickness_misfit_gradients(cuda):
    """The thickness misfit of a step's dynamics: the EVP subcycle moves the
    ice (velocities depend on P*, the drags and the initial thickness through
    mass and strength), the advection carries hIceMean and Area; loss =
    |h' - h_target|^2 + |A' - A_target|^2 over the interior.  Gradients w.r.t.
    P* (per cell), the initial hIceMean and the four drags against
    torch.autograd of the torch restatements (1e-9), and central differences
    of the GPU forward along 10 random directions (1e-5)."""
    st = _state("global")
    n = 40
    gst = st.to(cuda)
    rng = np.random.default_rng(23)
    ht = np.asarray(st.variables.hIceMean) * (1 + 0.1 * rng.standard_normal(st.shape))
    At = np.asarray(st.variables.Area) * (1 + 0.1 * rng.standard_normal(st.shape))
    ht_g, At_g = torch.from_numpy(ht).to(cuda), torch.from_numpy(At).to(cuda)

    def gpu_loss(p):
        u, v, *_ = A.evp_subcycle(gst, n_substeps=n, ckpt_every=10, **p)
        h1, _hs1, A1 = A.advect(gst, uIce=u, vIce=v, hIceMean=p["hIceMean"])
        return (((h1 - ht_g)[2:-2, 2:-2] ** 2).sum() + ((A1 - At_g)[2:-2, 2:-2] ** 2).sum())

    p = _params(st, cuda)
    gpu_loss(p).backward()
    got = {k: g2n(v.grad) for k, v in p.items()}

    vs = {k: torch.from_numpy(np.asarray(v, dtype=np.float64)) for k, v in st.variables.items()}
    s = st.settings
    pc = _params(st, "cpu")
    f = vs["fCori"]
    ca = torch.where(f < 0, pc["airIceDrag_south"], pc["airIceDrag"])
    cw = torch.where(f < 0, pc["waterIceDrag_south"], pc["waterIceDrag"])
    pre = TR.pre_evp(vs, s, pc["pStar"], pc["hIceMean"], ca)
    u, v, *_ = TR.evp(vs, s, pre, n, cw)
    h1 = TR.advection(vs, s, u, v, pc["hIceMean"])
    A1 = TR.advection(vs, s, u, v, vs["Area"])
    J = (((h1 - torch.from_numpy(ht))[2:-2, 2:-2] ** 2).sum()
         + ((A1 - torch.from_numpy(At))[2:-2, 2:-2] ** 2).sum())
    J.backward()
    for k, t in pc.items():
        w = t.grad.numpy()
        err = float(np.linalg.norm(got[k] - w) / max(np.linalg.norm(w), 1e-300))
        print(f"thickness misfit d/d{k}: rel L2 err {err:.3e} |grad| {np.linalg.norm(w):.3e}",
              flush=True)
        assert np.linalg.norm(w) > 0 and err <= 1e-9, (k, err)

    base = _params(st, cuda, requires_grad=False)
    rng = np.random.default_rng(29)
    worst = 0.0
    for d in range(10):
        dirs = {k: torch.tensor(rng.standard_normal(tuple(v.shape)), device=cuda,
                                dtype=torch.float64) * v for k, v in base.items()}
        eps = 1e-6
        with torch.no_grad():
            jp = float(gpu_loss({k: base[k] + eps * dirs[k] for k in base}))
            jm = float(gpu_loss({k: base[k] - eps * dirs[k] for k in base}))
        fd = (jp - jm) / (2 * eps)
        ad = sum(float(np.sum(got[k] * g2n(dirs[k]))) for k in base)
        rel = abs(fd - ad) / max(abs(ad), 1e-300)
        worst = max(worst, rel)
        print(f"direction {d}: fd={fd:.10e} adjoint={ad:.10e} rel={rel:.2e}", flush=True)
    assert worst <= 1e-5
