"""GPU parity: each HIP drop-in against the NumPy FP64 oracle on the same
seeded inputs.

Bar: the FP64 kernels that use only +, -, *, /, sqrt and comparisons (mass,
area, wind forcing, the EVP subcycle, ocean stress, advection, clean-up,
ridging, fill) are compared BIT FOR BIT on every padded cell, halos included.
Kernels that call exp/exp10/log (strength, basal drag, growth) are compared with
the relative tolerance stated in each test (max|diff| / max|oracle|).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from gpu_helpers import assert_close, assert_exact, evp_shapes_from_source, g2n, pre_evp
from oracle import veris_np as O
from veris_amd import _native as N
from veris_amd import kernels as K
from veris_amd import synth
from veris_amd.step import seaice_step

pytestmark = pytest.mark.gpu

EVP_NAMES = ("uIce", "vIce", "sigma1", "sigma2", "sigma12")


def _evp_pair(st_np, n, dev, **kw):
    want = O.evp_solver(st_np, n_substeps=n, **kw)
    got = K.evp_solver(st_np.to(dev), n_substeps=n)
    return [g2n(t) for t in got], want


def _spun_up(config_state, dev, n_spin=25):
    """State with non-zero velocities and stresses (oracle EVP spin-up)."""
    st = pre_evp(config_state)
    vs = st.variables
    vs.uIce, vs.vIce, vs.sigma1, vs.sigma2, vs.sigma12 = O.evp_solver(st, n_substeps=n_spin)
    return pre_evp(st)


def test_pre_evp_kernels(cuda):
    st = _spun_up(synth.config0(), cuda)
    g = st.to(cuda)
    for name, got, want in zip(("SeaIceMassC", "SeaIceMassU", "SeaIceMassV"), K.SeaIceMass(g),
                               O.SeaIceMass(st)):
        assert_exact(name, g2n(got), want)
    for name, got, want in zip(("AreaW", "AreaS"), K.AreaWS(g), O.AreaWS(st)):
        assert_exact(name, g2n(got), want)
    for name, got, want in zip(("WindForcingX", "WindForcingY"), K.WindForcingXY(g),
                               O.WindForcingXY(st)):
        assert_exact(name, g2n(got), want)
    # exp() of the device math library: a few ulp
    assert_close("SeaIceStrength", g2n(K.SeaIceStrength(g)), O.SeaIceStrength(st), 1e-15)
    pstar = np.random.default_rng(5).lognormal(np.log(27500.0), 0.3, st.shape)
    assert_close("SeaIceStrength(P* field)",
                 g2n(K.SeaIceStrength(g, torch.from_numpy(pstar).to(cuda))),
                 O.SeaIceStrength(st, pStar=pstar), 1e-15)


def test_fused_pre_evp_equals_separate_kernels(cuda):
    """vh_pre_evp: every output equals the separate drop-in bit for bit (the
    same device exp in the strength), with and without a P* field."""
    st = _spun_up(synth.config0(), cuda)
    g = st.to(cuda)
    sep = (*K.SeaIceMass(g), *K.AreaWS(g), *K.WindForcingXY(g), K.SeaIceStrength(g))
    names = ("SeaIceMassC", "SeaIceMassU", "SeaIceMassV", "AreaW", "AreaS", "WindForcingX",
             "WindForcingY", "SeaIceStrength")
    for name, got, want in zip(names, K.pre_evp(g), sep):
        assert_exact(f"fused {name}", g2n(got), g2n(want))
    pstar = torch.from_numpy(np.random.default_rng(5).lognormal(np.log(27500.0), 0.3,
                                                                st.shape)).to(cuda)
    assert_exact("fused SeaIceStrength(P* field)", g2n(K.pre_evp(g, pstar)[-1]),
                 g2n(K.SeaIceStrength(g, pstar)))


def test_ocean_coupling(cuda):
    """vh_ocean_to_ice / vh_ice_to_ocean against the oracle, bit for bit."""
    from veris_amd import coupling

    st = _spun_up(synth.config0(), cuda)
    vs = st.variables
    rng = np.random.default_rng(9)
    shp = st.shape
    u, v = rng.normal(0, 0.1, shp), rng.normal(0, 0.1, shp)
    t, sal = rng.normal(5, 8, shp), rng.normal(34.5, 1, shp)
    ft = rng.normal(0, 2e-5, shp)
    vs.OceanStressU, vs.OceanStressV = O.OceanStressUV(st)
    g = st.to(cuda)
    dev = lambda a: torch.from_numpy(a).to(cuda)
    got = coupling.ocean_to_ice(g, dev(u), dev(v), dev(t), dev(sal), dev(ft))
    want = O.ocean_to_ice(st, u, v, t, sal, ft)
    for name, a, b in zip(("uOcean", "vOcean", "theta", "ocSalt", "Qnet", "Qsw", "wSpeed"),
                          got, want):
        assert_exact(name, g2n(a), b)
    tx, ty = rng.normal(0, 0.1, shp), rng.normal(0, 0.1, shp)
    fi, fr = rng.normal(0, 1e-6, shp), rng.normal(0, 1e-6, shp)
    got = coupling.ice_to_ocean(g, dev(tx), dev(ty), dev(fi), dev(fr))
    want = O.ice_to_ocean(st, tx, ty, fi, fr)
    for name, a, b in zip(("surface_taux", "surface_tauy", "forc_temp_surface",
                           "forc_salt_surface"), got, want):
        assert_exact(name, g2n(a), b)


@pytest.mark.parametrize("n", [1, 2, 3, 40])
def test_evp_substeps_bit_exact(cuda, n):
    st = _spun_up(synth.config0(), cuda)
    got, want = _evp_pair(st, n, cuda)
    for name, a, b in zip(EVP_NAMES, got, want):
        assert_exact(f"{name}[n={n}]", a, b)


def test_evp_full_subcycle_config0(cuda):
    """400 substeps from rest, the 4-degree workload."""
    st = pre_evp(synth.config0())
    got, want = _evp_pair(st, 400, cuda)
    for name, a, b in zip(EVP_NAMES, got, want):
        assert_exact(f"{name}[400]", a, b)


def test_evp_config1_1deg(cuda):
    st = _spun_up(synth.config1(), cuda, n_spin=3)
    got, want = _evp_pair(st, 12, cuda)
    for name, a, b in zip(EVP_NAMES, got, want):
        assert_exact(f"{name}[1deg]", a, b)


def test_evp_config4_quarter_degree(cuda):
    """The full 0.25-degree grid (1440 x 720 + halo, about a million cells,
    the largest single domain here): EVP substeps bit-exact on every cell
    with the default shape, and with the two-substep shapes the autotuner
    picks there (16 x 32 x 2 in FP64, 24 x 32 x 2 in FP32) on every velocity
    and every stress outside the periodic artefact columns."""
    st = _spun_up(synth.config4(), cuda, n_spin=3)
    got, want = _evp_pair(st, 4, cuda)
    for name, a, b in zip(EVP_NAMES, got, want):
        assert_exact(f"{name}[0.25deg]", a, b)
    g = st.to(cuda)
    assert K.overlaps_periodic(g)
    keep = np.setdiff1d(np.arange(st.shape[0]), _periodic_artefact_cols(st))
    shapes = N_shapes()
    for tile in (t for t, sh in shapes.items() if sh in ((16, 32, 2), (24, 32, 2))):
        got = [g2n(x) for x in K.evp_solver(g, n_substeps=4, tile=tile)]
        for q, (name, a, b) in enumerate(zip(EVP_NAMES, got, want)):
            if q >= 2:
                a, b = a[keep], b[keep]
            assert_exact(f"{name}[0.25deg, tile={tile}]", a, b)


def test_full_step_config4_quarter_degree(cuda):
    """One whole step on the full 0.25-degree grid (two EVP substeps to keep
    the oracle quick); tolerance 1e-9 of each field's maximum, as for the
    4-degree steps (device exp/exp10 in strength and growth)."""
    st = synth.config4()
    q0 = st.variables.Qnet.copy()
    sw = st.variables.SWdown.copy()
    g = st.to(cuda)
    O.seaice_step(st, q0, sw, n_substeps=2)
    seaice_step(g, torch.from_numpy(q0).to(cuda), torch.from_numpy(sw).to(cuda), n_substeps=2)
    for name in ("uIce", "vIce", "sigma1", "sigma2", "sigma12", "OceanStressU", "OceanStressV",
                 "hIceMean", "hSnowMean", "Area", "TSurf", "Qnet", "EmPmR", "SeaIceLoad"):
        assert_close(f"0.25deg step {name}", g2n(g.variables[name]), st.variables[name], 1e-9)


@pytest.mark.parametrize("variant", ["regional", "own_fill", "turning", "kterms", "tensile",
                                     "no_repl_pressure", "small_domain", "adaptive",
                                     "adaptive_regional", "side_drag", "side_drag_regional",
                                     "coastline", "side_drag_adaptive", "second_order",
                                     "second_order_regional", "second_order_kterms"])
def test_evp_setting_variants(cuda, variant):
    if variant == "regional":
        st = synth.regional(7, 0, 37, 29)
    elif variant == "second_order":
        st = synth.tiny(20, 14, secondOrderBC=True)
    elif variant == "second_order_regional":
        st = synth.regional(7, 2, 37, 29, secondOrderBC=True, useAdaptiveEVP=True)
    elif variant == "second_order_kterms":
        st = synth.tiny(20, 14, secondOrderBC=True, sinWat=0.42261826174069944,
                        cosWat=0.9063077870366499)
        r = np.random.default_rng(4)
        for k in ("k1AtC", "k2AtC", "k1AtZ", "k2AtZ"):
            st.variables[k] = r.normal(0, 1e-7, st.shape)
    elif variant == "side_drag":
        st = synth.tiny(20, 14, noSlip=False)
    elif variant == "side_drag_regional":
        st = synth.regional(7, 1, 37, 29, noSlip=False, sideDragCoeff=0.004, sideDragU0=0.02)
    elif variant == "coastline":
        st = synth.tiny(20, 14, noSlip=False, use_coastline=True)
        r = np.random.default_rng(8)
        st.variables.Fu = r.uniform(0.0, 2.0, st.shape) * st.variables.iceMaskU
        st.variables.Fv = r.uniform(0.0, 2.0, st.shape) * st.variables.iceMaskV
    elif variant == "side_drag_adaptive":
        st = synth.tiny(20, 14, noSlip=False, useAdaptiveEVP=True)
    elif variant == "adaptive":
        st = synth.tiny(20, 14, useAdaptiveEVP=True)
    elif variant == "adaptive_regional":
        st = synth.regional(7, 1, 37, 29, useAdaptiveEVP=True, aEvpCoeff=0.2, aEVPalphaMin=20.0)
    elif variant == "own_fill":
        st = synth.tiny(20, 14, veros_fill=False)
    elif variant == "turning":
        st = synth.tiny(20, 14, sinWat=0.42261826174069944, cosWat=0.9063077870366499)
    elif variant == "tensile":
        st = synth.tiny(20, 14, tensileStrFac=0.05)
    elif variant == "no_repl_pressure":
        st = synth.tiny(20, 14, pressReplFac=0)
    elif variant == "small_domain":
        st = synth.tiny(4, 4)
    else:
        st = synth.tiny(20, 14)
        r = np.random.default_rng(3)
        for k in ("k1AtC", "k2AtC", "k1AtZ", "k2AtZ"):
            st.variables[k] = r.normal(0, 1e-7, st.shape)
    st = _spun_up(st, cuda, n_spin=10)
    got, want = _evp_pair(st, 30, cuda)
    for name, a, b in zip(EVP_NAMES, got, want):
        assert_exact(f"{name}[{variant}]", a, b)


def _assert_bits(name, got, want):
    """Equal bit patterns (also tells +0 from -0)."""
    got = np.ascontiguousarray(got, dtype=np.float64)
    want = np.ascontiguousarray(want, dtype=np.float64)
    nbad = int(np.count_nonzero(got.view(np.int64) != want.view(np.int64)))
    assert nbad == 0, f"{name}: {nbad} cells differ in their bits from the oracle"


def _packed_case(case):
    st = synth.tiny(20, 14)
    vs = st.variables
    NX, NY = st.shape
    if case == "half_mask":  # a mask value other than 0 or 1
        vs.maskInC = vs.maskInC.copy()
        wet = np.argwhere(vs.maskInC[2:-2, 2:-2] == 1.0)[7] + 2
        vs.maskInC[tuple(wet)] = 0.5
    elif case == "negative_zero_mask":  # -0.0 must not become +0.0
        vs.iceMask = vs.iceMask.copy()
        vs.iceMask[vs.iceMask == 0.0] = -0.0
    elif case == "curvilinear":  # metrics that vary along i as well as j
        wig = 1.0 + 1e-3 * np.cos(np.arange(NX))[:, None]
        for k in ("dxU", "dyV", "recip_dxU", "recip_rAu"):
            vs[k] = vs[k] * wig
    elif case == "one_cell_metric":  # one cell off its row by one ulp
        vs.recip_rAv = vs.recip_rAv.copy()
        vs.recip_rAv[9, 6] = np.nextafter(vs.recip_rAv[9, 6], np.inf)
    return st


@pytest.mark.parametrize("case", ["half_mask", "negative_zero_mask", "curvilinear",
                                  "one_cell_metric", "packed", "min_workspace"])
def test_evp_packed_constants(cuda, case):
    """The subcycle reads its masks as bits and its metrics as rows along j
    only when every mask is exactly +0 or 1 and every metric equals its row
    value in every bit; any other grid, or a workspace of the minimal size
    (without room for the compact form), reads the full fields.  Every case
    equals the oracle bit for bit, halos included."""
    st = _packed_case(case if case not in ("packed", "min_workspace") else "none")
    st = _spun_up(st, cuda, n_spin=5)
    want = O.evp_solver(st, n_substeps=30)
    g = st.to(cuda)
    if case == "min_workspace":
        from veris_amd import _native as N

        c = K.Ctx(g)
        NX, NY = c.shape
        full = N.workspace_bytes("vh_evp_workspace_bytes", c.dom)
        align = lambda b: (b + 255) // 256 * 256  # noqa: E731
        minimal = full - align(NX * NY) - align(12 * NY * 8)
        ws = torch.empty(minimal, dtype=torch.uint8, device=cuda)
        out = c.empty(5)
        N.call("vh_evp_solver", *c.args(), 30, *[t.data_ptr() for t in out], ws.data_ptr(),
               minimal, c.stream)
        with pytest.raises(RuntimeError):
            N.call("vh_evp_solver", *c.args(), 30, *[t.data_ptr() for t in out],
                   ws.data_ptr(), minimal - 256, c.stream)
        got = [g2n(t) for t in out]
    else:
        got = [g2n(t) for t in K.evp_solver(g, n_substeps=30)]
    for name, a, b in zip(EVP_NAMES, got, want):
        _assert_bits(f"{name}[{case}]", a, b)


def test_concurrent_first_calls_capture_one_graph_at_a_time(cuda):
    """Two host threads making their first subcycle calls at once (each on its
    own stream and workspace, so each captures its own graph) get the same
    bits as a call made alone."""
    import threading

    st = _spun_up(synth.tiny(20, 14, config=31), cuda, n_spin=3)
    want = [g2n(t) for t in K.evp_solver(st.to(cuda), n_substeps=37)]
    results, errors = [None, None], []

    def worker(k):
        try:
            s_ = torch.cuda.Stream(cuda)
            with torch.cuda.stream(s_):
                g = st.to(cuda)
                out = K.evp_solver(g, n_substeps=41 + k)  # new substep counts: new graphs
                out = K.evp_solver(g, n_substeps=37)
            s_.synchronize()
            results[k] = [g2n(t) for t in out]
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for k in range(2):
        for name, a, b in zip(EVP_NAMES, results[k], want):
            assert_exact(f"{name}[thread {k}]", a, b)


@pytest.mark.parametrize("kind", ["global", "regional"])
def test_freedrift_solver(cuda, kind):
    """Free drift (closed-form c-point solution with atan2, cos, sin): the
    device's transcendental functions differ from the host's by an ulp, so
    the bar is 1e-13 of each field's scale; IceVelocities dispatches to it
    when only useFreedrift is set and keeps the stresses."""
    if kind == "global":
        st = synth.config0(useFreedrift=True, useEVP=False, waterIceDrag_south=0.0042)
    else:
        st = synth.regional(7, 3, 37, 29, useFreedrift=True, useEVP=False)
    st = _spun_up(st, cuda, n_spin=3)
    want = O.freedrift_solver(st)
    g = st.to(cuda)
    got = K.freedrift_solver(g)
    for name, a, b in zip(("uIce", "vIce"), got, want):
        assert_close(f"freedrift {name}", g2n(a), b, 1e-13)
    u, v, s1, s2, s12 = K.IceVelocities(g)
    assert_close("IceVelocities uIce", g2n(u), want[0], 1e-13)
    assert s1 is g.variables.sigma1 and s12 is g.variables.sigma12


@pytest.mark.parametrize("variant", ["default", "adaptive", "side_drag", "second_order"])
def test_evp_residuals(cuda, variant):
    """computeEvpResidual: per-substep stress and velocity residual norms
    against the oracle (FP64 sums over the interior in a different order:
    1e-12 relative per substep); the subcycle itself equals vh_evp_solver's
    bit for bit."""
    kw = {"default": {}, "adaptive": {"useAdaptiveEVP": True},
          "side_drag": {"noSlip": False}, "second_order": {"secondOrderBC": True}}[variant]
    st = _spun_up(synth.config0(computeEvpResidual=True, **kw), cuda, n_spin=5)
    g = st.to(cuda)
    n = 23
    want = O.evp_solver_residuals(st, n_substeps=n)
    got = K.evp_solver_residuals(g, n_substeps=n)
    plain = K.evp_solver(g, n_substeps=n)
    for q, name in enumerate(EVP_NAMES):
        assert_exact(f"residual run {name}", g2n(got[q]), want[q])
        assert_exact(f"residual run vs plain {name}", g2n(got[q]), g2n(plain[q]))
    for name, a, b in (("resU", got[5], want[5]), ("resSig", got[6], want[6])):
        a = g2n(a)
        assert a.shape == (n,) and np.all(b > 0), name
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=0, err_msg=name)


def test_bulk_fluxes(cuda):
    """bulkf_formula_lanl on seeded open-ocean inputs spanning stable, neutral
    and unstable stratification and calm to gale winds: log/exp/atan/pow of
    the device library, 1e-12 of each field's scale."""
    from veris_amd.heat_flux_MITgcm import bulkf_formula_lanl

    st = synth.config0()
    r = np.random.default_rng(77)
    shp = st.shape
    tsf = r.uniform(271.0, 303.0, shp)
    ins = {"uw": r.normal(0, 8, shp), "vw": r.normal(0, 8, shp),
           "ta": tsf + r.normal(0, 6, shp), "qa": r.uniform(1e-3, 2e-2, shp), "tsf": tsf,
           "ocn_mask": st.variables.iceMask.copy()}
    ins["uw"][:3, :3] = 0.0
    ins["vw"][:3, :3] = 0.0
    want = O.bulkf_formula_lanl(st, *ins.values(), grav=st.settings.gravity)
    g = st.to(cuda)
    got = bulkf_formula_lanl(g, *[torch.from_numpy(v).to(cuda) for v in ins.values()])
    from veris_amd._native import BULK_FLUX_NAMES

    for name, a, b in zip(BULK_FLUX_NAMES, got, want):
        assert_close(f"bulk {name}", g2n(a), b, 1e-12)


def test_forcing_and_free_drift_fp32(cuda):
    """FP32 storage of the free-drift solver and the bulk fluxes against the
    FP64 oracle on the same (FP32-representable) inputs: 2e-5 of each field's
    scale (FP32 rounding through the Monin-Obukhov iterations and atan2)."""
    from veris_amd._native import BULK_FLUX_NAMES
    from veris_amd.heat_flux_MITgcm import bulkf_formula_lanl

    st = _spun_up(synth.config0(useFreedrift=True, useEVP=False), cuda, n_spin=3)
    for k in st.variables.keys():
        st.variables[k] = np.asarray(st.variables[k], dtype=np.float32).astype(np.float64)
    want = O.freedrift_solver(st)
    g32 = st.to(cuda, dtype=torch.float32)
    got = K.freedrift_solver(g32)
    for name, a, b in zip(("uIce", "vIce"), got, want):
        assert a.dtype == torch.float32
        assert_close(f"freedrift fp32 {name}", g2n(a).astype(np.float64), b, 2e-5)
    r = np.random.default_rng(78)
    shp = st.shape
    tsf = r.uniform(271.0, 303.0, shp)
    ins = {"uw": r.normal(0, 8, shp), "vw": r.normal(0, 8, shp), "ta": tsf + r.normal(0, 6, shp),
           "qa": r.uniform(1e-3, 2e-2, shp), "tsf": tsf, "ocn_mask": st.variables.iceMask.copy()}
    ins = {k: v.astype(np.float32).astype(np.float64) for k, v in ins.items()}
    want = O.bulkf_formula_lanl(st, *ins.values(), grav=st.settings.gravity)
    got = bulkf_formula_lanl(g32, *[torch.from_numpy(v).float().to(cuda) for v in ins.values()])
    for name, a, b in zip(BULK_FLUX_NAMES, got, want):
        assert_close(f"bulk fp32 {name}", g2n(a).astype(np.float64), b, 2e-5)


def test_evp_basal_drag(cuda):
    """basalDragK2 != 0 switches on the exp/log soft-max: tolerance 1e-12."""
    st = synth.tiny(20, 14, basalDragK2=15.0)
    st.variables.R_low = st.variables.R_low * 0.002  # shallow: keels ground
    st = _spun_up(st, cuda, n_spin=10)
    got, want = _evp_pair(st, 30, cuda)
    for name, a, b in zip(EVP_NAMES, got, want):
        assert_close(f"{name}[basal]", a, b, 1e-12)


def test_evp_zero_substeps_and_land(cuda):
    st = _spun_up(synth.tiny(12, 10), cuda, n_spin=5)
    got, want = _evp_pair(st, 0, cuda)
    for name, a, b in zip(EVP_NAMES, got, want):
        assert_exact(f"{name}[n=0]", a, b)
    # all land: masks zero everywhere, denominators hit the ==0 guard
    vs = st.variables
    for k in ("iceMask", "iceMaskU", "iceMaskV", "maskInC", "maskInU", "maskInV"):
        vs[k] = np.zeros(st.shape)
    st = pre_evp(st)
    got, want = _evp_pair(st, 5, cuda)
    for name, a, b in zip(EVP_NAMES, got, want):
        assert_exact(f"{name}[land]", a, b)


def test_ocean_stress(cuda):
    for angle in (0.0, 25.0):
        st = _spun_up(synth.tiny(24, 16, waterTurnAngle=angle), cuda, n_spin=20)
        g = st.to(cuda)
        want = O.OceanStressUV(st)
        got = K.OceanStressUV(g)
        for name, a, b in zip(("OceanStressU", "OceanStressV"), got, want):
            if angle == 0.0:
                assert_exact(f"{name}[{angle}]", g2n(a), b)
            else:  # sin/cos of the host libm vs numpy: 1 ulp in the turning factors
                assert_close(f"{name}[{angle}]", g2n(a), b, 1e-14)


def _after_evp(st):
    vs = st.variables
    vs.uIce, vs.vIce, vs.sigma1, vs.sigma2, vs.sigma12 = O.evp_solver(st, n_substeps=60)
    return st


def test_advection(cuda):
    for st in (synth.config0(), synth.regional(8, 1, 40, 33), synth.tiny(16, 12, veros_fill=False)):
        st = _after_evp(pre_evp(st))
        want = O.Advection(st)
        got = K.Advection(st.to(cuda))
        for name, a, b in zip(("hIceMean", "hSnowMean", "Area"), got, want):
            assert_exact(f"adv {name}", g2n(a), b)


def test_clean_up_and_ridging(cuda):
    st = synth.config0()
    vs = st.variables
    r = np.random.default_rng(11)
    vs.hIceMean = vs.hIceMean + r.normal(0, 0.05, st.shape)
    vs.hSnowMean = vs.hSnowMean + r.normal(0, 0.05, st.shape)
    vs.Area = vs.Area * 1.3 - 0.1
    vs.hIceMean[5:9, 5:9] = 5e-6
    g = st.to(cuda)
    names = ("hIceMean", "hSnowMean", "Area", "TSurf", "os_hIceMean", "os_hSnowMean")
    for name, a, b in zip(names, K.clean_up_advection(g), O.clean_up_advection(st)):
        assert_exact(f"clean_up {name}", g2n(a), b)
    assert_exact("ridging", g2n(K.ridging(g)), O.ridging(st))


def _growth_inputs(**settings):
    st = synth.config0(**settings)
    vs = st.variables
    r = np.random.default_rng(2)
    vs.TSurf = 250.0 + 25.0 * r.random(st.shape)
    vs.snowfall = np.abs(r.normal(0, 1e-8, st.shape))
    vs.precip = np.abs(r.normal(0, 1e-8, st.shape))
    vs.evap = r.normal(0, 1e-8, st.shape)
    vs.os_hIceMean = np.abs(r.normal(0, 1e-3, st.shape))
    return st


def test_growth(cuda):
    """exp/exp10 of the device math library: tolerance 1e-12 relative to the
    largest value of each output."""
    st = _growth_inputs()
    g = st.to(cuda)
    want = O.Growth(st)
    got = K.Growth(g)
    names = ("hIceMean", "hSnowMean", "Area", "TSurf", "EmPmR", "forc_salt_surface", "Qsw",
             "Qnet", "SeaIceLoad", "IcePenetSW", "recip_hIceMean")
    for name, a, b in zip(names, got, want):
        assert_close(f"growth {name}", g2n(a), b, 1e-12)


@pytest.mark.parametrize("nitc", [1, 12])
def test_growth_category_counts(cuda, nitc):
    """Other numbers of thickness categories against the oracle."""
    st = _growth_inputs(nITC=nitc)
    want = O.Growth(st)
    got = K.Growth(st.to(cuda))
    names = ("hIceMean", "hSnowMean", "Area", "TSurf", "EmPmR", "forc_salt_surface", "Qsw",
             "Qnet", "SeaIceLoad", "IcePenetSW", "recip_hIceMean")
    for name, a, b in zip(names, got, want):
        assert_close(f"growth nITC={nitc} {name}", g2n(a), b, 1e-12)


def test_fill_overlaps_equals_single_fills(cuda):
    """One launch filling several fields in place equals fill_overlap of
    each; more than eight fields or a repeated field is refused."""
    for st in (synth.config0(), synth.config0(veros_fill=False)):
        g = st.to(cuda)
        r = np.random.default_rng(5)
        fields = [torch.tensor(r.normal(size=st.shape), device=cuda) for _ in range(5)]
        want = [K.fill_overlap(g, a) for a in fields]
        K.fill_overlaps_(g, fields)
        for a, b in zip(fields, want):
            assert torch.equal(a, b)
    extra = [torch.zeros(st.shape, dtype=torch.float64, device=cuda) for _ in range(9)]
    with pytest.raises(N.VerisHipError, match="at most 8"):
        K.fill_overlaps_(g, extra)
    with pytest.raises(N.VerisHipError, match="repeats"):
        K.fill_overlaps_(g, [extra[0], extra[1], extra[0]])


def test_fill_and_global_sum(cuda):
    st = synth.config0()
    a = np.random.default_rng(4).normal(size=st.shape)
    g = st.to(cuda)
    for fill_mode in (True, False):
        st.settings.veros_fill = fill_mode
        g.settings.veros_fill = fill_mode
        assert_exact(f"fill[{fill_mode}]", g2n(K.fill_overlap(g, torch.from_numpy(a).to(cuda))),
                     O.fill_overlap(st, a))
    w = st.variables.rA
    s = K.global_sum(g, torch.from_numpy(a).to(cuda), g.variables.rA).item()
    ref = O.global_sum(a[2:-2, 2:-2] * w[2:-2, 2:-2])
    assert abs(s - ref) <= 1e-13 * O.global_sum(np.abs(a[2:-2, 2:-2] * w[2:-2, 2:-2]))


STEP_SOLVERS = {"evp": {}, "freedrift": {"useFreedrift": True, "useEVP": False},
                "adaptive_side_drag": {"useAdaptiveEVP": True, "noSlip": False},
                "second_order": {"secondOrderBC": True}}


@pytest.mark.parametrize("n_steps,solver", [(1, "evp"), (2, "evp"), (2, "freedrift"),
                                            (2, "adaptive_side_drag"), (2, "second_order")])
def test_full_step_config0(cuda, n_steps, solver):
    """Whole step (pre-EVP, 400 EVP substeps or free drift, ocean stress,
    advection, clean-up, ridging, growth, fills), default and alternative
    dynamics settings.  exp/pow in strength and growth make the GPU differ
    from the oracle by ulps that the next step carries on; tolerance 1e-9
    relative to each field's maximum."""
    kw = STEP_SOLVERS[solver]
    st = synth.config0(**kw)
    q0 = st.variables.Qnet.copy()
    sw = st.variables.SWdown.copy()
    g = st.to(cuda)
    qg = torch.from_numpy(q0).to(cuda)
    swg = torch.from_numpy(sw).to(cuda)
    for _ in range(n_steps):
        O.seaice_step(st, q0, sw)
        seaice_step(g, qg, swg)
    for name in ("uIce", "vIce", "sigma1", "sigma2", "sigma12", "OceanStressU", "OceanStressV",
                 "hIceMean", "hSnowMean", "Area", "TSurf", "Qnet", "Qsw", "EmPmR",
                 "forc_salt_surface_ice", "SeaIceLoad", "IcePenetSW", "recip_hIceMean"):
        assert_close(f"step{n_steps} {solver} {name}", g2n(g.variables[name]), st.variables[name],
                     1e-9)


def test_errors_are_reported(cuda):
    from veris_amd._native import VerisHipError

    st = pre_evp(synth.tiny(12, 10)).to(cuda)
    with pytest.raises(VerisHipError, match="VH_EUNSUPPORTED.*secondOrderBC"):
        st.settings.secondOrderBC = True
        K.evp_solver(st, 3, tile=8)
    st.settings.secondOrderBC = False
    st.variables.uIce = st.variables.uIce[:, :-1].contiguous()
    with pytest.raises(ValueError):
        K.evp_solver(st, 3)


ALL_TILES = sorted(evp_shapes_from_source())  # every launch shape of the library


def _periodic_artefact_cols(st):
    """x-overlap columns whose reference stresses are roll wrap-around
    artefacts on an x-periodic grid (nothing else depends on them)."""
    if not (st.settings.enable_cyclic_x or not st.settings.veros_fill):
        return []
    NX = st.shape[0]
    return [0, 1, NX - 2, NX - 1]


BRANCHES = {"adaptive": {"useAdaptiveEVP": True}, "side_drag": {"noSlip": False},
            "second_order": {"secondOrderBC": True}}


@pytest.mark.parametrize("tile", [1, 5, 8, 9, 13])
@pytest.mark.parametrize("branch", sorted(BRANCHES))
def test_branch_tile_shapes(cuda, branch, tile):
    """Adaptive EVP (per-cell alpha, beta), free slip with lateral drag and
    the second-order no-slip boundary term on single- and multi-substep
    shapes: bit-exact against the oracle (stresses outside the periodic
    artefact columns for k > 1 on the x-cyclic grid).  The second-order term
    reads two cells beyond the face, so multi-substep shapes refuse it."""
    from veris_amd._native import VerisHipError, evp_tile_shapes

    k = evp_tile_shapes()[tile][2]
    st = _spun_up(synth.config0(**BRANCHES[branch]), cuda, n_spin=5)
    g = st.to(cuda)
    if branch == "second_order" and k > 1:
        with pytest.raises(VerisHipError, match="VH_EUNSUPPORTED.*secondOrderBC"):
            K.evp_solver(g, n_substeps=3, tile=tile)
        return
    skip = _periodic_artefact_cols(st) if k > 1 else []
    keep = np.setdiff1d(np.arange(st.shape[0]), skip)
    want = O.evp_solver(st, n_substeps=21)
    got = K.evp_solver(g, n_substeps=21, tile=tile)
    for q, (name, a, b) in enumerate(zip(EVP_NAMES, got, want)):
        a = g2n(a)
        if q >= 2:
            a, b = a[keep], b[keep]
        assert_exact(f"{branch} {name}[tile={tile}]", a, b)


@pytest.mark.parametrize("tile", ALL_TILES)
def test_evp_every_tile_shape(cuda, tile):
    """Every workgroup shape against the oracle.  Single-substep shapes (k=1)
    are bit-exact on every cell; multi-substep shapes (k>1) are bit-exact on
    every velocity and on every stress outside the x-overlap columns whose
    reference values are wrap-around artefacts; on a closed regional domain
    (no periodic fill) they are bit-exact everywhere."""
    from veris_amd._native import evp_tile_shapes

    k = evp_tile_shapes()[tile][2]
    for st in (_spun_up(synth.config0(), cuda, n_spin=5),
               _spun_up(synth.regional(9, 0, 45, 61), cuda, n_spin=5)):
        g = st.to(cuda)
        skip = _periodic_artefact_cols(st) if k > 1 else []
        keep = np.setdiff1d(np.arange(st.shape[0]), skip)
        for n in (3, 17):
            want = O.evp_solver(st, n_substeps=n)
            got = K.evp_solver(g, n_substeps=n, tile=tile)
            for q, (name, a, b) in enumerate(zip(EVP_NAMES, got, want)):
                a = g2n(a)
                if q >= 2:
                    a, b = a[keep], b[keep]
                assert_exact(f"{name}[tile={tile}, k={k}, n={n}]", a, b)


def test_overlaps_periodic_guards_multi_substep_shapes(cuda):
    """vh_evp_overlaps_periodic: 1 on the x-cyclic 4-degree state (its
    inputs are filled), unchanged by a wrap-around value in an overlap stress
    that nothing interior reads (sigma1 in column 0), 0 for a value the
    interior reads (sigma1 in column 1, sigma12 in column NX-2, an ocean
    current anywhere in the overlap); always 1 without a filled overlap.
    Where it gives 0, a multi-substep shape indeed departs from the oracle."""
    from veris_amd.autotune import candidate_tiles

    st = _spun_up(synth.config0(), cuda, n_spin=5)
    g = st.to(cuda)
    assert K.overlaps_periodic(g)
    NX = st.shape[0]
    for name, col, want in (("sigma1", 0, True), ("sigma1", 1, False), ("sigma12", NX - 2, False),
                            ("sigma12", 1, True), ("uOcean", NX - 1, False), ("uIce", 0, False)):
        keep = g.variables[name]
        t = keep.clone()
        t[col, 5] += 0.125
        g.variables[name] = t
        assert K.overlaps_periodic(g) == want, (name, col)
        g.variables[name] = keep
    assert K.overlaps_periodic(g)
    # a current that is not periodic in the overlap: k > 1 shapes would read the
    # interior partner where the reference reads the overlap value
    uoc = g.variables.uOcean.clone()
    uoc[1, :] += 0.01
    g.variables.uOcean = uoc
    st.variables.uOcean = g2n(uoc)
    assert not K.overlaps_periodic(g)
    assert all(k == 1 for t, (_, _, k) in N_shapes().items()
               if t in candidate_tiles(False, st.settings, K.overlaps_periodic(g)))
    want = O.evp_solver(st, n_substeps=4)
    got = K.evp_solver(g, n_substeps=4, tile=13)
    assert np.count_nonzero(g2n(got[0]) != want[0]) > 0
    # without a filled overlap the check always passes
    r = _spun_up(synth.regional(9, 0, 45, 61), cuda, n_spin=3)
    assert K.overlaps_periodic(r.to(cuda))


def N_shapes():
    from veris_amd._native import evp_tile_shapes

    return evp_tile_shapes()


@pytest.mark.parametrize("tile", ALL_TILES)
def test_plain_instantiation_equals_general(cuda, tile, monkeypatch):
    """Under the default settings the subcycle launches the kernel
    instantiation with every optional branch compiled out; the general
    instantiation (VERIS_EVP_PLAIN=0) gives the same bits on every cell, for
    every launch shape, and both equal the oracle (velocities everywhere,
    stresses outside the periodic artefact columns for k > 1)."""
    from veris_amd._native import evp_tile_shapes

    k = evp_tile_shapes()[tile][2]
    st = _spun_up(synth.config0(), cuda, n_spin=5)
    g = st.to(cuda)
    monkeypatch.setenv("VERIS_EVP_PLAIN", "1")
    plain = [g2n(t) for t in K.evp_solver(g, n_substeps=17, tile=tile)]
    monkeypatch.setenv("VERIS_EVP_PLAIN", "0")
    general = [g2n(t) for t in K.evp_solver(g, n_substeps=17, tile=tile)]
    want = O.evp_solver(st, n_substeps=17)
    keep = np.setdiff1d(np.arange(st.shape[0]), _periodic_artefact_cols(st) if k > 1 else [])
    for q, (name, a, b, c) in enumerate(zip(EVP_NAMES, plain, general, want)):
        assert_exact(f"plain vs general {name}[tile={tile}]", a, b)
        if q >= 2:
            a, c = a[keep], c[keep]
        assert_exact(f"plain {name}[tile={tile}]", a, c)


def test_argument_structs_follow_the_state(cuda):
    """The drop-ins reuse their ABI structs between calls; every change a
    caller can make must still reach the kernels: a reassigned variable (a
    new tensor, possibly at a freed tensor's address), a changed setting, and
    an in-place change of the metric-term fields (which decides whether they
    are read at all).  Each call equals the oracle bit for bit."""
    st = _spun_up(synth.tiny(20, 14), cuda, n_spin=5)
    g = st.to(cuda)
    for name, a, b in zip(EVP_NAMES, K.evp_solver(g, n_substeps=9), O.evp_solver(st, n_substeps=9)):
        assert_exact(f"first call {name}", g2n(a), b)
    # reassign the ice velocities (new tensors) and the ocean currents
    r = np.random.default_rng(41)
    for k in ("uIce", "vIce", "uOcean"):
        st.variables[k] = st.variables[k] + r.normal(0, 0.01, st.shape) * st.variables.iceMaskU
        g.variables[k] = torch.from_numpy(st.variables[k]).to(cuda)
    for name, a, b in zip(EVP_NAMES, K.evp_solver(g, n_substeps=9), O.evp_solver(st, n_substeps=9)):
        assert_exact(f"after reassignment {name}", g2n(a), b)
    # a setting
    st.settings.evpAlpha = g.settings.evpAlpha = 450.0
    for name, a, b in zip(EVP_NAMES, K.evp_solver(g, n_substeps=9), O.evp_solver(st, n_substeps=9)):
        assert_exact(f"after a setting {name}", g2n(a), b)
    # the metric terms, zero so far, changed in place
    k = r.normal(0, 1e-7, st.shape)
    for name in ("k1AtC", "k2AtC", "k1AtZ", "k2AtZ"):
        st.variables[name] = k.copy()
        g.variables[name].copy_(torch.from_numpy(k))
    for name, a, b in zip(EVP_NAMES, K.evp_solver(g, n_substeps=9), O.evp_solver(st, n_substeps=9)):
        assert_exact(f"after metric terms {name}", g2n(a), b)
