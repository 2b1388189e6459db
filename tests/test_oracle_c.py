"""The NumPy oracle against the scalar-loop C restatement (``oracle/c/``).

Two restatements of ``veris/evp_solver.py:22-319`` and
``veris/freedrift_solver.py:5-62`` written in different idioms -- whole-array
``numpy.roll`` shifts in the oracle, explicit wrapped (i, j) indices cell by
cell in C -- must give the same bits on every padded cell, halos included,
for every solver branch.  A mistake in a shift direction, a metric, a mask or
an association order in either one breaks the equality.  Where NumPy's
transcendental functions may come from a different implementation than the C
library's (exp/log of basal drag, atan2/cos/sin of free drift), the bar is a
relative 1e-12 of each field's scale instead.

The second test plants each EVP and free-drift mistake of
``test_oracle_mutations`` into a copy of the oracle and requires the C
comparison alone to see it.

The whole step (``oracle/c/step_ref.c``: pre-EVP, IceVelocities, ocean
stress, advection, clean-up, ridging, Growth with solve4temp, fills) is
compared the same way.  The sea-ice strength calls exp(), so the C step is
handed the oracle's strength field; then every field computed before Growth
must be equal bit for bit, and Growth's outputs, whose solve4temp and melt
taper call pow() and exp(), must be bit-equal in at least 99% of the cells
and within 1e-12 of scale everywhere.  Every planted mistake of the step
functions must break that.

The residual norms of ``computeEvpResidual`` are compared within 1e-12 per
substep (FP64 sums in another order), the subcycle that carries them bit for
bit, and every planted residual mistake must break that.

The open-ocean bulk fluxes (the forcing producer of the step), all of whose
outputs go through exp/log/atan/pow, are compared within 1e-13 of scale
with at least 90% of the cells bit-equal, and every planted bulk-flux
mistake must break that too.
"""

from __future__ import annotations

import importlib.util
import inspect

import numpy as np
import pytest

from gpu_helpers import pre_evp
from oracle import c_ref
from oracle import veris_np as O
from test_oracle_mutations import MUTATIONS, ORACLE
from veris_amd import synth

N_SUB = 12


def _shallow(st):
    """Water 8 m deep, given as a negative depth (the drag takes |R_low|):
    keels of the 0-3 m ice ground (hCrit = 8 A / K1 < 1 m)."""
    st.variables.R_low = -st.variables.R_low * 0.002
    return st


def _kterms(st, seed):
    r = np.random.default_rng(seed)
    for k in ("k1AtC", "k2AtC", "k1AtZ", "k2AtZ"):
        st.variables[k] = r.normal(0, 1e-7, st.shape)
    return st


def _coast(st):
    r = np.random.default_rng(8)
    st.variables.Fu = r.uniform(0.0, 2.0, st.shape) * st.variables.iceMaskU
    st.variables.Fv = r.uniform(0.0, 2.0, st.shape) * st.variables.iceMaskV
    return st


def _ice_edge(st):
    """Open water in scattered cells: ice edges in both directions, so some
    faces have AreaW > 0 = AreaS and others AreaS > 0 = AreaW."""
    water = np.random.default_rng(11).random(st.shape) < 0.35
    for k in ("Area", "hIceMean", "hSnowMean"):
        st.variables[k] = np.where(water, 0.0, st.variables[k])
    return st


TURN = {"sinWat": 0.42261826174069944, "cosWat": 0.9063077870366499}
VARIANTS = {
    "default": lambda: synth.tiny(20, 14),
    "regional": lambda: synth.regional(7, 0, 37, 29),
    "own_fill": lambda: synth.tiny(20, 14, veros_fill=False),
    "turning": lambda: synth.tiny(20, 14, **TURN),
    "kterms": lambda: _kterms(synth.tiny(20, 14), 3),
    "tensile": lambda: synth.tiny(20, 14, tensileStrFac=0.05),
    "no_repl_pressure": lambda: synth.tiny(20, 14, pressReplFac=0),
    "adaptive": lambda: synth.tiny(20, 14, useAdaptiveEVP=True),
    "adaptive_regional": lambda: synth.regional(7, 1, 37, 29, useAdaptiveEVP=True, aEvpCoeff=0.2,
                                                aEVPalphaMin=20.0),
    "side_drag": lambda: synth.tiny(20, 14, noSlip=False),
    "side_drag_adaptive": lambda: synth.regional(7, 1, 37, 29, noSlip=False, useAdaptiveEVP=True,
                                                 sideDragCoeff=0.004, sideDragU0=0.02),
    "side_drag_ice_edge": lambda: _ice_edge(synth.tiny(20, 14, noSlip=False)),
    "coastline": lambda: _coast(synth.tiny(20, 14, noSlip=False, use_coastline=True)),
    "second_order": lambda: synth.tiny(20, 14, secondOrderBC=True),
    "second_order_regional": lambda: synth.regional(7, 2, 37, 29, secondOrderBC=True,
                                                    useAdaptiveEVP=True),
    "second_order_kterms": lambda: _kterms(synth.tiny(20, 14, secondOrderBC=True, **TURN), 4),
    "basal": lambda: _shallow(synth.tiny(20, 14, basalDragK2=15.0)),
}
INEXACT = {"basal"}
_STATES = {}


def _state(name):
    """Spun-up state of a variant (non-zero velocities and stresses)."""
    if name not in _STATES:
        st = pre_evp(VARIANTS[name]())
        vs = st.variables
        vs.uIce, vs.vIce, vs.sigma1, vs.sigma2, vs.sigma12 = O.evp_solver(st, n_substeps=10)
        _STATES[name] = pre_evp(st)
    return _STATES[name]


def _freedrift_state(kind):
    key = "freedrift_" + kind
    if key not in _STATES:
        if kind == "global":
            st = synth.config0(useFreedrift=True, useEVP=False, waterIceDrag_south=0.0042)
        else:
            st = synth.regional(7, 3, 37, 29, useFreedrift=True, useEVP=False)
        st = pre_evp(st)
        vs = st.variables
        vs.uIce, vs.vIce, vs.sigma1, vs.sigma2, vs.sigma12 = O.evp_solver(st, n_substeps=3)
        _STATES[key] = pre_evp(st)
    return _STATES[key]


def _rel(a, b):
    return float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-300)


def _differs(oracle, name):
    st = _state(name)
    want = c_ref.evp_solver(st, N_SUB)
    got = oracle.evp_solver(st, n_substeps=N_SUB)
    if name in INEXACT:
        return any(not _rel(a, b) <= 1e-12 for a, b in zip(got, want))
    return any(not np.array_equal(a, b) for a, b in zip(got, want))


@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_evp_matches_c_restatement(variant):
    st = _state(variant)
    want = c_ref.evp_solver(st, N_SUB)
    got = O.evp_solver(st, n_substeps=N_SUB)
    for name, a, b in zip(("uIce", "vIce", "sigma1", "sigma2", "sigma12"), got, want):
        assert np.all(np.isfinite(b)), name
        assert float(np.max(np.abs(b))) > 0, f"{name}[{variant}] is all zero"
        if variant in INEXACT:
            assert _rel(a, b) <= 1e-12, f"{name}[{variant}] rel {_rel(a, b):.3e}"
        else:
            nbad = int(np.count_nonzero(a != b))
            assert nbad == 0, f"{name}[{variant}]: {nbad} cells differ, rel {_rel(a, b):.3e}"


@pytest.mark.parametrize("kind", ["global", "regional"])
def test_freedrift_matches_c_restatement(kind):
    st = _freedrift_state(kind)
    want = c_ref.freedrift_solver(st)
    got = O.freedrift_solver(st)
    for name, a, b in zip(("uIce", "vIce"), got, want):
        assert float(np.max(np.abs(b))) > 0
        assert _rel(a, b) <= 1e-12, f"freedrift {name}[{kind}] rel {_rel(a, b):.3e}"


_EVP_FUNCS = (O.strainrates, O.side_drag, O.viscosities, O.c_point_to_z_point, O.stressdiv,
              O.evp_cfac, O.evp_substep, O.ocean_drag_coeffs, O.basal_drag_coeffs,
              O.fill_overlap)
_EVP_SRC = [inspect.getsource(f) for f in _EVP_FUNCS]
_FD_SRC = inspect.getsource(O.freedrift_solver)
# planted mistakes that the closed-form pins catch but that no C-comparison
# variant can show: no variant cell sits exactly at Area = 0.01 or has zero
# ice mass at an ice-covered u-point
PIN_ONLY = {"basal area gate", "turning mass gate"}
EVP_MUTATIONS = sorted(n for n, (a, _) in MUTATIONS.items()
                       if any(a in s for s in _EVP_SRC) and n not in PIN_ONLY)
FD_MUTATIONS = sorted(n for n, (a, _) in MUTATIONS.items() if a in _FD_SRC)


def _mutant(name, tmp_path):
    a, b = MUTATIONS[name]
    src = ORACLE.read_text()
    assert src.count(a) == 1, name
    path = tmp_path / "mutant.py"
    path.write_text(src.replace(a, b))
    spec = importlib.util.spec_from_file_location("veris_np_mutant", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_mutation_lists_are_not_empty():
    assert len(EVP_MUTATIONS) >= 25 and len(FD_MUTATIONS) >= 5


@pytest.mark.parametrize("mutation", EVP_MUTATIONS)
def test_c_comparison_catches_evp_mistake(mutation, tmp_path):
    mod = _mutant(mutation, tmp_path)
    assert any(_differs(mod, v) for v in VARIANTS), f"C comparison missed: {mutation}"


@pytest.mark.parametrize("mutation", FD_MUTATIONS)
def test_c_comparison_catches_free_drift_mistake(mutation, tmp_path):
    mod = _mutant(mutation, tmp_path)
    caught = False
    for kind in ("global", "regional"):
        st = _freedrift_state(kind)
        want = c_ref.freedrift_solver(st)
        got = mod.freedrift_solver(st)
        caught = caught or any(not _rel(a, b) <= 1e-12 for a, b in zip(got, want))
    assert caught, f"C comparison missed: {mutation}"


# ---------------------------------------------------------------------------
# the whole step
# ---------------------------------------------------------------------------

def _load_state(st):
    st = synth.tiny(20, 14) if st is None else st
    vs = st.variables
    vs.SeaIceLoad = st.settings.rhoIce * vs.hIceMean + st.settings.rhoSnow * vs.hSnowMean
    return st


def _moving_ice(st):
    """Non-zero initial ice velocities (relative wind, ocean stress)."""
    r = np.random.default_rng(12)
    st.variables.uIce = r.normal(0, 0.1, st.shape) * st.variables.iceMaskU
    st.variables.vIce = r.normal(0, 0.1, st.shape) * st.variables.iceMaskV
    return st


def _melting(st):
    """Warm air over thin snow: the snow melts away within the step."""
    vs = st.variables
    vs.ATemp = np.full(st.shape, 285.0)
    vs.hSnowMean = vs.hSnowMean * 0.002
    return st


def _at_melting_point(st):
    """Every surface starts the step at the melting point: wet albedos,
    with southern values that differ from the northern ones."""
    st.variables.TSurf = np.full(st.shape, st.settings.celsius2K)
    return st


def _thin_ice(st):
    """Ice exactly at, just below and just above hIce_min; with no motion
    (no substeps from rest) advection leaves these thicknesses unchanged."""
    vs, hmin = st.variables, st.settings.hIce_min
    h = vs.hIceMean.copy()
    h[2:-2:6, 2:-2] = hmin
    h[3:-2:6, 2:-2] = np.nextafter(hmin, 0)
    h[4:-2:6, 2:-2] = np.nextafter(hmin, 1)
    vs.hIceMean = h
    return st


STEP_VARIANTS = {
    "tiny": lambda: synth.tiny(20, 14),
    "moving_ice": lambda: _moving_ice(synth.tiny(20, 14)),
    "melting": lambda: _melting(synth.tiny(20, 14)),
    "thin_ice": lambda: _thin_ice(synth.tiny(20, 14)),
    "regional": lambda: synth.regional(7, 0, 37, 29),
    "config0": lambda: synth.config0(),
    "own_fill": lambda: synth.tiny(20, 14, veros_fill=False),
    "absolute_wind": lambda: synth.tiny(20, 14, useRelativeWind=False),
    "sea_ice_load": lambda: _load_state(synth.tiny(20, 14, useRealFreshWaterFlux=True)),
    "adaptive_side_drag": lambda: synth.tiny(20, 14, useAdaptiveEVP=True, noSlip=False),
    "ice_edge": lambda: _ice_edge(synth.tiny(20, 14)),
    # both hemispheres, turning angles, and southern dry albedos of their own
    "turning_angles": lambda: _moving_ice(synth.global_latlon(
        98, 20, 14, -70.0, 70.0, airTurnAngle=25.0, waterTurnAngle=20.0,
        dryIceAlb_south=0.72, drySnowAlb_south=0.8)),
    "wet_albedo": lambda: _at_melting_point(synth.global_latlon(
        97, 20, 14, -70.0, 70.0, wetIceAlb_south=0.6, wetSnowAlb_south=0.64,
        dryIceAlb_south=0.72, drySnowAlb_south=0.8)),
}
GROWTH_OUT = ("hIceMean", "hSnowMean", "Area", "TSurf", "EmPmR", "forc_salt_surface_ice", "Qsw",
              "Qnet", "SeaIceLoad", "IcePenetSW", "recip_hIceMean")


STEP_SUBSTEPS = {"thin_ice": 0}


def _step_pair(oracle, variant):
    """(C step, oracle step) from two fresh copies of the same seeded state."""
    mk, n_sub = STEP_VARIANTS[variant], STEP_SUBSTEPS.get(variant, 15)
    st, want = mk(), mk()
    q0 = st.variables.Qnet.copy()
    sw = st.variables.SWdown.copy()
    strength = O.SeaIceStrength(pre_evp(mk()))
    got = c_ref.seaice_step(st, q0, sw, n_substeps=n_sub, strength=strength)
    oracle.seaice_step(want, q0, sw, n_substeps=n_sub)
    return got, want.variables


def _step_mismatches(got, want):
    """Names of the fields that break the bar stated in the module docstring."""
    bad = []
    for k, a in got.items():
        if k not in want.keys() or not isinstance(want[k], np.ndarray):
            continue
        b = np.ascontiguousarray(np.broadcast_to(np.asarray(want[k], dtype=np.float64), a.shape))
        neq = int(np.count_nonzero(a.view(np.int64) != b.view(np.int64)))
        if not neq:
            continue
        rel = float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-300)
        if k not in GROWTH_OUT or neq > 0.01 * a.size or not rel <= 1e-12:
            bad.append((k, neq, rel))
    return bad


@pytest.mark.parametrize("variant", sorted(STEP_VARIANTS))
def test_step_matches_c_restatement(variant):
    got, want = _step_pair(O, variant)
    moving = STEP_SUBSTEPS.get(variant, 15) > 0
    assert float(np.max(np.abs(got["uIce"]))) > 0 or not moving
    assert float(np.max(got["hIceMean"])) > 0
    assert not _step_mismatches(got, want), _step_mismatches(got, want)


def test_step_strength_matches_c_restatement():
    """The C library's exp() in the strength: within 1e-15 of scale."""
    mk = STEP_VARIANTS["config0"]
    st = mk()
    got = c_ref.seaice_step(st, st.variables.Qnet.copy(), st.variables.SWdown.copy(),
                            n_substeps=1)
    want = O.SeaIceStrength(pre_evp(mk()))
    assert _rel(got["SeaIceStrength"], want) <= 1e-15


def test_free_drift_step_matches_c_restatement():
    """Free drift calls atan2/cos/sin: every field within 1e-12 of scale."""
    STEP_VARIANTS["free_drift"] = lambda: synth.config0(useFreedrift=True, useEVP=False)
    try:
        got, want = _step_pair(O, "free_drift")
    finally:
        del STEP_VARIANTS["free_drift"]
    for k, a in got.items():
        if k in want.keys() and isinstance(want[k], np.ndarray):
            b = np.broadcast_to(np.asarray(want[k], dtype=np.float64), a.shape)
            assert _rel(a, b) <= 1e-12, k


_STEP_FUNCS = (O.SeaIceMass, O.AreaWS, O.SeaIceStrength, O.tauXY, O.WindForcingXY,
               O.OceanStressUV, O.limiter, O._slope_ratio, O.calc_ZonalFlux,
               O.calc_MeridionalFlux, O.calc_Advection, O.clean_up_advection, O.ridging,
               O.solve4temp, O.Growth, O.seaice_step)
_STEP_SRC = [inspect.getsource(f) for f in _STEP_FUNCS]
STEP_MUTATIONS = sorted(n for n, (a, _) in MUTATIONS.items() if any(a in s for s in _STEP_SRC))


def test_step_mutation_list_is_not_empty():
    assert len(STEP_MUTATIONS) >= 20


@pytest.mark.parametrize("mutation", STEP_MUTATIONS)
def test_c_comparison_catches_step_mistake(mutation, tmp_path):
    mod = _mutant(mutation, tmp_path)
    for variant in ("tiny", "moving_ice", "melting", "thin_ice", "config0", "sea_ice_load",
                    "ice_edge", "turning_angles", "wet_albedo"):
        try:
            got, want = _step_pair(mod, variant)
        except Exception:  # a planted mistake may also make the oracle raise
            return
        if _step_mismatches(got, want):
            return
    pytest.fail(f"C comparison missed: {mutation}")


# ---------------------------------------------------------------------------
# the open-ocean bulk fluxes (the forcing producer of the step)
# ---------------------------------------------------------------------------

def _bulk_inputs(seed=77):
    """Seeded open-ocean inputs spanning stable, neutral and unstable
    stratification and calm (including exactly zero) to gale winds."""
    st = synth.config0()
    r = np.random.default_rng(seed)
    shp = st.shape
    tsf = r.uniform(271.0, 303.0, shp)
    ins = [r.normal(0, 8, shp), r.normal(0, 8, shp), tsf + r.normal(0, 6, shp),
           r.uniform(1e-3, 2e-2, shp), tsf]
    ins[0][:3, :3] = 0.0
    ins[1][:3, :3] = 0.0
    return st, ins


def _bulk_mismatches(oracle):
    st, ins = _bulk_inputs()
    want = oracle.bulkf_formula_lanl(st, *ins, st.variables.iceMask, grav=st.settings.gravity)
    got = c_ref.bulkf_formula_lanl(st, *ins, grav=st.settings.gravity)
    bad = []
    for k, (a, b) in enumerate(zip(got, want)):
        b = np.ascontiguousarray(b, dtype=np.float64)
        neq = int(np.count_nonzero(a.view(np.int64) != b.view(np.int64)))
        if neq > 0.1 * a.size or not _rel(a, b) <= 1e-13:
            bad.append((k, neq, _rel(a, b)))
    return bad


def test_bulk_fluxes_match_c_restatement():
    """Every output within 1e-13 of scale and bit-equal in at least 90% of the
    cells (exp/log/atan/pow of NumPy and of libm differ by an ulp here and
    there)."""
    assert not _bulk_mismatches(O), _bulk_mismatches(O)


BULK_MUTATIONS = sorted(n for n, (a, _) in MUTATIONS.items()
                        if a in inspect.getsource(O.bulkf_formula_lanl))


def test_bulk_mutation_list_is_not_empty():
    assert len(BULK_MUTATIONS) >= 8


@pytest.mark.parametrize("mutation", BULK_MUTATIONS)
def test_c_comparison_catches_bulk_flux_mistake(mutation, tmp_path):
    assert _bulk_mismatches(_mutant(mutation, tmp_path)), f"C comparison missed: {mutation}"


# ---------------------------------------------------------------------------
# residual norms of the subcycle (computeEvpResidual)
# ---------------------------------------------------------------------------

RES_VARIANTS = ("default", "adaptive", "adaptive_regional", "side_drag_adaptive")


def _residual_mismatches(oracle, variant, n=9):
    st = _state(variant)
    want = oracle.evp_solver_residuals(st, n)
    got = c_ref.evp_solver_residuals(st, n)
    bad = [k for k in range(5) if not np.array_equal(got[k], want[k])]
    for k in (5, 6):  # FP64 sums in another order: 1e-12 per substep
        rel = np.abs(got[k] - want[k]) / np.maximum(np.abs(want[k]), 1e-300)
        if not float(rel.max()) <= 1e-12:
            bad.append(k)
    return bad


@pytest.mark.parametrize("variant", RES_VARIANTS)
def test_residuals_match_c_restatement(variant):
    """The subcycle bit for bit, the per-substep residual norms (sums in
    another order) within 1e-12."""
    assert not _residual_mismatches(O, variant)


RES_MUTATIONS = sorted(n for n, (a, _) in MUTATIONS.items()
                       if a in inspect.getsource(O.evp_residuals))


@pytest.mark.parametrize("mutation", RES_MUTATIONS)
def test_c_comparison_catches_residual_mistake(mutation, tmp_path):
    mod = _mutant(mutation, tmp_path)
    assert any(_residual_mismatches(mod, v) for v in RES_VARIANTS), f"missed: {mutation}"


def test_openmp_threads_change_no_bit():
    """On a grid large enough for the OpenMP cell loops (1 degree, 59,696
    padded cells), a whole step of the C restatement on one thread and on
    four threads gives the same bits in every field."""
    st = synth.config1()
    q0, sw = st.variables.Qnet.copy(), st.variables.SWdown.copy()
    outs = []
    try:
        for threads in (1, 4):
            assert c_ref.set_threads(threads) == threads
            outs.append(c_ref.seaice_step(synth.config1(), q0, sw, n_substeps=6))
    finally:
        c_ref.set_threads(1)
    one, four = outs
    assert one.keys() == four.keys()
    for k in one:
        assert np.array_equal(one[k].view(np.int64), four[k].view(np.int64)), k
