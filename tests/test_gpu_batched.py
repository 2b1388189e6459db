"""GPU parity of the batched ensemble path.

FP64 members of uneven sizes (closed regional domains and one x-periodic
global grid) run through vh_evp_solver_batched and must equal the NumPy
oracle bit for bit, member by member, for single-substep shapes, and for
multi-substep shapes on every member without a periodic fill.  The FP32 path
is compared with the FP64 oracle at the tolerance stated in its test.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from gpu_helpers import assert_exact, evp_shapes_from_source, g2n, pre_evp, report
from oracle import veris_np as O
from veris_amd import synth
from veris_amd.batched import Ensemble

pytestmark = pytest.mark.gpu

EVP_NAMES = ("uIce", "vIce", "sigma1", "sigma2", "sigma12")


def _members():
    sizes = [(37, 29), (64, 23), (9, 71), (130, 41), (17, 17)]
    states = [pre_evp(synth.regional(21, m, nx, ny)) for m, (nx, ny) in enumerate(sizes)]
    states.append(pre_evp(synth.tiny(20, 14, config=22)))  # x-periodic member
    for st in states:  # spin up so stresses and velocities are non-zero
        vs = st.variables
        vs.uIce, vs.vIce, vs.sigma1, vs.sigma2, vs.sigma12 = O.evp_solver(st, n_substeps=5)
        pre_evp(st)
    return states


@pytest.mark.parametrize("tile", sorted(evp_shapes_from_source()))
def test_batched_fp64_matches_oracle(cuda, tile):
    from veris_amd._native import evp_tile_shapes

    k = evp_tile_shapes()[tile][2]
    states = _members()
    ens = Ensemble(states, device=cuda, dtype=torch.float64)
    for n in (3, 17):
        out = ens.evp(n, tile=tile)
        for m, st in enumerate(states):
            want = O.evp_solver(st, n_substeps=n)
            for q, name in enumerate(EVP_NAMES):
                got = g2n(ens.view(out[q], m))
                ref = want[q]
                if k > 1 and st.settings.enable_cyclic_x and q >= 2:
                    keep = np.setdiff1d(np.arange(st.shape[0]), [0, 1, st.shape[0] - 2, st.shape[0] - 1])
                    got, ref = got[keep], ref[keep]
                assert_exact(f"member {m} {name}[tile={tile}, n={n}]", got, ref)


@pytest.mark.parametrize("case", ["half_mask", "curvilinear"])
def test_batched_packed_constants_fallback(cuda, case):
    """One member whose grid does not allow the compact step constants (a mask
    other than 0/1, metrics varying along i) sends the whole batch to the full
    fields; every member still equals the oracle bit for bit."""
    states = _members()
    vs = states[1].variables
    if case == "half_mask":
        vs.maskInV = vs.maskInV.copy()
        vs.maskInV[10, 10] = 0.5
    else:
        wig = 1.0 + 1e-3 * np.cos(np.arange(states[1].shape[0]))[:, None]
        vs.dyU = vs.dyU * wig
        vs.rAz = vs.rAz * wig
    ens = Ensemble(states, device=cuda, dtype=torch.float64)
    out = ens.evp(9, tile=5)
    for m, st in enumerate(states):
        want = O.evp_solver(st, n_substeps=9)
        for q, name in enumerate(EVP_NAMES):
            assert_exact(f"member {m} {name}[{case}]", g2n(ens.view(out[q], m)), want[q])


def test_batched_smaller_workspaces(cuda):
    """A workspace without room for the tile tables (each workgroup then
    searches the member table), and one without room for the packed step
    constants either (full fields), give the same bits as the full one."""
    import ctypes

    from veris_amd import _native as N

    states = _members()
    ens = Ensemble(states, device=cuda, dtype=torch.float64)
    tile, n = 5, 11
    want = [o.clone() for o in ens.evp(n, tile=tile)]
    b = ens._batch(tile)
    full = ctypes.c_size_t(0)
    N.call("vh_evp_batched_workspace_bytes", ctypes.byref(b), ctypes.byref(full))
    ri, rj, k = N.evp_tile_shapes()[tile]
    r5, j5, k5 = N.evp_tile_shapes()[5]

    def tiles(ri_, rj_, k_):
        TI, TJ = ri_ - 4 * k_, rj_ - 4 * k_
        return sum(-(-a // TI) * -(-b_ // TJ) for a, b_ in ens.shapes)

    align = lambda x: (x + 255) // 256 * 256  # noqa: E731
    no_tables = full.value - align((tiles(ri, rj, k) + tiles(r5, j5, k5)) * 48)
    rows = sum(b_ for _, b_ in ens.shapes)
    no_pack = no_tables - align(ens.total) - align(12 * rows * 8)
    sett = N.settings_struct(ens.settings)
    fields = ens._fields_struct()
    stream = torch.cuda.current_stream(cuda).cuda_stream
    for size in (no_tables, no_pack):
        ws = torch.empty(size, dtype=torch.uint8, device=cuda)
        out = [torch.empty(ens.total, dtype=torch.float64, device=cuda) for _ in range(5)]
        N.call("vh_evp_solver_batched", ctypes.byref(b), ctypes.byref(sett), ctypes.byref(fields),
               n, *[t.data_ptr() for t in out], ws.data_ptr(), size, stream)
        for q, name in enumerate(EVP_NAMES):
            assert torch.equal(out[q], want[q]), f"{name} differs with a {size}-byte workspace"
    ws = torch.empty(no_pack - 256, dtype=torch.uint8, device=cuda)
    with pytest.raises(N.VerisHipError, match="VH_EWORKSPACE"):
        N.call("vh_evp_solver_batched", ctypes.byref(b), ctypes.byref(sett), ctypes.byref(fields),
               n, *[t.data_ptr() for t in out], ws.data_ptr(), no_pack - 256, stream)


def test_batched_pre_evp_and_budgets(cuda):
    states = [synth.regional(23, m, nx, ny) for m, (nx, ny) in enumerate([(40, 30), (25, 61)])]
    ens = Ensemble(states, device=cuda, dtype=torch.float64)
    ens.pre_evp()
    for m, st in enumerate(states):
        pre_evp(st)
        for name in ("SeaIceMassU", "AreaS", "WindForcingX"):
            assert_exact(f"member {m} {name}", g2n(ens.view(ens.fields[name], m)), st.variables[name])
    b = ens.budgets(("hIceMean", "Area"))
    for m, st in enumerate(states):
        vs = st.variables
        want = O.global_sum((vs.rA * vs.hIceMean)[2:-2, 2:-2])
        assert abs(b["hIceMean"][m].item() - want) <= 1e-13 * abs(want)


def test_batched_fp32_against_fp64_oracle(cuda):
    """FP32 bulk with FP64 Delta/zeta/press, 400 substeps from rest, against
    the FP64 oracle, relative L2 error over interior ice-covered cells.

    Measured: velocities 0.4-1.5e-4, stresses 0.5-3e-3.  The subcycle
    amplifies rounding chaotically: rounding only the inputs to FP32 already
    moves the FP64 stresses by ~5e-4 (velocities by ~4e-6), and FP64
    arithmetic on FP32-stored state reaches ~1e-5 in velocity.  Bounds:
    velocities 3e-4, stresses 1e-2."""
    sizes = [(64, 64), (96, 80), (70, 112)]
    states = [pre_evp(synth.regional(2, m, nx, ny)) for m, (nx, ny) in enumerate(sizes)]
    ens = Ensemble(states, device=cuda, dtype=torch.float32)
    out = ens.evp(400, tile=3)
    worst = {}
    for m, st in enumerate(states):
        want = O.evp_solver(st, n_substeps=400)
        ice = (st.variables.iceMask > 0)[2:-2, 2:-2]
        for q, name in enumerate(EVP_NAMES):
            got = g2n(ens.view(out[q], m))[2:-2, 2:-2][ice]
            ref = want[q][2:-2, 2:-2][ice]
            rel = float(np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300))
            worst[name] = max(worst.get(name, 0.0), rel)
            print(f"member {m} {name}: rel L2 {rel:.2e}", flush=True)
    assert worst["uIce"] <= 3e-4 and worst["vIce"] <= 3e-4, worst
    assert max(worst["sigma1"], worst["sigma2"], worst["sigma12"]) <= 1e-2, worst


def test_batched_fp32_storage_fp64_arith_against_fp64_oracle(cuda):
    """FP32-stored fields with FP64 arithmetic in every substep (arith="fp64"),
    400 substeps, against the FP64 oracle over interior ice-covered cells.
    Only the FP32 rounding of the stored state between substeps remains; a
    NumPy emulation of that rounding gives ~1e-5 in velocity.  Bounds:
    velocities 1e-4, stresses 5e-3."""
    sizes = [(64, 64), (96, 80), (70, 112)]
    states = [pre_evp(synth.regional(2, m, nx, ny)) for m, (nx, ny) in enumerate(sizes)]
    ens = Ensemble(states, device=cuda, dtype=torch.float32, arith="fp64")
    for tile in (3, 9):
        out = ens.evp(400, tile=tile)
        worst = {}
        for m, st in enumerate(states):
            want = O.evp_solver(st, n_substeps=400)
            ice = (st.variables.iceMask > 0)[2:-2, 2:-2]
            for q, name in enumerate(EVP_NAMES):
                got = g2n(ens.view(out[q], m))[2:-2, 2:-2][ice]
                ref = want[q][2:-2, 2:-2][ice]
                rel = float(np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300))
                worst[name] = max(worst.get(name, 0.0), rel)
        print(f"tile {tile}: " + ", ".join(f"{k} {v:.2e}" for k, v in worst.items()), flush=True)
        assert worst["uIce"] <= 1e-4 and worst["vIce"] <= 1e-4, worst
        assert max(worst["sigma1"], worst["sigma2"], worst["sigma12"]) <= 5e-3, worst


@pytest.mark.parametrize("tile", [3, 9])
def test_batched_fp64_state_equals_fp64_on_fp32_constants(cuda, tile):
    """arith="fp64_state" (FP32-stored step constants, FP64 state and
    arithmetic) on step constants that are FP32-representable and a spun-up
    FP64 state equals the FP64 batched subcycle on the same values: the FP32
    storage is then lossless, so only the multiply-adds its instantiation
    fuses separate the two (measured ~1e-14 after 17 substeps; bound 1e-11
    of each field's scale, per member)."""
    states = _members()
    for st in states:
        vs = st.variables
        for k in list(vs.keys()):
            a = vs[k]
            if k not in EVP_NAMES and isinstance(a, np.ndarray) and a.dtype == np.float64:
                vs[k] = a.astype(np.float32).astype(np.float64)
    e64 = Ensemble(states, device=cuda, dtype=torch.float64)
    emx = Ensemble(states, device=cuda, dtype=torch.float32, arith="fp64_state")
    ref = e64.evp(17, tile=tile)
    got = emx.evp(17, tile=tile)
    worst = 0.0
    for q, name in enumerate(EVP_NAMES):
        assert got[q].dtype == torch.float64
        for m in range(len(states)):
            r = g2n(e64.view(ref[q], m))
            d = float(np.max(np.abs(g2n(emx.view(got[q], m)) - r)))
            rel = d / max(float(np.max(np.abs(r))), 1e-300)
            worst = max(worst, rel)
            assert rel <= 1e-11, (name, m, rel)
    print(f"tile {tile}: worst relative difference {worst:.2e}", flush=True)


BRANCHES = {"adaptive": {"useAdaptiveEVP": True}, "side_drag": {"noSlip": False},
            "second_order": {"secondOrderBC": True}}


@pytest.mark.parametrize("tile", [3, 9])
@pytest.mark.parametrize("branch", sorted(BRANCHES))
def test_batched_branches_fp64_match_oracle(cuda, branch, tile):
    """Adaptive EVP (per-cell alpha/beta), free slip with lateral drag and
    the second-order no-slip boundary term in the batched launch: every
    member bit for bit against the oracle (closed regional members); the
    second-order term refuses multi-substep shapes."""
    sizes = [(37, 29), (64, 23), (30, 61)]
    states = [pre_evp(synth.regional(21, m, nx, ny, **BRANCHES[branch]))
              for m, (nx, ny) in enumerate(sizes)]
    for st in states:
        vs = st.variables
        vs.uIce, vs.vIce, vs.sigma1, vs.sigma2, vs.sigma12 = O.evp_solver(st, n_substeps=5)
        pre_evp(st)
    ens = Ensemble(states, device=cuda, dtype=torch.float64)
    if branch == "second_order" and tile == 9:
        from veris_amd._native import VerisHipError

        with pytest.raises(VerisHipError, match="VH_EUNSUPPORTED.*secondOrderBC"):
            ens.evp(17, tile=tile)
        return
    out = ens.evp(17, tile=tile)
    for m, st in enumerate(states):
        want = O.evp_solver(st, n_substeps=17)
        for q, name in enumerate(EVP_NAMES):
            assert_exact(f"{branch} member {m} {name}[tile={tile}]", g2n(ens.view(out[q], m)),
                         want[q])


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batched_pre_evp_and_budgets_one_launch(cuda, dtype):
    """vh_pre_evp_batched equals vh_pre_evp on every member bit for bit (with
    a packed P* field too), and vh_global_sums_batched gives each member's
    FP64 budgets of h, hs, A and Qnet (weight rA) within 1e-13 of math.fsum;
    70 members, so the launches run in two chunks of at most 64."""
    import math

    from veris_amd import kernels as K

    sizes = [(9 + (m * 7) % 23, 8 + (m * 5) % 19) for m in range(70)]
    states = [synth.regional(24, m, nx, ny) for m, (nx, ny) in enumerate(sizes)]
    states[3] = synth.tiny(20, 14, config=25)  # one x-periodic member
    ens = Ensemble(states, device=cuda, dtype=dtype)
    rng = np.random.default_rng(3)
    pstar = torch.from_numpy(rng.lognormal(np.log(27500.0), 0.3, ens.total)).to(cuda, dtype)
    for ps in (None, pstar):
        ens.pre_evp(ps)
        for m in range(ens.n):
            st = ens.member_state(m)
            want = K.pre_evp(st, None if ps is None else ens.view(ps, m))
            for name, w in zip(("SeaIceMassC", "SeaIceMassU", "SeaIceMassV", "AreaW", "AreaS",
                                "WindForcingX", "WindForcingY", "SeaIceStrength"), want):
                got = ens.view(ens.fields[name], m)
                assert torch.equal(got, w), f"member {m} {name} (P* field: {ps is not None})"
    b = ens.budgets()
    for m, st in enumerate(states):
        vs = ens.member_state(m).variables
        rA = vs.rA[2:-2, 2:-2].double().cpu().numpy().ravel()
        for name in ("hIceMean", "hSnowMean", "Area", "Qnet"):
            x = vs[name][2:-2, 2:-2].double().cpu().numpy().ravel()
            want = math.fsum(x * rA)
            scale = math.fsum(np.abs(x * rA))
            assert abs(b[name][m].item() - want) <= 1e-13 * max(scale, 1e-300), (m, name)
