"""The C ABI library loads on the CPU host and matches its header.

No compute call is made here (there is no GPU): only symbol presence, struct
layouts against include/veris_hip.h, and host-side argument validation.
"""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

from veris_amd import _native as N
from veris_amd.settings import SETTINGS
from veris_amd.variables import VARIABLE_NAMES

HEADER = Path(__file__).resolve().parent.parent / "include" / "veris_hip.h"


def _header():
    return HEADER.read_text()


def _struct_members(text, name):
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    members = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        names = decl.split(None, 1)[1] if not decl.startswith("const") else decl.split(None, 2)[2]
        for n in names.split(","):
            members.append(n.strip().lstrip("*").strip())
    return members


def test_every_declared_function_is_exported_and_bound():
    lib = N.lib()
    decls = set(re.findall(r"^int (vh_\w+)\(", _header(), re.M))
    assert decls, "no declarations parsed"
    for name in decls:
        assert hasattr(lib, name), f"{name} not exported by {N.LIB_PATH}"
    assert decls == set(N.SIGNATURES), decls ^ set(N.SIGNATURES)


def test_struct_layouts_match_header():
    text = _header()
    assert _struct_members(text, "vh_settings") == list(SETTINGS)
    assert _struct_members(text, "vh_fields") == list(VARIABLE_NAMES)
    assert _struct_members(text, "vh_domain") == [f[0] for f in N.Domain._fields_]
    assert _struct_members(text, "vh_growth_out") == [f[0] for f in N.GrowthOut._fields_]
    assert ctypes.sizeof(N.Settings) == 8 * len(SETTINGS)
    assert ctypes.sizeof(N.Fields) == 8 * len(VARIABLE_NAMES)
    assert ctypes.sizeof(N.Domain) == 32


def test_version_and_host_side_validation():
    lib = N.lib()
    assert lib.vh_version() == N.ABI_VERSION
    dom = N.Domain(nx=90, ny=40, olx=2, oly=2, cyclic_x=1, fill_mode=0, dtype=N.VH_F64, tile=0)
    assert N.workspace_bytes("vh_evp_workspace_bytes", dom) >= 5 * 94 * 44 * 8
    assert N.workspace_bytes("vh_advection_workspace_bytes", dom) >= 6 * 94 * 44 * 8
    bad = N.Domain(nx=90, ny=40, olx=3, oly=2, cyclic_x=1, fill_mode=0, dtype=N.VH_F64, tile=0)
    with pytest.raises(N.VerisHipError, match="VH_ESHAPE.*olx=3"):
        N.workspace_bytes("vh_evp_workspace_bytes", bad)
    bad = N.Domain(nx=90, ny=40, olx=2, oly=2, cyclic_x=1, fill_mode=0, dtype=7, tile=0)
    with pytest.raises(N.VerisHipError, match="VH_EDTYPE"):
        N.workspace_bytes("vh_evp_workspace_bytes", bad)
    shapes = N.evp_tile_shapes()
    assert len(shapes) == lib.vh_evp_tile_count()
    # the GPU tests parametrize over the shape list read from the kernel source
    from gpu_helpers import evp_shapes_from_source

    assert shapes == evp_shapes_from_source()
    for ri, rj, k in shapes.values():
        # rj a power of two (LDS row stride); ri free, so a workgroup may end in a
        # partial wave (13 x 32 = 6.5 waves)
        assert ri * rj <= 1024 and rj & (rj - 1) == 0 and ri > 4 * k and rj > 4 * k


def test_null_arguments_are_rejected_without_touching_the_gpu():
    dom = N.Domain(nx=8, ny=6, olx=2, oly=2, cyclic_x=1, fill_mode=0, dtype=N.VH_F64, tile=0)
    sett = N.Settings()
    fields = N.Fields()
    with pytest.raises(N.VerisHipError, match="VH_EINVAL|VH_EWORKSPACE"):
        N.call("vh_evp_solver", ctypes.byref(dom), ctypes.byref(sett), ctypes.byref(fields), 5,
               None, None, None, None, None, None, 0, None)
    with pytest.raises(N.VerisHipError, match="VH_EINVAL"):
        N.call("vh_seaice_mass", ctypes.byref(dom), ctypes.byref(sett), ctypes.byref(fields),
               None, None, None, None)


def test_adjoint_and_batch_structs_match_header():
    text = _header()
    assert _struct_members(text, "vh_evp_adjoint_out") == [f[0] for f in N.EvpAdjointOut._fields_]
    assert _struct_members(text, "vh_pre_evp_adjoint_out") == [
        f[0] for f in N.PreEvpAdjointOut._fields_]
    assert _struct_members(text, "vh_batch") == [f[0] for f in N.Batch._fields_]
    assert _struct_members(text, "vh_member") == [f[0] for f in N.Member._fields_]
    assert _struct_members(text, "vh_pre_evp_out") == [f[0] for f in N.PreEvpOut._fields_]
    assert _struct_members(text, "vh_ocean_to_ice_out") == [f[0] for f in N.OceanToIceOut._fields_]
    assert _struct_members(text, "vh_ice_to_ocean_out") == [f[0] for f in N.IceToOceanOut._fields_]


def test_adjoint_host_side_validation():
    from veris_amd.settings import make_settings

    dom = N.Domain(nx=360, ny=160, olx=2, oly=2, cyclic_x=1, fill_mode=0, dtype=N.VH_F64, tile=0)
    n = ctypes.c_size_t(0)
    N.call("vh_evp_adjoint_workspace_bytes", ctypes.byref(dom), 120, 10, ctypes.byref(n))
    fb = 364 * 164 * 8
    # checkpoints, ping-pong, two segment sets and the state adjoints at least
    assert n.value >= (13 + 2 + 22 + 10) * 5 * fb // 5
    m = ctypes.c_size_t(0)
    N.call("vh_evp_adjoint_workspace_bytes", ctypes.byref(dom), 120, 1, ctypes.byref(m))
    assert m.value > n.value  # every state kept
    with pytest.raises(N.VerisHipError, match="VH_EINVAL"):
        N.call("vh_evp_adjoint_workspace_bytes", ctypes.byref(dom), 0, 10, ctypes.byref(n))
    fields = N.Fields()
    for kw, msg in (({"useEVP": False}, "VH_EUNSUPPORTED.*EVP solver"),
                    ({"basalDragK2": 8.0}, "VH_EINVAL.*basal drag needs")):
        sett = N.settings_struct(make_settings(**kw))
        with pytest.raises(N.VerisHipError, match=msg):
            N.call("vh_evp_forward_checkpointed", ctypes.byref(dom), ctypes.byref(sett),
                   ctypes.byref(fields), 120, 10, None, None, None, None, None, None, 0, None)
    sett = N.settings_struct(make_settings())
    with pytest.raises(N.VerisHipError, match="VH_EINVAL.*missing input"):
        N.call("vh_evp_forward_checkpointed", ctypes.byref(dom), ctypes.byref(sett),
               ctypes.byref(fields), 120, 10, None, None, None, None, None, None, 0, None)
    multi = N.Domain(nx=360, ny=160, olx=2, oly=2, cyclic_x=1, fill_mode=0, dtype=N.VH_F64, tile=11)
    with pytest.raises(N.VerisHipError, match="VH_EINVAL.*one or two substeps"):
        N.call("vh_evp_adjoint", ctypes.byref(multi), ctypes.byref(sett), ctypes.byref(fields),
               120, 10, None, None, None, None, None, ctypes.byref(N.EvpAdjointOut()), None, 0,
               None)


def test_batch_arith_validation():
    members = (N.Member * 1)(N.Member(nx=64, ny=64, cyclic_x=0, fill_mode=0, offset=0))
    n = ctypes.c_size_t(0)
    ok = N.Batch(n_members=1, dtype=N.VH_F32, tile=0, arith=1, members=members,
                 total_elems=68 * 68)
    N.call("vh_evp_batched_workspace_bytes", ctypes.byref(ok), ctypes.byref(n))
    bad = N.Batch(n_members=1, dtype=N.VH_F64, tile=0, arith=1, members=members,
                  total_elems=68 * 68)
    with pytest.raises(N.VerisHipError, match="VH_EINVAL.*arith"):
        N.call("vh_evp_batched_workspace_bytes", ctypes.byref(bad), ctypes.byref(n))
    # arith 2: FP32 step constants around an FP64 state, whose five ping-pong
    # slots are FP64 (4 more bytes per element each)
    n2 = ctypes.c_size_t(0)
    st64 = N.Batch(n_members=1, dtype=N.VH_F32, tile=0, arith=2, members=members,
                   total_elems=68 * 68)
    N.call("vh_evp_batched_workspace_bytes", ctypes.byref(st64), ctypes.byref(n2))
    assert n2.value >= n.value + 5 * (68 * 68 * 4 - 256)
    for dtype, arith in ((N.VH_F64, 2), (N.VH_F32, 3)):
        bad = N.Batch(n_members=1, dtype=dtype, tile=0, arith=arith, members=members,
                      total_elems=68 * 68)
        with pytest.raises(N.VerisHipError, match="VH_EINVAL.*arith"):
            N.call("vh_evp_batched_workspace_bytes", ctypes.byref(bad), ctypes.byref(n))


def test_forcing_residual_and_free_drift_validation():
    assert _struct_members(_header(), "vh_bulk_flux_out") == [f[0] for f in N.BulkFluxOut._fields_]
    dom = N.Domain(nx=90, ny=40, olx=2, oly=2, cyclic_x=1, fill_mode=0, dtype=N.VH_F64, tile=0)
    sett = N.Settings()
    fields = N.Fields()
    with pytest.raises(N.VerisHipError, match="VH_EINVAL.*output struct"):
        N.call("vh_bulkf_formula_lanl", ctypes.byref(dom), ctypes.byref(sett), 9.81,
               None, None, None, None, None, None, None, None)
    with pytest.raises(N.VerisHipError, match="VH_EINVAL.*null input"):
        N.call("vh_bulkf_formula_lanl", ctypes.byref(dom), ctypes.byref(sett), 9.81,
               None, None, None, None, None, None, ctypes.byref(N.BulkFluxOut()), None)
    with pytest.raises(N.VerisHipError, match="VH_EINVAL.*freedrift"):
        N.call("vh_freedrift_solver", ctypes.byref(dom), ctypes.byref(sett), ctypes.byref(fields),
               None, None, None)
    base = N.workspace_bytes("vh_evp_workspace_bytes", dom)
    n = ctypes.c_size_t(0)
    N.call("vh_evp_residuals_workspace_bytes", ctypes.byref(dom), 400, ctypes.byref(n))
    assert n.value > base
    with pytest.raises(N.VerisHipError, match="VH_EINVAL"):
        N.call("vh_evp_residuals_workspace_bytes", ctypes.byref(dom), 0, ctypes.byref(n))
    with pytest.raises(N.VerisHipError, match="VH_EINVAL.*residual outputs"):
        N.call("vh_evp_solver_residuals", ctypes.byref(dom), ctypes.byref(sett),
               ctypes.byref(fields), 10, None, None, None, None, None, None, None, None, 0, None)
