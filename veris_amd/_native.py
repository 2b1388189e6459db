"""ctypes binding of ``libveris_hip.so`` (declared in ``include/veris_hip.h``).

PyTorch tensors are only the carrier: the binding passes their device
pointers and the current HIP stream to the C ABI.  There is no fallback: if
the library is missing or a call fails, a :class:`VerisHipError` is raised.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

from .settings import SETTINGS
from .variables import VARIABLE_NAMES

LIB_PATH = Path(os.environ.get("VERIS_HIP_LIB", Path(__file__).resolve().parent / "libveris_hip.so"))

VH_F64, VH_F32 = 0, 1
ABI_VERSION = 1

_ERRORS = {
    -1: "VH_EINVAL",
    -2: "VH_ESHAPE",
    -3: "VH_EDTYPE",
    -4: "VH_EUNSUPPORTED",
    -5: "VH_EHIP",
    -6: "VH_EWORKSPACE",
}


class VerisHipError(RuntimeError):
    def __init__(self, func: str, code: int, message: str):
        super().__init__(f"{func}: {_ERRORS.get(code, code)}: {message}")
        self.code = code


class Domain(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("nx", "ny", "olx", "oly", "cyclic_x", "fill_mode", "dtype", "tile")]


class Settings(ctypes.Structure):
    _fields_ = [(name, ctypes.c_double) for name in SETTINGS]


class Fields(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in VARIABLE_NAMES]


GROWTH_OUT_NAMES = ("hIceMean", "hSnowMean", "Area", "TSurf", "EmPmR", "forc_salt_surface", "Qsw",
                    "Qnet", "SeaIceLoad", "IcePenetSW", "recip_hIceMean")


class GrowthOut(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in GROWTH_OUT_NAMES]


EVP_ADJOINT_NAMES = ("uIce", "vIce", "sigma1", "sigma2", "sigma12", "uNm1", "vNm1",
                     "SeaIceStrength", "SeaIceMassC", "SeaIceMassU", "SeaIceMassV", "AreaW",
                     "AreaS", "WindForcingX", "WindForcingY", "waterDrag_north",
                     "waterDrag_south", "hIceMean")


class EvpAdjointOut(ctypes.Structure):
    _fields_ = [("adj_" + name, ctypes.c_void_p) for name in EVP_ADJOINT_NAMES]


PRE_EVP_ADJOINT_NAMES = ("pStar", "hIceMean", "hSnowMean", "airDrag_north", "airDrag_south")


class PreEvpAdjointOut(ctypes.Structure):
    _fields_ = [("adj_" + name, ctypes.c_void_p) for name in PRE_EVP_ADJOINT_NAMES]


PRE_EVP_NAMES = ("SeaIceMassC", "SeaIceMassU", "SeaIceMassV", "AreaW", "AreaS", "WindForcingX",
                 "WindForcingY", "SeaIceStrength")


class PreEvpOut(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in PRE_EVP_NAMES]


OCEAN_TO_ICE_NAMES = ("uOcean", "vOcean", "theta", "ocSalt", "Qnet", "Qsw", "wSpeed")


class OceanToIceOut(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in OCEAN_TO_ICE_NAMES]


ICE_TO_OCEAN_NAMES = ("surface_taux", "surface_tauy", "forc_temp_surface", "forc_salt_surface")


class IceToOceanOut(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in ICE_TO_OCEAN_NAMES]


BULK_FLUX_NAMES = ("flwupa", "flha", "fsha", "df0dT", "ust", "vst", "evp", "ssq", "dEvdT")


class BulkFluxOut(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in BULK_FLUX_NAMES]


class Member(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("cyclic_x", ctypes.c_int32),
                ("fill_mode", ctypes.c_int32), ("offset", ctypes.c_int64)]


class Batch(ctypes.Structure):
    _fields_ = [("n_members", ctypes.c_int32), ("dtype", ctypes.c_int32), ("tile", ctypes.c_int32),
                ("arith", ctypes.c_int32), ("members", ctypes.POINTER(Member)),
                ("total_elems", ctypes.c_int64)]


_P = ctypes.c_void_p
_DP = ctypes.POINTER(Domain)
_SP = ctypes.POINTER(Settings)
_FP = ctypes.POINTER(Fields)
_SZ = ctypes.c_size_t
_I32 = ctypes.c_int32

# name -> argtypes (all return int)
SIGNATURES = {
    "vh_version": [],
    "vh_last_error": [ctypes.c_char_p, _SZ],
    "vh_seaice_mass": [_DP, _SP, _FP, _P, _P, _P, _P],
    "vh_area_ws": [_DP, _SP, _FP, _P, _P, _P],
    "vh_seaice_strength": [_DP, _SP, _FP, _P, _P, _P],
    "vh_wind_forcing_xy": [_DP, _SP, _FP, _P, _P, _P],
    "vh_evp_tile_count": [],
    "vh_evp_tile_shape": [_I32, ctypes.POINTER(_I32), ctypes.POINTER(_I32), ctypes.POINTER(_I32)],
    "vh_evp_workspace_bytes": [_DP, ctypes.POINTER(_SZ)],
    "vh_evp_overlaps_periodic": [_DP, _SP, _FP, _P, _P],
    "vh_evp_solver": [_DP, _SP, _FP, _I32, _P, _P, _P, _P, _P, _P, _SZ, _P],
    "vh_evp_solver_noalloc": [_DP, _SP, _FP, _I32, _P, _P, _P, _P, _P, _P],
    "vh_evp_batched_workspace_bytes": [ctypes.POINTER(Batch), ctypes.POINTER(_SZ)],
    "vh_evp_solver_batched": [ctypes.POINTER(Batch), _SP, _FP, _I32, _P, _P, _P, _P, _P, _P, _SZ,
                              _P],
    "vh_evp_adjoint_workspace_bytes": [_DP, _I32, _I32, ctypes.POINTER(_SZ)],
    "vh_evp_forward_checkpointed": [_DP, _SP, _FP, _I32, _I32, _P, _P, _P, _P, _P, _P, _SZ, _P],
    "vh_evp_adjoint": [_DP, _SP, _FP, _I32, _I32, _P, _P, _P, _P, _P,
                       ctypes.POINTER(EvpAdjointOut), _P, _SZ, _P],
    "vh_pre_evp_adjoint": [_DP, _SP, _FP, _P, ctypes.POINTER(EvpAdjointOut),
                           ctypes.POINTER(PreEvpAdjointOut), _P],
    "vh_evp_adjoint_batched_workspace_bytes": [_DP, _I32, _I32, _I32, ctypes.POINTER(_SZ)],
    "vh_evp_forward_checkpointed_batched": [_DP, _I32, _SP, _FP, _I32, _I32, _P, _P, _P, _P, _P,
                                            _P, _SZ, _P],
    "vh_evp_adjoint_batched": [_DP, _I32, _SP, _FP, _I32, _I32, _P, _P, _P, _P, _P,
                               ctypes.POINTER(EvpAdjointOut), _P, _SZ, _P],
    "vh_pre_evp_adjoint_batched": [_DP, _I32, _SP, _FP, _P, ctypes.POINTER(EvpAdjointOut),
                                   ctypes.POINTER(PreEvpAdjointOut), _P],
    "vh_pre_evp": [_DP, _SP, _FP, _P, ctypes.POINTER(PreEvpOut), _P],
    "vh_advection_adjoint_workspace_bytes": [_DP, ctypes.POINTER(_SZ)],
    "vh_advection_adjoint": [_DP, _SP, _FP, _P, _P, _P, _P, _P, _P, _P, _P, _P, _SZ, _P],
    "vh_pre_evp_batched": [ctypes.POINTER(Batch), _SP, _FP, _P, ctypes.POINTER(PreEvpOut), _P],
    "vh_global_sums_batched": [ctypes.POINTER(Batch), _I32, ctypes.POINTER(_P), _P, _P, _P],
    "vh_ocean_to_ice": [_DP, _SP, _FP, _P, _P, _P, _P, _P, ctypes.POINTER(OceanToIceOut), _P],
    "vh_ice_to_ocean": [_DP, _SP, _FP, _P, _P, _P, _P, ctypes.POINTER(IceToOceanOut), _P],
    "vh_ocean_stress_uv": [_DP, _SP, _FP, _P, _P, _P],
    "vh_freedrift_solver": [_DP, _SP, _FP, _P, _P, _P],
    "vh_bulkf_formula_lanl": [_DP, _SP, ctypes.c_double] + [_P] * 6 +
                             [ctypes.POINTER(BulkFluxOut), _P],
    "vh_evp_residuals_workspace_bytes": [_DP, _I32, ctypes.POINTER(_SZ)],
    "vh_evp_solver_residuals": [_DP, _SP, _FP, _I32] + [_P] * 7 + [_P, _SZ, _P],
    "vh_advection_workspace_bytes": [_DP, ctypes.POINTER(_SZ)],
    "vh_advection": [_DP, _SP, _FP, _P, _P, _P, _P, _SZ, _P],
    "vh_clean_up_advection": [_DP, _SP, _FP, _P, _P, _P, _P, _P, _P, _P],
    "vh_ridging": [_DP, _SP, _FP, _P, _P],
    "vh_growth": [_DP, _SP, _FP, ctypes.POINTER(GrowthOut), _P],
    "vh_fill_overlap": [_DP, _P, _P],
    "vh_fill_overlaps": [_DP, _I32, ctypes.POINTER(_P), _P],
    "vh_global_sum_workspace_bytes": [_DP, ctypes.POINTER(_SZ)],
    "vh_global_sum_f64": [_DP, _P, _P, _P, _P, _SZ, _P],
}

_lib = None


def lib():
    """Load the library once; raise if it is missing or its ABI differs."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise VerisHipError("load", -4, f"{LIB_PATH} not found: build it with "
                            "`python -m veris_amd._build` (or __graft_entry__.build())")
    L = ctypes.CDLL(str(LIB_PATH))
    for name, argtypes in SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = argtypes
        fn.restype = ctypes.c_int
    if L.vh_version() != ABI_VERSION:
        raise VerisHipError("load", -4, f"ABI version {L.vh_version()} != {ABI_VERSION}")
    _lib = L
    return L


def last_error() -> str:
    buf = ctypes.create_string_buffer(512)
    lib().vh_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


_fns: dict = {}


def call(name: str, *args) -> None:
    fn = _fns.get(name)
    if fn is None:
        fn = _fns[name] = getattr(lib(), name)
    rc = fn(*args)
    if rc != 0:
        raise VerisHipError(name, rc, last_error())


def settings_struct(sett) -> Settings:
    s = Settings()
    for name in SETTINGS:
        setattr(s, name, float(sett[name]))
    return s


def domain_struct(sett, dtype_code: int, tile: int = 0) -> Domain:
    return Domain(nx=int(sett.nx), ny=int(sett.ny), olx=int(sett.olx), oly=int(sett.oly),
                  cyclic_x=int(bool(sett.enable_cyclic_x)),
                  fill_mode=0 if sett.veros_fill else 1, dtype=dtype_code, tile=int(tile))


def evp_tile_shapes() -> dict:
    """EVP tile index -> (ri, rj, k): workgroup shape and substeps per launch."""
    out = {}
    for t in range(1, lib().vh_evp_tile_count() + 1):
        ri, rj, k = _I32(0), _I32(0), _I32(0)
        call("vh_evp_tile_shape", t, ctypes.byref(ri), ctypes.byref(rj), ctypes.byref(k))
        out[t] = (ri.value, rj.value, k.value)
    return out


def workspace_bytes(name: str, dom: Domain) -> int:
    n = _SZ(0)
    call(name, ctypes.byref(dom), ctypes.byref(n))
    return int(n.value)
