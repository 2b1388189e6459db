"""ctypes loader of the scalar-loop C restatement (``oracle/c/``) -- TEST
INFRASTRUCTURE ONLY.

The C code restates ``veris/evp_solver.py:22-319`` and
``veris/freedrift_solver.py:5-62`` cell by cell with wrapped indices; the
tests run it beside the NumPy oracle (``oracle/veris_np.py``) on the same
state and require the two to agree bit for bit where only +, -, *, /, sqrt
and comparisons are involved.  ``build()`` compiles it with gcc through
``oracle/c/Makefile``.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c")
LIB = os.path.join(HERE, "_build", "libevp_ref.so")

# field and setting orders of oracle/c/evp_ref.h
SETTINGS = ("NX", "NY", "noSlip", "secondOrderBC", "PlasDefCoeff", "deltaMin", "tensileStrFac",
            "pressReplFac", "waterIceDrag", "waterIceDrag_south", "rhoSea", "cDragMin",
            "basalDragK1", "basalDragK2", "basalDragU0", "cBasalStar", "use_coastline",
            "sideDragCoeff", "sideDragU0", "veros_fill", "enable_cyclic_x", "useAdaptiveEVP",
            "deltatDyn", "aEvpCoeff", "evpAlpha", "aEVPalphaMin", "evpBeta", "recip_deltatDyn",
            "cosWat", "sinWat", "rhoIce")
FIELDS = ("uIce", "vIce", "sigma1", "sigma2", "sigma12",
          "iceMask", "iceMaskU", "iceMaskV", "maskInC", "maskInU", "maskInV",
          "recip_dxU", "recip_dyU", "recip_dxV", "recip_dyV", "dxU", "dyU", "dxV", "dyV",
          "recip_rA", "rAz", "recip_rAu", "recip_rAv",
          "k1AtC", "k2AtC", "k1AtZ", "k2AtZ",
          "SeaIceStrength", "fCori", "uOcean", "vOcean", "Area", "hIceMean", "R_low",
          "AreaW", "AreaS", "SeaIceMassC", "SeaIceMassU", "SeaIceMassV",
          "WindForcingX", "WindForcingY", "Fu", "Fv")

_lib = None


def build() -> str:
    """Compile the C restatement (gcc, -ffp-contract=off); returns the library path."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        lib.evp_solver_ref.argtypes = [dp, ctypes.POINTER(dp), ctypes.c_int] + [dp] * 5
        lib.evp_solver_ref.restype = ctypes.c_int
        lib.freedrift_ref.argtypes = [dp, ctypes.POINTER(dp), dp, dp]
        lib.freedrift_ref.restype = ctypes.c_int
        _lib = lib
    return _lib


def _pack(state):
    vs, s = state.variables, state.settings
    NX, NY = np.shape(vs.uIce)
    vals = {"NX": NX, "NY": NY}
    sett = np.array([float(vals[k]) if k in vals else float(s[k]) for k in SETTINGS])
    fields = [np.ascontiguousarray(np.broadcast_to(np.asarray(vs[k], dtype=np.float64), (NX, NY)))
              for k in FIELDS]
    dp = ctypes.POINTER(ctypes.c_double)
    ptrs = (dp * len(fields))(*[f.ctypes.data_as(dp) for f in fields])
    return (NX, NY), sett, fields, ptrs


def evp_solver(state, n_substeps):
    """Returns (uIce, vIce, sigma1, sigma2, sigma12) after n_substeps."""
    lib = _load()
    shape, sett, keep, ptrs = _pack(state)
    out = [np.empty(shape) for _ in range(5)]
    dp = ctypes.POINTER(ctypes.c_double)
    rc = lib.evp_solver_ref(sett.ctypes.data_as(dp), ptrs, int(n_substeps),
                            *[o.ctypes.data_as(dp) for o in out])
    if rc != 0:
        raise MemoryError("evp_solver_ref: allocation failed")
    del keep
    return tuple(out)


def freedrift_solver(state):
    """Returns (uIce, vIce) of the closed-form free drift."""
    lib = _load()
    shape, sett, keep, ptrs = _pack(state)
    out = [np.empty(shape) for _ in range(2)]
    dp = ctypes.POINTER(ctypes.c_double)
    rc = lib.freedrift_ref(sett.ctypes.data_as(dp), ptrs, *[o.ctypes.data_as(dp) for o in out])
    if rc != 0:
        raise MemoryError("freedrift_ref: allocation failed")
    del keep
    return tuple(out)
