"""ctypes loader of the scalar-loop C restatement (``oracle/c/``) -- TEST
INFRASTRUCTURE ONLY.

The C code restates the sea-ice step cell by cell with wrapped indices:
``veris/evp_solver.py:22-319`` and ``veris/freedrift_solver.py:5-62``
(``evp_ref.c``) and the rest of the step of
``seaice_global_4deg.py:874-932`` (``step_ref.c``).  The tests run it beside
the NumPy oracle (``oracle/veris_np.py``) on the same state and require the
two to agree bit for bit where only +, -, *, /, sqrt and comparisons are
involved.  ``build()`` compiles it with gcc through ``oracle/c/Makefile``.
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

STEP_OUTPUTS = ("SeaIceMassC", "SeaIceMassU", "SeaIceMassV", "AreaW", "AreaS", "WindForcingX",
                "WindForcingY", "SeaIceStrength", "OceanStressU", "OceanStressV", "os_hIceMean",
                "os_hSnowMean", "EmPmR", "forc_salt_surface_ice", "Qsw", "Qnet", "SeaIceLoad",
                "IcePenetSW", "recip_hIceMean")

_lib = None


def build() -> str:
    """Compile the C restatement (gcc, -ffp-contract=off); returns the library path."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def set_threads(n: int = 0) -> int:
    """OpenMP threads of the C restatement's cell loops (``n`` > 0 sets
    them); returns the count in effect."""
    return int(_load().evp_ref_set_threads(int(n)))


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        lib.evp_solver_ref.argtypes = [dp, ctypes.POINTER(dp), ctypes.c_int] + [dp] * 5
        lib.evp_solver_ref.restype = ctypes.c_int
        lib.evp_solver_residuals_ref.argtypes = ([dp, ctypes.POINTER(dp), ctypes.c_int] +
                                                 [dp] * 7)
        lib.evp_solver_residuals_ref.restype = ctypes.c_int
        lib.freedrift_ref.argtypes = [dp, ctypes.POINTER(dp), dp, dp]
        lib.freedrift_ref.restype = ctypes.c_int
        cpp = ctypes.POINTER(ctypes.c_char_p)
        lib.seaice_step_ref.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, cpp,
                                        ctypes.POINTER(dp), ctypes.c_int, cpp, dp, ctypes.c_int]
        lib.seaice_step_ref.restype = ctypes.c_int
        lib.step_ref_missing.restype = ctypes.c_char_p
        lib.bulkf_ref.argtypes = ([ctypes.c_int] + [dp] * 5 + [ctypes.c_double] * 11 +
                                  [ctypes.POINTER(dp)])
        lib.bulkf_ref.restype = ctypes.c_int
        lib.evp_ref_set_threads.argtypes = [ctypes.c_int]
        lib.evp_ref_set_threads.restype = ctypes.c_int
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


def evp_solver_residuals(state, n_substeps):
    """Returns (uIce, vIce, sigma1, sigma2, sigma12, resU, resSig) after
    n_substeps, resU and resSig one value per substep."""
    lib = _load()
    shape, sett, keep, ptrs = _pack(state)
    out = [np.empty(shape) for _ in range(5)] + [np.empty(n_substeps) for _ in range(2)]
    dp = ctypes.POINTER(ctypes.c_double)
    rc = lib.evp_solver_residuals_ref(sett.ctypes.data_as(dp), ptrs, int(n_substeps),
                                      *[o.ctypes.data_as(dp) for o in out])
    if rc != 0:
        raise MemoryError("evp_solver_residuals_ref: allocation failed")
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


def seaice_step(state, qnet_forcing=None, swdown=None, n_substeps=None, strength=None):
    """The whole step on a copy of the state's variables; returns the updated
    variables as a dict (the oracle's ``seaice_step`` updates them in place).
    ``strength``: a SeaIceStrength field to use instead of the C library's
    exp() (for bit-for-bit comparisons of everything else)."""
    from oracle import veris_np as O

    lib = _load()
    vs, s = state.variables, state.settings
    arrays = {k: np.array(vs[k], dtype=np.float64, copy=True) for k in vs.keys()
              if isinstance(vs[k], np.ndarray)}
    if qnet_forcing is not None:
        arrays["Qnet"] = np.array(qnet_forcing, dtype=np.float64, copy=True)
    if swdown is not None:
        arrays["Qsw"] = -np.asarray(swdown, dtype=np.float64)
    NX, NY = arrays["uIce"].shape
    for k in STEP_OUTPUTS:  # fields the step writes (zero until it does)
        arrays.setdefault(k, np.zeros((NX, NY)))
    for k, a in arrays.items():
        arrays[k] = np.ascontiguousarray(np.broadcast_to(a, (NX, NY))).copy()
    sett = {k: float(s[k]) for k in s.keys() if isinstance(s[k], (bool, int, float))}
    # the oracle's module constants of solve4temp and the McPhee taper (growth.py:258-259)
    sett.update(AA1=O.AA1, AA2=O.AA2, BB1=O.BB1, PPASCALS=float(O.PPASCALS), CC1=O.CC1,
                CC2=O.CC2, McPheeT0=0.4, McPheeT1=7 / 0.4)
    if strength is not None:
        arrays["SeaIceStrength"] = np.array(strength, dtype=np.float64, copy=True)
        sett["keep_strength"] = 1.0
    n = int(s.nEVPsteps if n_substeps is None else n_substeps)
    fnames = list(arrays)
    dp = ctypes.POINTER(ctypes.c_double)
    fn = (ctypes.c_char_p * len(fnames))(*[k.encode() for k in fnames])
    fp = (dp * len(fnames))(*[arrays[k].ctypes.data_as(dp) for k in fnames])
    snames = list(sett)
    sn = (ctypes.c_char_p * len(snames))(*[k.encode() for k in snames])
    sv = np.array([sett[k] for k in snames], dtype=np.float64)
    rc = lib.seaice_step_ref(NX, NY, len(fnames), fn, fp, len(snames), sn,
                             sv.ctypes.data_as(dp), n)
    if rc == -2:
        raise KeyError(f"seaice_step_ref: no field or setting {lib.step_ref_missing().decode()!r}")
    if rc != 0:
        raise RuntimeError(f"seaice_step_ref returned {rc}")
    return arrays


def bulkf_formula_lanl(state, uw, vw, ta, qa, tsf, grav=9.81):
    """The open-ocean bulk fluxes (flwupa, flha, fsha, df0dT, ust, vst, evp,
    ssq, dEvdT) of ``veris/heat_flux_MITgcm.py:5-123``, cell by cell."""
    lib = _load()
    s = state.settings
    ins = [np.ascontiguousarray(a, dtype=np.float64) for a in (uw, vw, ta, qa, tsf)]
    n = ins[0].size
    out = [np.empty(ins[0].shape) for _ in range(9)]
    dp = ctypes.POINTER(ctypes.c_double)
    outs = (dp * 9)(*[o.ctypes.data_as(dp) for o in out])
    aln = float(np.log(2.0 / 10.0))
    rdn = float(s.karman / np.log(10.0 / 0.0005))
    rc = lib.bulkf_ref(n, *[a.ctypes.data_as(dp) for a in ins], float(grav), float(s.karman),
                       float(s.zvir), float(s.latvap), float(s.gamma_blk), float(s.rhoAir),
                       float(s.cpdair), float(s.ocean_emissivity), float(s.stefBoltz), aln, rdn,
                       outs)
    if rc != 0:
        raise RuntimeError(f"bulkf_ref returned {rc}")
    return tuple(out)
