"""Ensembles of sea-ice domains of uneven sizes in one launch sequence.

Members (each a :class:`~veris_amd.state.VerisState` with its own grid, masks,
ice and forcing) are packed back to back into one device array per variable;
the batched EVP entry point of the C ABI runs every member's subcycle in the
same launches (one member table, tiles of all members in one grid), replayed
from a hipGraph.  With float32 the bulk is FP32 while Delta, zeta and the
replacement pressure stay FP64; budgets are summed in FP64 on the device.

The reference has no batching; each member follows ``evp_solver``
(veris/evp_solver.py:257-319) exactly as the single-domain drop-in does.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native as N
from . import kernels as K
from .state import VerisState
from .variables import VARIABLE_NAMES

_KTERMS = ("k1AtC", "k2AtC", "k1AtZ", "k2AtZ")
_PRE_EVP = ("SeaIceMassC", "SeaIceMassU", "SeaIceMassV", "AreaW", "AreaS", "WindForcingX",
            "WindForcingY", "SeaIceStrength")
EVP_OUT = ("uIce", "vIce", "sigma1", "sigma2", "sigma12")


class Ensemble:
    """Packed ensemble on one device.

    ``states``: member states (NumPy or torch variables); their settings must
    agree except for ``nx``, ``ny`` and ``enable_cyclic_x``.  ``arith``
    chooses, for float32 fields, FP32 arithmetic with FP64 Delta, zeta and
    press ("native"), FP64 arithmetic on the FP32-stored fields ("fp64"), or
    FP32-stored step constants around an FP64 state with FP64 arithmetic
    ("fp64_state": every field is also held in FP64, ``fields64``; the
    pre-EVP forcing runs in FP64 on those and is rounded once into the FP32
    step constants; the EVP state uIce, vIce, sigma1, sigma2, sigma12 stays
    FP64, ``state64``, and the subcycle returns FP64).
    """

    _ARITH = {"native": 0, "fp64": 1, "fp64_state": 2}

    def __init__(self, states, device="cuda", dtype=torch.float32, arith="native"):
        if arith not in self._ARITH:
            raise ValueError("arith must be 'native', 'fp64' or 'fp64_state'")
        if arith != "native" and dtype != torch.float32:
            raise ValueError(f"arith={arith!r} applies to float32 ensembles")
        self.arith = arith
        if not states:
            raise ValueError("empty ensemble")
        self.device = torch.device(device)
        self.dtype = dtype
        self.settings = states[0].settings
        skip = {"nx", "ny", "enable_cyclic_x"}
        for st in states[1:]:
            for k in self.settings.keys():
                if k not in skip and st.settings[k] != self.settings[k]:
                    raise ValueError(f"members disagree on setting {k}")
        self.shapes = [tuple(st.shape) for st in states]
        sizes = [a * b for a, b in self.shapes]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        self.total = int(sum(sizes))
        self.n = len(states)
        self.cyclic = [bool(st.settings.enable_cyclic_x) for st in states]
        self.fill_mode = 0 if self.settings.veros_fill else 1
        self.fields = self._pack(states, dtype)
        self.fields64 = self._pack(states, torch.float64) if arith == "fp64_state" else {}
        self.state64 = {name: self.fields64[name] for name in EVP_OUT} if self.fields64 else {}
        self._members = (N.Member * self.n)()
        for m, (NX, NY) in enumerate(self.shapes):
            self._members[m] = N.Member(nx=NX - 4, ny=NY - 4, cyclic_x=int(self.cyclic[m]),
                                        fill_mode=self.fill_mode, offset=int(self.offsets[m]))
        self._ws = None
        self.member_settings = [st.settings for st in states]

    def _pack(self, states, dtype):
        fields = {}
        for name in VARIABLE_NAMES:
            buf = torch.empty(self.total, dtype=dtype, device=self.device)
            for m, st in enumerate(states):
                a = st.variables[name]
                t = a if isinstance(a, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(a))
                self.view(buf, m).copy_(t.to(device=self.device, dtype=dtype))
            fields[name] = buf
        return fields

    # -- views --------------------------------------------------------------

    def view(self, buf, m):
        """Member ``m``'s padded box of a packed 1-D buffer, as (NX, NY)."""
        o = int(self.offsets[m])
        NX, NY = self.shapes[m]
        return buf[o: o + NX * NY].view(NX, NY)

    def member_state(self, m) -> VerisState:
        """A state whose variables are views into the packed arrays."""
        vals = {name: self.view(buf, m) for name, buf in self.fields.items()}
        return VerisState(self.member_settings[m], vals)

    def unpack(self, name_or_buf):
        buf = self.fields[name_or_buf] if isinstance(name_or_buf, str) else name_or_buf
        return [self.view(buf, m) for m in range(self.n)]

    # -- kernels ------------------------------------------------------------

    def pre_evp(self, pStar: torch.Tensor | None = None):
        """SeaIceMass, AreaWS, WindForcingXY and SeaIceStrength of every member
        (veris/area_mass.py:5-28, veris/dynsolver.py:7-84,
        veris/dynamics_routines.py:5-18) in one batched launch, written into
        the packed arrays; ``pStar`` optionally a packed per-cell P* field.
        With arith "fp64_state" the forcing is computed in FP64 from the FP64
        fields and rounded once into the FP32 step constants."""
        fdt = torch.float64 if self.fields64 else self.dtype
        if pStar is not None and (pStar.shape != (self.total,) or pStar.dtype != fdt):
            raise ValueError("pStar must be a packed field of the ensemble's (forcing) dtype")
        fields = self.fields64 or self.fields
        if self.fields64:
            b = N.Batch(n_members=self.n, dtype=N.VH_F64, tile=0, arith=0, members=self._members,
                        total_elems=self.total)
        else:
            b = self._batch(0)
        out = N.PreEvpOut(*[fields[name].data_ptr() for name in _PRE_EVP])
        N.call("vh_pre_evp_batched", ctypes.byref(b), ctypes.byref(N.settings_struct(self.settings)),
               ctypes.byref(self._fields_struct(fields)), None if pStar is None else pStar.data_ptr(),
               ctypes.byref(out), torch.cuda.current_stream(self.device).cuda_stream)
        for name in _PRE_EVP if self.fields64 else ():
            self.fields[name].copy_(self.fields64[name])

    def _batch(self, tile):
        return N.Batch(n_members=self.n, dtype=N.VH_F64 if self.dtype == torch.float64 else N.VH_F32,
                       tile=int(tile), arith=self._ARITH[self.arith],
                       members=self._members, total_elems=self.total)

    def _fields_struct(self, fields=None, state64=False):
        """Packed field pointers (of ``fields``, default the ensemble's own);
        with ``state64`` the EVP state points at the FP64 state arrays (arith
        "fp64_state")."""
        fields = self.fields if fields is None else fields
        f = N.Fields()
        for name in VARIABLE_NAMES:
            setattr(f, name, fields[name].data_ptr())
        if state64:
            for name, buf in self.state64.items():
                setattr(f, name, buf.data_ptr())
        if all(K._all_zero(fields[k]) for k in _KTERMS):
            for k in _KTERMS:
                setattr(f, k, None)
        return f

    def evp(self, n_substeps=None, tile=0, out=None):
        """Batched EVP subcycle; returns the five packed outputs
        (uIce, vIce, sigma1, sigma2, sigma12), FP64 with arith "fp64_state"."""
        n = int(self.settings.nEVPsteps if n_substeps is None else n_substeps)
        b = self._batch(tile)
        nbytes = ctypes.c_size_t(0)
        N.call("vh_evp_batched_workspace_bytes", ctypes.byref(b), ctypes.byref(nbytes))
        if self._ws is None or self._ws.numel() < nbytes.value:
            self._ws = torch.empty(nbytes.value, dtype=torch.uint8, device=self.device)
        if out is None:
            sdt = torch.float64 if self.state64 else self.dtype
            out = [torch.empty(self.total, dtype=sdt, device=self.device) for _ in range(5)]
        sett = N.settings_struct(self.settings)
        fields = self._fields_struct(state64=True)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        N.call("vh_evp_solver_batched", ctypes.byref(b), ctypes.byref(sett), ctypes.byref(fields),
               n, *[t.data_ptr() for t in out], self._ws.data_ptr(), self._ws.numel(), stream)
        return tuple(out)

    def interior_cells(self) -> int:
        return int(sum((a - 4) * (b - 4) for a, b in self.shapes))

    def budgets(self, names=("hIceMean", "hSnowMean", "Area", "Qnet")):
        """FP64 budgets of every member in one launch: the sums of rA * field
        over each member's interior (ice and snow volume, ice area, and the
        heat budget of the net heat flux Qnet), accumulated in FP64 whatever
        the storage type; returns {name: float64 tensor of shape (n_members,)}."""
        if not 1 <= len(names) <= 8:
            raise ValueError("1 to 8 budget fields per call")
        res = torch.empty((len(names), self.n), dtype=torch.float64, device=self.device)
        ptrs = (ctypes.c_void_p * len(names))(*[self.fields[n].data_ptr() for n in names])
        b = self._batch(0)
        N.call("vh_global_sums_batched", ctypes.byref(b), len(names), ptrs,
               self.fields["rA"].data_ptr(), res.data_ptr(),
               torch.cuda.current_stream(self.device).cuda_stream)
        return dict(zip(names, res.unbind(0)))
