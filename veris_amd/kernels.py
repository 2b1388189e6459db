"""Drop-in replacements of the reference's sea-ice kernels on the GPU.

Every function takes the Veros-style ``state`` (``state.variables`` holds
device tensors in the reference layout, ``state.settings`` the scalars) and
returns fresh tensors in the reference's return order, exactly like the
``@veros_kernel`` functions it replaces; the caller assigns the results.  All
arithmetic runs in the HIP kernels of ``libveris_hip.so`` on the tensors'
device; there is no CPU path.
"""

from __future__ import annotations

import ctypes
import weakref

import torch

from . import _native as N
from .variables import VARIABLE_NAMES

_KTERMS = ("k1AtC", "k2AtC", "k1AtZ", "k2AtZ")
_ws_cache: dict = {}
_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def current_stream(index: int) -> int:
    """The handle of torch's current HIP stream on device ``index`` (the
    cheap raw query where this torch build has it)."""
    if _raw_stream is not None:
        return _raw_stream(index)
    return torch.cuda.current_stream(index).cuda_stream


def _dtype_code(t: torch.Tensor) -> int:
    if t.dtype == torch.float64:
        return N.VH_F64
    if t.dtype == torch.float32:
        return N.VH_F32
    raise TypeError(f"unsupported dtype {t.dtype}; use float64 or float32")


def _all_zero(t: torch.Tensor) -> bool:
    """Whether ``t`` is all zeros, remembered on the tensor object that owns
    the storage (a view's base: views share its version counter) for the
    current version and the view's extent (an address-keyed cache would be
    reused by a new tensor that the allocator places at the same address)."""
    owner = t._base if t._base is not None else t
    key = (t.storage_offset(), tuple(t.shape), tuple(t.stride()))
    tags = getattr(owner, "_vh_all_zero", None)
    if tags is None or tags[0] != t._version:
        tags = (t._version, {})
        owner._vh_all_zero = tags
    res = tags[1].get(key)
    if res is None:
        res = bool(torch.count_nonzero(t).item() == 0)
        tags[1][key] = res
    return res


class Ctx:
    """Per-call ABI arguments derived from a state."""

    def __init__(self, state, ref: torch.Tensor | None = None):
        vs, sett = state.variables, state.settings
        ref = vs.iceMask if ref is None else ref
        if not isinstance(ref, torch.Tensor) or ref.device.type != "cuda":
            raise TypeError("state variables must be torch tensors on a GPU device")
        self.device = ref.device
        self.index = ref.get_device()
        self.dtype = ref.dtype
        self.shape = (int(sett.nx) + 2 * int(sett.olx), int(sett.ny) + 2 * int(sett.oly))
        self.field_bytes = self.shape[0] * self.shape[1] * ref.element_size()
        self.dom = N.domain_struct(sett, _dtype_code(ref))
        self.sett = N.settings_struct(sett)
        self.fields = N.Fields()
        self.field_names = []  # variables whose device pointers are in self.fields
        self._src = {}  # name -> weak reference to the tensor behind its pointer
        for name in VARIABLE_NAMES:
            t = vs.get(name)
            if t is None:
                continue
            self._check(name, t)
            setattr(self.fields, name, t.data_ptr())
            self.field_names.append(name)
            self._src[name] = weakref.ref(t)
        self._kterms(vs)
        self.stream = current_stream(self.index)

    def _kterms(self, vs):
        """Metric-term fields that are absent or all zero are passed as NULL
        (the kernels then skip them)."""
        if all(vs.get(k) is None or _all_zero(vs.get(k)) for k in _KTERMS):
            for k in _KTERMS:
                setattr(self.fields, k, None)
        else:
            for k in _KTERMS:
                setattr(self.fields, k, vs.get(k).data_ptr())

    def _refresh(self, vs) -> bool:
        """Re-point the fields of the variables that were reassigned since
        this Ctx was built (each new tensor checked as in the constructor);
        False if the set of variables changed (then rebuild)."""
        names = []
        for name in VARIABLE_NAMES:
            t = vs.get(name)
            if t is None:
                continue
            names.append(name)
            r = self._src.get(name)
            if r is not None and r() is t:
                continue
            self._check(name, t)
            setattr(self.fields, name, t.data_ptr())
            self._src[name] = weakref.ref(t)
        if names != self.field_names:
            return False
        self._kterms(vs)
        return True

    @classmethod
    def of(cls, state, ref: torch.Tensor | None = None) -> "Ctx":
        """The Ctx of ``state``, built once and reused by later calls on the
        same stream while no variable or setting of the state has been
        assigned since (the namespaces count assignments); in-place changes of
        the metric-term fields, which decide whether they are read at all, are
        noticed through their version counters.  Callers must not modify the
        returned structs except ``dom.tile``, which every EVP call sets."""
        vs, sett = state.variables, state.settings
        t = vs.get("iceMask") if ref is None else ref
        vver, sver = getattr(vs, "_version", None), getattr(sett, "_version", None)
        if not isinstance(t, torch.Tensor) or not t.is_cuda or vver is None or sver is None:
            return cls(state, ref)
        get = vs.get
        kv = tuple([None if x is None else (id(x), x._version) for x in
                    (get("k1AtC"), get("k2AtC"), get("k1AtZ"), get("k2AtZ"))])
        index = t.get_device()
        slot = (t.dtype, index, current_stream(index))
        cache = vs.__dict__.setdefault("_ctx_cache", {})
        hit = cache.get(slot)
        if hit is not None and hit[1] is sett and hit[0][1] == sver:
            if hit[0] == (vver, sver, kv):
                return hit[2]
            c = hit[2]
            if c._refresh(vs):  # only some variables were reassigned
                cache[slot] = ((vver, sver, kv), sett, c)
                return c
        c = cls(state, ref)
        cache[slot] = ((vver, sver, kv), sett, c)
        return c

    def clone(self) -> "Ctx":
        """A copy with ABI structs of its own, which the caller may modify
        (the cached Ctx of :meth:`of` must stay as it is)."""
        c = object.__new__(Ctx)
        c.__dict__.update(self.__dict__)
        c.dom = type(self.dom).from_buffer_copy(self.dom)
        c.sett = type(self.sett).from_buffer_copy(self.sett)
        c.fields = type(self.fields).from_buffer_copy(self.fields)
        c.field_names = list(self.field_names)
        c._src = dict(self._src)
        return c

    def _check(self, name, t):
        if (isinstance(t, torch.Tensor) and t.dtype is self.dtype and t.shape == self.shape
                and t.get_device() == self.index and t.is_contiguous()):
            return
        self._check_failed(name, t)

    def _check_failed(self, name, t):
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"variable {name} is not a torch tensor")
        if t.device != self.device or t.dtype != self.dtype:
            raise TypeError(f"variable {name}: {t.dtype} on {t.device}, expected "
                            f"{self.dtype} on {self.device}")
        raise ValueError(f"variable {name}: shape {tuple(t.shape)} (contiguous="
                         f"{t.is_contiguous()}), expected contiguous {self.shape}")

    def empty(self, n=1):
        """One output field, or a list of n (views of one allocation)."""
        if n == 1:
            return torch.empty(self.shape, dtype=self.dtype, device=self.device)
        return list(torch.empty((n,) + self.shape, dtype=self.dtype, device=self.device).unbind(0))

    def empty_addressed(self, n):
        """n output fields (views of one allocation) and their device
        addresses."""
        block = torch.empty((n,) + self.shape, dtype=self.dtype, device=self.device)
        p0, nb = block.data_ptr(), self.field_bytes
        return block.unbind(0), range(p0, p0 + n * nb, nb)

    def workspace(self, query: str) -> torch.Tensor:
        nbytes = N.workspace_bytes(query, self.dom)
        key = (query, self.device, self.stream, self.shape, self.dtype)
        ws = _ws_cache.get(key)
        if ws is None or ws.numel() < nbytes:
            ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=self.device)
            _ws_cache[key] = ws
        return ws

    def args(self):
        return ctypes.byref(self.dom), ctypes.byref(self.sett), ctypes.byref(self.fields)


def _p(t):
    return None if t is None else t.data_ptr()


# --------------------------------------------------------------------------


def SeaIceMass(state):
    """veris/area_mass.py:17-28 -> (SeaIceMassC, SeaIceMassU, SeaIceMassV)."""
    c = Ctx.of(state)
    out = c.empty(3)
    N.call("vh_seaice_mass", *c.args(), *map(_p, out), c.stream)
    return tuple(out)


def AreaWS(state):
    """veris/area_mass.py:5-14 -> (AreaW, AreaS)."""
    c = Ctx.of(state)
    out = c.empty(2)
    N.call("vh_area_ws", *c.args(), *map(_p, out), c.stream)
    return tuple(out)


def SeaIceStrength(state, pStar: torch.Tensor | None = None):
    """veris/dynamics_routines.py:5-18 -> SeaIceStrength.

    ``pStar`` optionally replaces the scalar setting by a per-cell field."""
    c = Ctx.of(state)
    if pStar is not None:
        c._check("pStar", pStar)
    out = c.empty()
    N.call("vh_seaice_strength", *c.args(), _p(pStar), _p(out), c.stream)
    return out


def WindForcingXY(state):
    """veris/dynsolver.py:50-84 (with tauXY :7-47) -> (WindForcingX, WindForcingY)."""
    c = Ctx.of(state)
    out = c.empty(2)
    N.call("vh_wind_forcing_xy", *c.args(), *map(_p, out), c.stream)
    return tuple(out)


def pre_evp(state, pStar: torch.Tensor | None = None):
    """SeaIceMass, AreaWS, WindForcingXY and SeaIceStrength in one fused launch
    (veris/area_mass.py:5-28, veris/dynsolver.py:7-84,
    veris/dynamics_routines.py:5-18) -> (SeaIceMassC, SeaIceMassU, SeaIceMassV,
    AreaW, AreaS, WindForcingX, WindForcingY, SeaIceStrength); results equal
    the four separate drop-ins, in that order of the reference setup."""
    c = Ctx.of(state)
    if pStar is not None:
        c._check("pStar", pStar)
    buf = torch.empty((len(N.PRE_EVP_NAMES),) + c.shape, dtype=c.dtype, device=c.device)
    out = buf.unbind(0)
    o = N.PreEvpOut(*[t.data_ptr() for t in out])
    N.call("vh_pre_evp", *c.args(), _p(pStar), ctypes.byref(o), c.stream)
    return tuple(out)


# EVP tile shape index used when a call does not name one (0 = library default),
# per evp_tile_key; set by veris_amd.autotune.
EVP_TILE = {}


def default_settings(state) -> bool:
    """Whether ``state``'s settings take none of the EVP subcycle's optional
    branches (no adaptive EVP, no-slip without the second-order term, no
    turning angle, basal drag or metric terms, tensileStrFac = 0): those the
    library runs on its default-settings kernels."""
    vs, s = state.variables, state.settings
    if any(vs.get(k) is not None and not _all_zero(vs.get(k)) for k in _KTERMS):
        return False
    return (not s.useAdaptiveEVP and bool(s.noSlip) and not s.secondOrderBC
            and float(s.sinWat) == 0.0 and float(s.basalDragK2) == 0.0
            and float(s.tensileStrFac) == 0.0)


def evp_tile_key(state) -> tuple:
    """Key of :data:`EVP_TILE`: the padded grid, the dtype and whether the
    settings are the default ones (the best shape of the default-settings
    kernels says nothing about the general ones, whose multi-substep shapes
    hold more registers)."""
    t = state.variables.iceMask
    return (tuple(t.shape), t.dtype, default_settings(state))


def overlaps_periodic(state) -> bool:
    """Whether multi-substep EVP tile shapes give the single-substep results
    on ``state`` (vh_evp_overlaps_periodic): every field they read at
    neighbouring cells holds, in each filled overlap, the value the fill
    copies there.  True for every domain without a filled overlap.  Waits for
    the device."""
    c = Ctx.of(state)
    flag = torch.zeros((), dtype=torch.int32, device=c.device)
    N.call("vh_evp_overlaps_periodic", *c.args(), flag.data_ptr(), c.stream)
    return bool(flag.item())


def single_substep_ok(settings, tile: int) -> bool:
    """False when ``settings`` need single-substep shapes (the second-order
    no-slip boundary term reads two cells beyond the face) and ``tile`` is a
    multi-substep shape."""
    if not (settings.noSlip and settings.secondOrderBC):
        return True
    return N.evp_tile_shapes()[int(tile)][2] == 1


def evp_solver(state, n_substeps: int | None = None, tile: int | None = None):
    """veris/evp_solver.py:257-319 -> (uIce, vIce, sigma1, sigma2, sigma12).

    ``tile`` selects the workgroup shape (see :func:`veris_amd.autotune.tune_evp`)."""
    c = Ctx.of(state)
    if tile is None:
        tile = EVP_TILE.get(evp_tile_key(state), 0)
        if tile and not single_substep_ok(state.settings, tile):
            tile = 0  # a tuned multi-substep shape cannot run the second-order boundary term
    c.dom.tile = int(tile)
    n = int(state.settings.nEVPsteps if n_substeps is None else n_substeps)
    out = c.empty(5)
    ws = c.workspace("vh_evp_workspace_bytes")
    N.call("vh_evp_solver", *c.args(), n, *map(_p, out), _p(ws), ws.numel(), c.stream)
    return tuple(out)


def evp_solver_residuals(state, n_substeps: int | None = None):
    """veris/evp_solver.py:257-319 with computeEvpResidual ->
    (uIce, vIce, sigma1, sigma2, sigma12, resU, resSig): the same subcycle as
    :func:`evp_solver` (bit for bit) plus the per-substep residual norms as
    FP64 device tensors of length n_substeps."""
    c = Ctx.of(state)
    n = int(state.settings.nEVPsteps if n_substeps is None else n_substeps)
    out = c.empty(5)
    res = torch.empty((2, n), dtype=torch.float64, device=c.device)
    nbytes = ctypes.c_size_t(0)
    N.call("vh_evp_residuals_workspace_bytes", ctypes.byref(c.dom), n, ctypes.byref(nbytes))
    ws = torch.empty(max(nbytes.value, 256), dtype=torch.uint8, device=c.device)
    N.call("vh_evp_solver_residuals", *c.args(), n, *map(_p, out), res[0].data_ptr(),
           res[1].data_ptr(), _p(ws), ws.numel(), c.stream)
    return (*out, res[0], res[1])


def freedrift_solver(state):
    """veris/freedrift_solver.py:5-62 -> (uIce, vIce): steady free drift."""
    c = Ctx.of(state)
    out = c.empty(2)
    N.call("vh_freedrift_solver", *c.args(), *map(_p, out), c.stream)
    return tuple(out)


def IceVelocities(state, n_substeps: int | None = None):
    """veris/dynsolver.py:87-102 -> (uIce, vIce, sigma1, sigma2, sigma12).

    EVP when useEVP is set (it overrides free drift, as in the reference);
    otherwise free drift with the state's stresses unchanged (the reference's
    free-drift branch reads them from an undefined name)."""
    s = state.settings
    if s.useEVP:
        return evp_solver(state, n_substeps)
    if s.useFreedrift:
        vs = state.variables
        u, v = freedrift_solver(state)
        return u, v, vs.sigma1, vs.sigma2, vs.sigma12
    raise ValueError("IceVelocities: neither useEVP nor useFreedrift is set")


def OceanStressUV(state):
    """veris/ocean_stress.py:8-40 -> (OceanStressU, OceanStressV)."""
    c = Ctx.of(state)
    out = c.empty(2)
    N.call("vh_ocean_stress_uv", *c.args(), *map(_p, out), c.stream)
    return tuple(out)


def Advection(state):
    """veris/advection.py:15-25 -> (hIceMean, hSnowMean, Area)."""
    c = Ctx.of(state)
    out = c.empty(3)
    ws = c.workspace("vh_advection_workspace_bytes")
    N.call("vh_advection", *c.args(), *map(_p, out), _p(ws), ws.numel(), c.stream)
    return tuple(out)


def clean_up_advection(state):
    """veris/clean_up.py:5-39 ->
    (hIceMean, hSnowMean, Area, TSurf, os_hIceMean, os_hSnowMean)."""
    c = Ctx.of(state)
    out = c.empty(6)
    N.call("vh_clean_up_advection", *c.args(), *map(_p, out), c.stream)
    return tuple(out)


def ridging(state):
    """veris/clean_up.py:42-47 -> Area."""
    c = Ctx.of(state)
    out = c.empty()
    N.call("vh_ridging", *c.args(), _p(out), c.stream)
    return out


def Growth(state):
    """veris/growth.py:8-397 -> (hIceMean, hSnowMean, Area, TSurf, EmPmR,
    forc_salt_surface, Qsw, Qnet, SeaIceLoad, IcePenetSW, recip_hIceMean)."""
    c = Ctx.of(state)
    out = c.empty(len(N.GROWTH_OUT_NAMES))
    go = N.GrowthOut(*[t.data_ptr() for t in out])
    N.call("vh_growth", *c.args(), ctypes.byref(go), c.stream)
    return tuple(out)


def fill_overlap(state, a: torch.Tensor) -> torch.Tensor:
    """veris/fill_overlap.py:6-18 -> filled copy of ``a``."""
    c = Ctx.of(state, ref=a)
    c._check("field", a)
    out = a.clone()
    N.call("vh_fill_overlap", ctypes.byref(c.dom), _p(out), c.stream)
    return out


def fill_overlaps_(state, fields) -> None:
    """fill_overlap (veris/fill_overlap.py:6-18) of up to 8 distinct fields
    of the state's grid and dtype, each in place, in one launch: the fills
    that close a step, which the reference makes one call each
    (veris/setup/seaice_global_4deg/seaice_global_4deg.py:928-932), on
    fields the step has just produced and owns."""
    fields = list(fields)
    if not fields:
        return
    c = Ctx.of(state, ref=fields[0])
    for q, a in enumerate(fields):
        c._check(f"field {q}", a)
    ptrs = (ctypes.c_void_p * len(fields))(*[a.data_ptr() for a in fields])
    N.call("vh_fill_overlaps", ctypes.byref(c.dom), len(fields), ptrs, c.stream)


def fill_overlap_uv(state, u, v):
    """veris/fill_overlap.py:21-23."""
    return fill_overlap(state, u), fill_overlap(state, v)


def global_sum(state, a: torch.Tensor, weight: torch.Tensor | None = None) -> torch.Tensor:
    """FP64 sum of ``a`` (times ``weight``) over the interior cells, as a 0-d
    float64 device tensor.  The reference's ``global_sum`` (veris/global_sum.py:1-4)
    is an identity stub; this is the reduction it stands for."""
    c = Ctx.of(state, ref=a)
    c._check("field", a)
    if weight is not None:
        c._check("weight", weight)
    res = torch.empty((), dtype=torch.float64, device=c.device)
    ws = c.workspace("vh_global_sum_workspace_bytes")
    N.call("vh_global_sum_f64", ctypes.byref(c.dom), _p(a), _p(weight), _p(res), _p(ws),
           ws.numel(), c.stream)
    return res
