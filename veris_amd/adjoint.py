"""Reverse mode of the EVP subcycle and its pre-EVP forcing on the GPU.

The reference obtains sensitivities by ``jax.grad`` through its JAX kernels
(veris/evp_solver.py:22-319 with the forcing of veris/dynamics_routines.py:5-18,
veris/area_mass.py:17-28 and veris/dynsolver.py:7-84).  Here the derivative
is the hand-written adjoint of ``libveris_hip.so``: the forward subcycle keeps
the state every ``ckpt_every`` substeps, and the backward pass recomputes each
segment from its checkpoint and runs the adjoint kernel backwards through it.

:func:`evp_subcycle` is differentiable (through ``torch.autograd``) with
respect to the ice-strength parameter P* (a scalar or a per-cell field), the
initial ice thickness hIceMean, the water drag coefficients (north, south) and
the air drag coefficients (north, south).  Everything else in the state is a
constant of the derivative.  Non-smooth points follow one convention:
d sqrt(x)/dx = 0 at x = 0 (no strain, zero relative velocity), and the
denominator guard and the drag regime switch contribute no derivative.
"""

from __future__ import annotations

import ctypes

import torch

from . import _native as N
from . import kernels as K

_DRAGS = ("waterIceDrag", "waterIceDrag_south", "airIceDrag", "airIceDrag_south")
_EVP_UNUSED = ("AreaW", "AreaS", "uNm1", "vNm1")
_PRE_OUT = ("SeaIceMassC", "SeaIceMassU", "SeaIceMassV", "AreaW", "AreaS", "WindForcingX",
            "WindForcingY", "SeaIceStrength")


def _scalars(values):
    """Host floats of 0-d tensors / numbers, with at most one device sync."""
    dev = [v for v in values if isinstance(v, torch.Tensor) and v.device.type != "cpu"]
    host = iter(torch.stack([v.detach().double().reshape(()) for v in dev]).tolist()) if dev else None
    out = []
    for v in values:
        if isinstance(v, torch.Tensor):
            out.append(next(host) if v.device.type != "cpu" else float(v))
        else:
            out.append(float(v))
    return out


def _forward_fields(state, tile, pstar, h, drags):
    """ABI arguments of the subcycle with the given parameters: a fresh Ctx of
    ``state`` whose hIceMean, drag settings and pre-EVP fields (computed here
    on the device) are the parameters', with the tensors behind them."""
    c = K.Ctx.of(state).clone()
    c.dom.tile = int(tile)
    dt, shape = c.dtype, c.shape
    for name, val in zip(_DRAGS, _scalars(drags)):
        setattr(c.sett, name, val)
    h = h.detach().to(dt).contiguous()
    c._check("hIceMean", h)
    c.fields.hIceMean = h.data_ptr()
    pstar_field = pstar.detach().to(device=c.device, dtype=dt).expand(shape).contiguous()
    pre = torch.empty((len(_PRE_OUT),) + shape, dtype=dt, device=c.device)
    mC, mU, mV, AW, AS, WX, WY, P = pre.unbind(0)
    N.call("vh_seaice_mass", *c.args(), mC.data_ptr(), mU.data_ptr(), mV.data_ptr(), c.stream)
    N.call("vh_area_ws", *c.args(), AW.data_ptr(), AS.data_ptr(), c.stream)
    for name, t in (("SeaIceMassC", mC), ("SeaIceMassU", mU), ("SeaIceMassV", mV),
                    ("AreaW", AW), ("AreaS", AS)):
        setattr(c.fields, name, t.data_ptr())
    N.call("vh_wind_forcing_xy", *c.args(), WX.data_ptr(), WY.data_ptr(), c.stream)
    N.call("vh_seaice_strength", *c.args(), pstar_field.data_ptr(), P.data_ptr(), c.stream)
    for name, t in (("WindForcingX", WX), ("WindForcingY", WY), ("SeaIceStrength", P)):
        setattr(c.fields, name, t.data_ptr())
    return c, h, pre, pstar_field


def pair_tile(state, *, pStar=None, hIceMean=None) -> int:
    """A two-substep tile shape for :func:`evp_subcycle` (``tile=``) when
    its forward and recompute may run two substeps per launch on these
    inputs, else 0: the default settings, no metric terms, and every field the
    launches read at neighbouring cells periodic in the filled overlaps
    (:func:`veris_amd.kernels.overlaps_periodic`, evaluated on the pre-EVP
    fields these parameters give).  Waits for the device."""
    vs, s = state.variables, state.settings
    if not K.default_settings(state):
        return 0
    pstar = _as_tensor(s.pStar if pStar is None else pStar)
    h = vs.hIceMean if hIceMean is None else hIceMean
    c, *keep = _forward_fields(state, 0, pstar, h, [float(s[k]) for k in _DRAGS])
    flag = torch.zeros((), dtype=torch.int32, device=c.device)
    N.call("vh_evp_overlaps_periodic", *c.args(), flag.data_ptr(), c.stream)
    if not flag.item():
        return 0
    shapes = N.evp_tile_shapes()
    want = (20, 32, 2) if c.shape[0] * c.shape[1] < 400_000 else (32, 32, 2)
    return next(t for t, sh in shapes.items() if sh == want)


class _EvpSubcycle(torch.autograd.Function):
    @staticmethod
    def forward(ctx, state, n, ckpt_every, tile, pstar, h, cwN, cwS, caN, caS):
        c, h, pre, pstar_field = _forward_fields(state, tile, pstar, h, [cwN, cwS, caN, caS])
        dt, shape = c.dtype, c.shape
        nbytes = ctypes.c_size_t(0)
        N.call("vh_evp_adjoint_workspace_bytes", ctypes.byref(c.dom), n, ckpt_every,
               ctypes.byref(nbytes))
        ws = torch.empty(max(int(nbytes.value), 256), dtype=torch.uint8, device=c.device)
        out = torch.empty((5,) + shape, dtype=dt, device=c.device).unbind(0)
        N.call("vh_evp_forward_checkpointed", *c.args(), n, ckpt_every,
               *(t.data_ptr() for t in out), ws.data_ptr(), ws.numel(), c.stream)
        # The ABI structs hold raw pointers into the state's tensors, the
        # pre-EVP fields and the workspace (its checkpoints).  Saving every one
        # of them for backward keeps their memory alive however the caller
        # reassigns the state's variables in between, makes autograd refuse a
        # backward after an in-place change of any of them, and frees them
        # after backward unless the graph is retained.
        ptrs = {t.data_ptr() for t in (ws, h, pre, pstar_field)}
        keep = [ws, h, pre, pstar_field]
        for name in c.field_names:
            t = state.variables.get(name)
            if t.data_ptr() not in ptrs:
                ptrs.add(t.data_ptr())
                keep.append(t)
        ctx.save_for_backward(*keep)
        ctx.c = c
        ctx.n, ctx.ckpt = n, ckpt_every
        ctx.meta = [(x.dtype, x.device, x.shape) for x in (pstar, cwN, cwS, caN, caS)]
        ctx.h_meta = (h.dtype, h.device)
        return tuple(out)

    @staticmethod
    def backward(ctx, gu, gv, gs1, gs2, gs12):
        # raises autograd's own error when the saved tensors were freed by an
        # earlier backward (retain_graph=False) or changed in place since forward
        ws, _h, _pre, pstar_field = ctx.saved_tensors[:4]
        c = ctx.c
        shape, dt = c.shape, c.dtype
        grads = [None if g is None else g.to(dt).contiguous() for g in (gu, gv, gs1, gs2, gs12)]
        ebuf = torch.empty((len(N.EVP_ADJOINT_NAMES),) + shape, dtype=dt, device=c.device)
        evp = dict(zip(N.EVP_ADJOINT_NAMES, ebuf.unbind(0)))
        # the AreaW/AreaS and uNm1/vNm1 adjoints feed no gradient returned here
        eo = N.EvpAdjointOut(**{"adj_" + k: (None if k in _EVP_UNUSED else v.data_ptr())
                                for k, v in evp.items()})
        N.call("vh_evp_adjoint", *c.args(), ctx.n, ctx.ckpt,
               *(None if g is None else g.data_ptr() for g in grads), ctypes.byref(eo),
               ws.data_ptr(), ws.numel(), c.stream)
        pbuf = torch.empty((len(N.PRE_EVP_ADJOINT_NAMES),) + shape, dtype=dt, device=c.device)
        pre = dict(zip(N.PRE_EVP_ADJOINT_NAMES, pbuf.unbind(0)))
        po = N.PreEvpAdjointOut(**{"adj_" + k: v.data_ptr() for k, v in pre.items()})
        N.call("vh_pre_evp_adjoint", *c.args(), pstar_field.data_ptr(), ctypes.byref(eo),
               ctypes.byref(po), c.stream)
        need = ctx.needs_input_grad  # (state, n, ckpt, tile, pstar, h, cwN, cwS, caN, caS)
        # sums of the per-cell parameter adjoints over the padded field, in FP64
        sums = torch.stack([pre["pStar"], evp["waterDrag_north"], evp["waterDrag_south"],
                            pre["airDrag_north"], pre["airDrag_south"]]).double().sum(dim=(1, 2))
        res = []
        for q, (x_dt, x_dev, x_shape) in enumerate(ctx.meta):
            if not need[4 if q == 0 else q + 5]:
                res.append(None)
            elif q == 0 and len(x_shape) == 2:
                res.append(pre["pStar"].to(device=x_dev, dtype=x_dt))
            else:
                res.append(sums[q].reshape(x_shape).to(device=x_dev, dtype=x_dt))
        g_h = pre["hIceMean"].to(device=ctx.h_meta[1], dtype=ctx.h_meta[0]) if need[5] else None
        return (None, None, None, None, res[0], g_h, res[1], res[2], res[3], res[4])


def _as_tensor(x):
    return x if isinstance(x, torch.Tensor) else torch.tensor(float(x), dtype=torch.float64)


def evp_subcycle(state, *, pStar=None, hIceMean=None, waterIceDrag=None,
                 waterIceDrag_south=None, airIceDrag=None, airIceDrag_south=None,
                 n_substeps: int | None = None, ckpt_every: int = 10, tile: int = 0):
    """Pre-EVP forcing + EVP subcycle -> (uIce, vIce, sigma1, sigma2, sigma12).

    The forward pass computes SeaIceMass, AreaWS, WindForcingXY and
    SeaIceStrength (veris/area_mass.py:5-28, veris/dynsolver.py:7-84,
    veris/dynamics_routines.py:5-18) from ``state`` with the given parameters,
    then ``n_substeps`` EVP substeps (veris/evp_solver.py:257-319), with the
    same results as the drop-in kernels.  Parameters left as None take the
    state's value and are constants; tensors with ``requires_grad`` receive
    gradients in ``backward``: ``pStar`` (0-d or per-cell), ``hIceMean``
    (field), the four drag coefficients (0-d).  Gradients come back on the
    parameter's device and dtype.  Every setting of the EVP subcycle is
    supported: fixed or adaptive relaxation; no-slip boundaries (with or
    without the second-order term) or free slip with lateral drag (land
    neighbours or coastline form factors); with or without basal drag (its
    thickness dependence included in the hIceMean gradient), the water-stress
    turning angle (sinWat, cosWat) and the metric terms (k1AtC, k2AtC, k1AtZ,
    k2AtZ); the pre-EVP part takes the sea-ice load in the surface tilt
    (useRealFreshWaterFlux) as a state input, as the reference's kernel reads it.

    ``tile`` (0: the library's choice) names the forward's workgroup shape.
    A two-substep shape (k = 2, e.g. from :func:`pair_tile`) runs the
    checkpointed forward and the recompute two substeps per launch; the
    caller states with it that the inputs are periodic in their filled
    overlaps (what :func:`pair_tile` checks), under which the gradients and
    velocities are those of one substep per launch bit for bit.  With tile 0
    the library does so by itself on domains without a filled overlap.
    """
    vs, s = state.variables, state.settings
    n = int(s.nEVPsteps if n_substeps is None else n_substeps)
    if n < 1 or ckpt_every < 1:
        raise ValueError("n_substeps and ckpt_every must be >= 1")
    args = (
        _as_tensor(s.pStar if pStar is None else pStar),
        vs.hIceMean if hIceMean is None else hIceMean,
        _as_tensor(s.waterIceDrag if waterIceDrag is None else waterIceDrag),
        _as_tensor(s.waterIceDrag_south if waterIceDrag_south is None else waterIceDrag_south),
        _as_tensor(s.airIceDrag if airIceDrag is None else airIceDrag),
        _as_tensor(s.airIceDrag_south if airIceDrag_south is None else airIceDrag_south),
    )
    if args[0].dim() not in (0, 2):
        raise ValueError("pStar must be 0-d or a field")
    return _EvpSubcycle.apply(state, n, int(ckpt_every), int(tile), *args)


class _EvpSubcycleBatched(torch.autograd.Function):
    """Forward and reverse of the subcycle of every member of an ensemble of
    equal members in one launch sequence each (the batched entry points)."""

    @staticmethod
    def forward(ctx, ens, n, ckpt_every, tile, pstar, h, cwN, cwS, caN, caS):
        from . import batched as B

        dev, dt, total = ens.device, ens.dtype, ens.total
        NX, NY = ens.shapes[0]
        m0 = ens.member_settings[0]
        dom = N.domain_struct(m0, K._dtype_code(ens.fields["iceMask"]), tile)
        sett = N.settings_struct(ens.settings)
        for name, val in zip(_DRAGS, _scalars([cwN, cwS, caN, caS])):
            setattr(sett, name, val)
        h = h.detach().to(device=dev, dtype=dt).contiguous()
        if h.shape != (total,):
            raise ValueError("hIceMean must be a packed field of the ensemble")
        pstar_field = pstar.detach().to(device=dev, dtype=dt).expand(total).contiguous()
        fields = ens._fields_struct()
        fields.hIceMean = h.data_ptr()
        pre = torch.empty((len(_PRE_OUT), total), dtype=dt, device=dev)
        po = N.PreEvpOut(*[t.data_ptr() for t in pre.unbind(0)])
        stream = torch.cuda.current_stream(dev).cuda_stream
        b = ens._batch(0)
        N.call("vh_pre_evp_batched", ctypes.byref(b), ctypes.byref(sett), ctypes.byref(fields),
               pstar_field.data_ptr(), ctypes.byref(po), stream)
        for name, t in zip(_PRE_OUT, pre.unbind(0)):
            setattr(fields, name, t.data_ptr())
        nbytes = ctypes.c_size_t(0)
        N.call("vh_evp_adjoint_batched_workspace_bytes", ctypes.byref(dom), ens.n, n, ckpt_every,
               ctypes.byref(nbytes))
        ws = torch.empty(max(int(nbytes.value), 256), dtype=torch.uint8, device=dev)
        out = torch.empty((5, total), dtype=dt, device=dev).unbind(0)
        N.call("vh_evp_forward_checkpointed_batched", ctypes.byref(dom), ens.n, ctypes.byref(sett),
               ctypes.byref(fields), n, ckpt_every, *(t.data_ptr() for t in out), ws.data_ptr(),
               ws.numel(), stream)
        # every tensor behind a pointer of the ABI structs stays alive until backward
        ctx.save_for_backward(ws, h, pre, pstar_field,
                              *[ens.fields[k] for k in B.VARIABLE_NAMES])
        ctx.abi = (dom, sett, fields)
        ctx.n, ctx.ckpt, ctx.nm, ctx.total = n, ckpt_every, ens.n, total
        ctx.meta = [(x.dtype, x.device, x.shape) for x in (pstar, cwN, cwS, caN, caS)]
        ctx.h_meta = (h.dtype, h.device)
        return tuple(out)

    @staticmethod
    def backward(ctx, gu, gv, gs1, gs2, gs12):
        ws, _h, _pre, pstar_field = ctx.saved_tensors[:4]
        dom, sett, fields = ctx.abi
        dev, dt, total = ws.device, pstar_field.dtype, ctx.total
        stream = torch.cuda.current_stream(dev).cuda_stream
        grads = [None if g is None else g.to(dt).contiguous() for g in (gu, gv, gs1, gs2, gs12)]
        ebuf = torch.empty((len(N.EVP_ADJOINT_NAMES), total), dtype=dt, device=dev)
        evp = dict(zip(N.EVP_ADJOINT_NAMES, ebuf.unbind(0)))
        # the AreaW/AreaS and uNm1/vNm1 adjoints feed no gradient returned here
        eo = N.EvpAdjointOut(**{"adj_" + k: (None if k in _EVP_UNUSED else v.data_ptr())
                                for k, v in evp.items()})
        N.call("vh_evp_adjoint_batched", ctypes.byref(dom), ctx.nm, ctypes.byref(sett),
               ctypes.byref(fields), ctx.n, ctx.ckpt,
               *(None if g is None else g.data_ptr() for g in grads), ctypes.byref(eo),
               ws.data_ptr(), ws.numel(), stream)
        pbuf = torch.empty((len(N.PRE_EVP_ADJOINT_NAMES), total), dtype=dt, device=dev)
        pre = dict(zip(N.PRE_EVP_ADJOINT_NAMES, pbuf.unbind(0)))
        po = N.PreEvpAdjointOut(**{"adj_" + k: v.data_ptr() for k, v in pre.items()})
        N.call("vh_pre_evp_adjoint_batched", ctypes.byref(dom), ctx.nm, ctypes.byref(sett),
               ctypes.byref(fields), pstar_field.data_ptr(), ctypes.byref(eo), ctypes.byref(po),
               stream)
        need = ctx.needs_input_grad  # (ens, n, ckpt, tile, pstar, h, cwN, cwS, caN, caS)
        sums = torch.stack([pre["pStar"], evp["waterDrag_north"], evp["waterDrag_south"],
                            pre["airDrag_north"], pre["airDrag_south"]]).double().sum(dim=1)
        res = []
        for q, (x_dt, x_dev, x_shape) in enumerate(ctx.meta):
            if not need[4 if q == 0 else q + 5]:
                res.append(None)
            elif q == 0 and len(x_shape) == 1:
                res.append(pre["pStar"].to(device=x_dev, dtype=x_dt))
            else:
                res.append(sums[q].reshape(x_shape).to(device=x_dev, dtype=x_dt))
        g_h = pre["hIceMean"].to(device=ctx.h_meta[1], dtype=ctx.h_meta[0]) if need[5] else None
        return (None, None, None, None, res[0], g_h, res[1], res[2], res[3], res[4])


def evp_subcycle_batched(ens, *, pStar=None, hIceMean=None, waterIceDrag=None,
                         waterIceDrag_south=None, airIceDrag=None, airIceDrag_south=None,
                         n_substeps: int | None = None, ckpt_every: int = 10, tile: int = 0):
    """:func:`evp_subcycle` of every member of an ensemble of equal members
    (a :class:`~veris_amd.batched.Ensemble` whose members share one padded
    shape, cyclicity and settings) -> packed (uIce, vIce, sigma1, sigma2,
    sigma12), each of the ensemble's packed length.

    Forward and reverse each run all members in one launch sequence; member
    by member the results and gradients equal :func:`evp_subcycle`'s bit for
    bit.  ``pStar`` is 0-d (shared) or a packed per-cell field, ``hIceMean``
    a packed field (default: the ensemble's), the four drag coefficients are
    0-d and shared by all members (their gradients sum over the members)."""
    if len(set(ens.shapes)) != 1 or len(set(ens.cyclic)) != 1:
        raise ValueError("evp_subcycle_batched needs members of one shape and cyclicity")
    s = ens.settings
    n = int(s.nEVPsteps if n_substeps is None else n_substeps)
    if n < 1 or ckpt_every < 1:
        raise ValueError("n_substeps and ckpt_every must be >= 1")
    args = (
        _as_tensor(s.pStar if pStar is None else pStar),
        ens.fields["hIceMean"] if hIceMean is None else hIceMean,
        _as_tensor(s.waterIceDrag if waterIceDrag is None else waterIceDrag),
        _as_tensor(s.waterIceDrag_south if waterIceDrag_south is None else waterIceDrag_south),
        _as_tensor(s.airIceDrag if airIceDrag is None else airIceDrag),
        _as_tensor(s.airIceDrag_south if airIceDrag_south is None else airIceDrag_south),
    )
    if args[0].dim() not in (0, 1):
        raise ValueError("pStar must be 0-d or a packed field")
    return _EvpSubcycleBatched.apply(ens, n, int(ckpt_every), int(tile), *args)


_ADV_IN = ("uIce", "vIce", "hIceMean", "hSnowMean", "Area")


class _Advect(torch.autograd.Function):
    @staticmethod
    def forward(ctx, state, u, v, h, hs, A):
        c = K.Ctx(state)
        keep = []
        for name, t in zip(_ADV_IN, (u, v, h, hs, A)):
            t = t.detach().to(c.dtype).contiguous()
            c._check(name, t)
            setattr(c.fields, name, t.data_ptr())
            keep.append(t)
        out = c.empty(3)
        ws = c.workspace("vh_advection_workspace_bytes")
        N.call("vh_advection", *c.args(), *(t.data_ptr() for t in out), ws.data_ptr(),
               ws.numel(), c.stream)
        # the reverse recomputes the forward fluxes from these inputs
        ctx.save_for_backward(*keep, *[state.variables.get(n) for n in c.field_names])
        ctx.c = c
        ctx.metas = [(x.dtype, x.device) for x in (u, v, h, hs, A)]
        return tuple(out)

    @staticmethod
    def backward(ctx, gh, ghs, gA):
        ctx.saved_tensors  # raises if freed by an earlier backward or changed in place
        c = ctx.c
        lo = [None if g is None else g.to(c.dtype).contiguous() for g in (gh, ghs, gA)]
        res = torch.empty((5,) + c.shape, dtype=c.dtype, device=c.device)
        lu, lv, lh, lhs, lA = res.unbind(0)
        ws = c.workspace("vh_advection_adjoint_workspace_bytes")
        N.call("vh_advection_adjoint", *c.args(), *(None if g is None else g.data_ptr() for g in lo),
               lh.data_ptr(), lhs.data_ptr(), lA.data_ptr(), lu.data_ptr(), lv.data_ptr(),
               ws.data_ptr(), ws.numel(), c.stream)
        grads = [t.to(dtype=dt, device=dev) if need else None
                 for t, (dt, dev), need in zip((lu, lv, lh, lhs, lA), ctx.metas,
                                               ctx.needs_input_grad[1:])]
        return (None, *grads)


def advect(state, uIce=None, vIce=None, hIceMean=None, hSnowMean=None, Area=None):
    """Advection (veris/advection.py:15-175) of hIceMean, hSnowMean and Area
    by the ice velocities -> (hIceMean, hSnowMean, Area), the drop-in's results,
    differentiable through ``torch.autograd`` with respect to all five inputs
    (each defaults to the state's variable).  The backward is the
    hand-written adjoint of the flux-limited scheme: with
    :func:`evp_subcycle` in front it gives the gradients of a thickness misfit
    with respect to P*, the drag coefficients and the initial thickness."""
    vs = state.variables
    args = [vs[n] if x is None else x for n, x in zip(_ADV_IN, (uIce, vIce, hIceMean, hSnowMean,
                                                                Area))]
    return _Advect.apply(state, *args)
