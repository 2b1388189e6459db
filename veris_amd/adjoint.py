"""Reverse mode of the EVP subcycle and its pre-EVP forcing on the GPU.

The reference obtains sensitivities by ``jax.grad`` through its JAX kernels
(veris/evp_solver.py:22-319 with the forcing of veris/dynamics_routines.py:5-18,
veris/area_mass.py:17-28 and veris/dynsolver.py:7-84).  Here the derivative
is the hand-written adjoint of ``libveris_hip.so``: the forward subcycle keeps
the state every ``ckpt_every`` substeps, and the backward pass recomputes each
segment from its checkpoint and runs the adjoint kernels backwards through it.

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
from .settings import SettingsNamespace
from .state import VerisState



def _with(state, settings: dict | None = None, variables: dict | None = None) -> VerisState:
    sv = state.settings.as_dict()
    sv.update(settings or {})
    vv = dict(state.variables.items())
    vv.update(variables or {})
    return VerisState(SettingsNamespace(sv), vv)


def _pre_evp(state, pstar_field):
    mC, mU, mV = K.SeaIceMass(state)
    st = _with(state, variables=dict(SeaIceMassC=mC, SeaIceMassU=mU, SeaIceMassV=mV))
    AW, AS = K.AreaWS(st)
    st = _with(st, variables=dict(AreaW=AW, AreaS=AS))
    WX, WY = K.WindForcingXY(st)
    P = K.SeaIceStrength(st, pStar=pstar_field)
    return _with(st, variables=dict(WindForcingX=WX, WindForcingY=WY, SeaIceStrength=P))


class _EvpSubcycle(torch.autograd.Function):
    @staticmethod
    def forward(ctx, state, n, ckpt_every, tile, pstar, h, cwN, cwS, caN, caS):
        vs = state.variables
        ref = vs.iceMask
        dt = ref.dtype
        pstar_field = pstar.detach().to(dt).expand(ref.shape).contiguous()
        scal = dict(waterIceDrag=float(cwN), waterIceDrag_south=float(cwS),
                    airIceDrag=float(caN), airIceDrag_south=float(caS))
        st = _with(state, settings=scal, variables=dict(hIceMean=h.detach().to(dt).contiguous()))
        st = _pre_evp(st, pstar_field)
        c = K.Ctx(st)
        c.dom.tile = int(tile)
        nbytes = ctypes.c_size_t(0)
        N.call("vh_evp_adjoint_workspace_bytes", ctypes.byref(c.dom), n, ckpt_every,
               ctypes.byref(nbytes))
        ws = torch.empty(max(int(nbytes.value), 256), dtype=torch.uint8, device=c.device)
        out = c.empty(5)
        N.call("vh_evp_forward_checkpointed", *c.args(), n, ckpt_every, *map(K._p, out),
               K._p(ws), ws.numel(), c.stream)
        ctx.st, ctx.n, ctx.ckpt, ctx.tile = st, n, ckpt_every, tile
        ctx.pstar_field, ctx.ws = pstar_field, ws
        ctx.in_dtypes = [x.dtype for x in (pstar, h, cwN, cwS, caN, caS)]
        ctx.pstar_shape = pstar.shape
        return tuple(out)

    @staticmethod
    def backward(ctx, gu, gv, gs1, gs2, gs12):
        st = ctx.st
        c = K.Ctx(st)
        c.dom.tile = int(ctx.tile)
        grads = [None if g is None else g.to(c.dtype).contiguous() for g in (gu, gv, gs1, gs2, gs12)]
        evp_out = {name: c.empty() for name in N.EVP_ADJOINT_NAMES}
        eo = N.EvpAdjointOut(**{"adj_" + k: K._p(v) for k, v in evp_out.items()})
        N.call("vh_evp_adjoint", *c.args(), ctx.n, ctx.ckpt, *map(K._p, grads), ctypes.byref(eo),
               K._p(ctx.ws), ctx.ws.numel(), c.stream)
        pre_out = {name: c.empty() for name in N.PRE_EVP_ADJOINT_NAMES}
        po = N.PreEvpAdjointOut(**{"adj_" + k: K._p(v) for k, v in pre_out.items()})
        N.call("vh_pre_evp_adjoint", *c.args(), K._p(ctx.pstar_field), ctypes.byref(eo),
               ctypes.byref(po), c.stream)
        dps, dh, dcwN, dcwS, dcaN, dcaS = ctx.in_dtypes
        g_pstar = pre_out["pStar"].double()
        if len(ctx.pstar_shape) == 0 or g_pstar.shape != ctx.pstar_shape:
            g_pstar = g_pstar.sum().reshape(ctx.pstar_shape)
        out = (
            g_pstar.to(dps),
            pre_out["hIceMean"].to(dh),
            evp_out["waterDrag_north"].double().sum().to(dcwN),
            evp_out["waterDrag_south"].double().sum().to(dcwS),
            pre_out["airDrag_north"].double().sum().to(dcaN),
            pre_out["airDrag_south"].double().sum().to(dcaS),
        )
        ctx.ws = None
        return (None, None, None, None) + out


def _as_tensor(x, device):
    if isinstance(x, torch.Tensor):
        return x
    return torch.tensor(float(x), dtype=torch.float64, device=device)


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
    gradients in ``backward``: ``pStar`` (0-d or per-cell, float64 or the
    state dtype), ``hIceMean`` (field), the four drag coefficients (0-d).
    Supported settings are those of the default configuration: non-adaptive
    noSlip EVP, no turning angles, no basal drag, no metric terms, no
    sea-ice load in the surface tilt.
    """
    vs, s = state.variables, state.settings
    dev = vs.iceMask.device
    n = int(s.nEVPsteps if n_substeps is None else n_substeps)
    if n < 1 or ckpt_every < 1:
        raise ValueError("n_substeps and ckpt_every must be >= 1")
    args = (
        _as_tensor(s.pStar if pStar is None else pStar, dev),
        vs.hIceMean if hIceMean is None else hIceMean,
        _as_tensor(s.waterIceDrag if waterIceDrag is None else waterIceDrag, dev),
        _as_tensor(s.waterIceDrag_south if waterIceDrag_south is None else waterIceDrag_south, dev),
        _as_tensor(s.airIceDrag if airIceDrag is None else airIceDrag, dev),
        _as_tensor(s.airIceDrag_south if airIceDrag_south is None else airIceDrag_south, dev),
    )
    return _EvpSubcycle.apply(state, n, int(ckpt_every), int(tile), *args)
