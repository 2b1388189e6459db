"""vh_evp_solver_noalloc (the EVP subcycle with its scratch in the library's
own per-GPU device array) against vh_evp_solver (caller's workspace) through
the C ABI: bit for bit, on two calls in a row, the second on a smaller
domain, then a third on a still smaller FP32 one; and the host-side refusal
of a domain beyond the scratch maximum."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from gpu_helpers import pre_evp
from oracle import veris_np as O
from veris_amd import _native as N
from veris_amd import kernels as K
from veris_amd import synth

pytestmark = pytest.mark.gpu


def _spun(st, n=5):
    pre_evp(st)
    vs = st.variables
    vs.uIce, vs.vIce, vs.sigma1, vs.sigma2, vs.sigma12 = O.evp_solver(st, n_substeps=n)
    return pre_evp(st)


def _both(g, n):
    c = K.Ctx(g)
    ws = torch.empty(N.workspace_bytes("vh_evp_workspace_bytes", c.dom), dtype=torch.uint8,
                     device=c.device)
    ref = c.empty(5)
    N.call("vh_evp_solver", *c.args(), n, *[t.data_ptr() for t in ref], ws.data_ptr(),
           ws.numel(), c.stream)
    got = c.empty(5)
    N.call("vh_evp_solver_noalloc", *c.args(), n, *[t.data_ptr() for t in got], c.stream)
    torch.cuda.synchronize()
    return ref, got


def test_noalloc_equals_workspace_entry_point(cuda):
    cases = [("1 degree FP64", _spun(synth.config1(), 3), torch.float64, 41),
             ("4 degree FP64", _spun(synth.config0()), torch.float64, 23),
             ("24x16 FP32", _spun(synth.tiny(24, 16)), torch.float32, 17)]
    for label, st, dt, n in cases:
        print(f"noalloc: {label}, {n} substeps", flush=True)
        ref, got = _both(st.to(cuda, dt), n)
        for q, (a, b) in enumerate(zip(got, ref)):
            assert torch.equal(a.view(torch.int64 if dt == torch.float64 else torch.int32),
                               b.view(torch.int64 if dt == torch.float64 else torch.int32)), \
                f"{label}: output {q} differs from vh_evp_solver"
            assert bool(torch.isfinite(a).all())
    # the 1-degree FP64 case also against the oracle (the shared kernels' bar)
    st = cases[0][1]
    want = O.evp_solver(st, n_substeps=41)
    _, got = _both(st.to(cuda), 41)
    for a, b in zip(got, want):
        assert np.array_equal(a.cpu().numpy(), b)


def test_noalloc_refuses_a_domain_beyond_its_scratch():
    dom = N.Domain(nx=2000, ny=1200, olx=2, oly=2, cyclic_x=1, fill_mode=0, dtype=N.VH_F64,
                   tile=0)
    sett, fields = N.Settings(), N.Fields()
    with pytest.raises(N.VerisHipError, match="VH_EWORKSPACE.*beyond its maximum"):
        N.call("vh_evp_solver_noalloc", ctypes.byref(dom), ctypes.byref(sett),
               ctypes.byref(fields), 3, None, None, None, None, None, None)
