"""The per-device graph cache is bounded: past 256 graphs, a call with a new
launch sequence runs its launches directly on the caller's stream, with the
same bits, and graphs captured before the bound keep replaying.

It fills the process-wide cache, so its file name keeps it last in the run.
"""

from __future__ import annotations

import pytest

from gpu_helpers import assert_exact, g2n, pre_evp
from oracle import veris_np as O
from veris_amd import kernels as K
from veris_amd import synth

pytestmark = pytest.mark.gpu

EVP_NAMES = ("uIce", "vIce", "sigma1", "sigma2", "sigma12")


def test_graph_cache_bound_keeps_results(cuda):
    st = pre_evp(synth.tiny(12, 10, config=31))
    vs = st.variables
    vs.uIce, vs.vIce, vs.sigma1, vs.sigma2, vs.sigma12 = O.evp_solver(st, n_substeps=5)
    st = pre_evp(st)
    g = st.to(cuda)
    early = [g2n(t) for t in K.evp_solver(g, n_substeps=9)]  # captured below the bound
    # every substep count from 8 up is a graph of its own: 300 more keys pass the bound
    for n in range(8, 308):
        K.evp_solver(g, n_substeps=n)
    for n in (9, 311):  # a cached graph, then a new key past the bound
        want = O.evp_solver(st, n_substeps=n)
        got = [g2n(t) for t in K.evp_solver(g, n_substeps=n)]
        for name, a, b in zip(EVP_NAMES, got, want):
            assert_exact(f"{name}[n={n}]", a, b)
    for name, a, b in zip(EVP_NAMES, early, O.evp_solver(st, n_substeps=9)):
        assert_exact(f"{name}[first call]", a, b)
