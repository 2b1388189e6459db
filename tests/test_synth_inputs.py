"""Seeded inputs of the benchmark workloads (CPU): shapes, ranges and
determinism of the adjoint inputs, the ragged ensemble sizes and the
0.25-degree adjoint members."""

from __future__ import annotations

import numpy as np

from veris_amd import synth


def test_adjoint_inputs_are_seeded_and_in_range():
    st = synth.config3()
    assert int(st.settings.nEVPsteps) == 120 and st.shape == (364, 164)
    p1, u1, v1 = synth.config3_adjoint_inputs(st)
    p2, u2, v2 = synth.config3_adjoint_inputs(st)
    assert np.array_equal(p1, p2) and np.array_equal(u1, u2) and np.array_equal(v1, v2)
    assert p1.shape == st.shape and p1.min() > 0
    med = np.median(p1[2:-2, 2:-2])
    assert 0.9 * 27500 < med < 1.1 * 27500  # lognormal with median 27500
    assert np.array_equal(p1[:2], p1[-4:-2]) and np.array_equal(p1[-2:], p1[2:4])  # x-cyclic
    assert np.all(u1[np.asarray(st.variables.maskInU) == 0] == 0)
    assert np.all(v1[np.asarray(st.variables.maskInV) == 0] == 0)
    assert 0.03 < u1[2:-2, 2:-2].std() < 0.15


def test_ensemble_and_adjoint_member_sizes():
    sizes = synth.config2_sizes()
    assert len(sizes) == 48
    assert all(64 <= a <= 512 and 64 <= b <= 384 for a, b in sizes)
    assert sum(a * b for a, b in sizes) == 3_202_380  # interior cells
    m = synth.config4_adjoint_member(3)
    assert m.shape == (244, 244) and not m.settings.enable_cyclic_x
    assert int(m.settings.nEVPsteps) == 120


def test_quarter_degree_adjoint_members_carry_smooth_thickness_noise():
    """The 240 x 240 members draw thickness from noise with a 10-cell
    correlation length: neighbouring ice cells are strongly correlated, where
    the uniform 0-3 m draw of the other workloads has no correlation."""
    def neighbour_corr(st):
        h = np.asarray(st.variables.hIceMean)[2:-2, 2:-2]
        ice = (h[1:] > 0) & (h[:-1] > 0)
        return float(np.corrcoef(h[1:][ice], h[:-1][ice])[0, 1])

    m = synth.config4_adjoint_member(0)
    h = np.asarray(m.variables.hIceMean)[2:-2, 2:-2]
    assert 0.0 <= h.min() and h.max() <= 3.0
    assert neighbour_corr(m) > 0.9
    assert abs(neighbour_corr(synth.regional(4, 100, 240, 240))) < 0.1
