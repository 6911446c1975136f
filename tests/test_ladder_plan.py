"""Host logic of the precision-ladder planner (CPU): split points against the
error model |dE| = kappa * (tail sum of the vector weights)."""
import numpy as np

from openmolcas_amd import ladder


def weights(m=400):
    # pivot-ordered weights decaying over 12 orders of magnitude
    return 10.0 ** (-12.0 * np.arange(m) / (m - 1))


def test_tail_sums():
    w = np.array([4.0, 2.0, 1.0])
    assert np.array_equal(ladder.tail_sums(w), [7.0, 3.0, 1.0, 0.0])


def test_segments_within_budget():
    w, tol, share = weights(), 1e-6, 0.25
    k32, kb3 = 1e-7, 1e-4
    p = ladder.plan(w, k32, kb3, tol, share)
    t = ladder.tail_sums(w)
    assert 0 <= p.j_fp32 <= p.j_bf16 <= w.size
    assert kb3 * t[p.j_bf16] <= share * tol
    assert k32 * (t[p.j_fp32] - t[p.j_bf16]) <= share * tol
    # one vector earlier would break a budget (the split points are the smallest)
    if p.j_bf16 > 0:
        assert kb3 * t[p.j_bf16 - 1] > share * tol or p.j_bf16 == p.j_fp32
    if p.j_fp32 > 0:
        assert k32 * (t[p.j_fp32 - 1] - t[p.j_bf16]) > share * tol
    assert p.predicted <= 2 * share * tol


def test_no_bf16_segment_when_excluded():
    w = weights()
    p = ladder.plan(w, 1e-7, float("inf"), 1e-6, 0.25)
    assert p.j_bf16 == w.size
    assert np.isfinite(p.predicted)


def test_tighter_tolerance_keeps_more_fp64():
    w = weights()
    loose = ladder.plan(w, 1e-7, 1e-4, 1e-5, 0.25)
    tight = ladder.plan(w, 1e-7, 1e-4, 1e-8, 0.25)
    assert tight.j_fp32 >= loose.j_fp32 and tight.j_bf16 >= loose.j_bf16


def test_speedup_bound_closed_forms():
    L = ladder.Ladder
    all64 = L(100, 100, 0.0, nchol=100)
    assert ladder.speedup_bound(all64) == 1.0
    # a quarter of the vectors in FP64, reduced segments free: the cap is 4
    q = L(25, 60, 0.0, nchol=100)
    assert abs(ladder.speedup_bound(q) - 4.0) < 1e-12
    # reduced modes as fast as FP64: no gain
    assert abs(ladder.speedup_bound(q, rate_fp32=1.0, rate_bf16=1.0) - 1.0) < 1e-12
    # 0.25 + 0.35/2 + 0.40/4 = 0.525
    assert abs(ladder.speedup_bound(q, rate_fp32=2.0, rate_bf16=4.0) - 1.0 / 0.525) < 1e-12
    # batch: shares weighted by per-vector work; the heavy system dominates
    a, b = L(10, 10, 0.0, nchol=10), L(0, 0, 0.0, nchol=10)  # all FP64 / all split-BF16
    assert abs(ladder.speedup_bound([a, b], work=[3.0, 1.0]) - 4.0 / 3.0) < 1e-12


def test_plan_records_nchol():
    w = weights(300)
    p = ladder.plan(w, 1e-7, 1e-4)
    assert p.nchol == 300
    assert 0 < p.j_fp32 < 300
    assert abs(ladder.speedup_bound(p) - 300 / p.j_fp32) < 1e-12 * 300
