"""Cross-period mode checks (modisp_amd.modes, the reference's
check_mode_reliability interface, python/modisp/reliability.py) on
hand-built velocity arrays: each flag fires on the case built for it and on
no smooth neighbour."""
import numpy as np
import pytest

from modisp_amd.modes import ModeReliabilityFlag as F, check_mode_reliability


def smooth(B=2, K=3, T=10):
    per = np.geomspace(1., 2.5, T)                    # log step ~0.1
    v = 3. + .4 * np.arange(K)[None, :, None] + .05 * np.log(per)[None, None, :]
    return np.broadcast_to(v, (B, K, T)).copy(), per


def test_smooth_curves_pass_after_the_first_period():
    v, per = smooth()
    r = check_mode_reliability(v, per)
    assert np.all(r.flags[:, :, 0] == F.UNASSESSED)
    assert np.all(r.flags[:, :, 1:] == 0) and r.assessed[:, :, 1:].all() and not r.assessed[:, :, 0].any()
    assert list(r.previous_index) == [-1] + list(range(9)) and list(r.sorted_order) == list(range(10))
    assert r.summary()["passes_checks"] == 2 * 3 * 9


def test_swapped_modes_are_rank_ambiguous_and_jump():
    v, per = smooth()
    v[0, [0, 1], 5] = v[0, [1, 0], 5]
    r = check_mode_reliability(v, per)
    assert r.flags[0, 0, 5] & F.RANK_AMBIGUOUS and r.flags[0, 1, 5] & F.RANK_AMBIGUOUS   # not ascending
    assert r.flags[0, 0, 5] & F.STEP_JUMP                 # 0.4 / 3.4 > 0.1
    assert r.flags[1].max() <= F.UNASSESSED               # the other model is untouched


def test_a_missing_mode_changes_the_count_and_matching():
    v, per = smooth()
    v[1, 0, 4] = np.nan
    r = check_mode_reliability(v, per)
    assert r.flags[1, 0, 4] & F.NONFINITE
    assert np.all(r.flags[1, :, 4] & F.ROOT_COUNT_CHANGE) and np.all(r.flags[1, :, 5] & F.ROOT_COUNT_CHANGE)
    assert not r.flags[1, 1, 4] & F.RANK_AMBIGUOUS        # the remaining modes keep their ranks
    v[1, :, 4] = [v[1, 1, 3], v[1, 2, 3], np.nan]          # a lost fundamental: the others shift down a rank
    r = check_mode_reliability(v, per)
    assert r.flags[1, 0, 4] & F.RANK_AMBIGUOUS and r.flags[1, 1, 4] & F.RANK_AMBIGUOUS


def test_trend_mismatch_gap_close_and_duplicates():
    v, per = smooth(B=1, K=2, T=8)
    v[0, 1, 6] *= 1.08                                    # off the trend of periods 4 and 5 by 8%
    r = check_mode_reliability(v, per)
    assert r.flags[0, 1, 6] & F.PREDICTION_MISMATCH and not r.flags[0, 1, 6] & F.STEP_JUMP
    gap = np.array([1., 1.1, 2., 2.2])
    r = check_mode_reliability(np.full((1, 1, 4), 3.), gap)
    assert r.flags[0, 0, 2] & F.PERIOD_GAP and not r.assessed[0, 0, 2] and r.assessed[0, 0, 3]
    close = np.array([[[3., 3.], [3.02, 3.001]]])         # gap 0.02 > 0.005 x 3.02; 0.001 <= 0.005 x 3.001
    r = check_mode_reliability(close, [1., 1.1])
    assert not r.flags[0, 0, 0] & F.CLOSE_ROOT and r.flags[0, 0, 1] & F.CLOSE_ROOT and r.flags[0, 1, 1] & F.CLOSE_ROOT
    dup = np.array([[[3., 3.2, 3.]]])
    r = check_mode_reliability(dup, [1., 2., 1.])         # the repeated period agrees with itself
    assert not np.any(r.flags & F.DUPLICATE_MISMATCH)
    dup[0, 0, 2] = 3.1
    r = check_mode_reliability(dup, [1., 2., 1.])
    assert r.flags[0, 0, 0] & F.DUPLICATE_MISMATCH and r.flags[0, 0, 2] & F.DUPLICATE_MISMATCH
    assert r.previous_index[2] == 0 and list(r.sorted_order) == [0, 2, 1]


def test_trace_flags_mark_failures_and_uncertain_searches():
    v, per = smooth(B=1)
    tf = np.zeros((1, 10), np.uint32)
    tf[0, 3], tf[0, 6] = 4, 2048
    r = check_mode_reliability(v, per, trace_flags=tf)
    assert np.all(r.flags[0, :, 3] & F.NUMERICAL_FAILURE) and not r.assessed[0, :, 3].any()
    assert not r.assessed[0, :, 4].any()                  # its incoming edge starts at a failed period
    assert np.all(r.flags[0, :, 6] & F.SEARCH_UNCERTAIN)


def test_argument_validation():
    v, per = smooth()
    with pytest.raises(ValueError, match="expected velocity"):
        check_mode_reliability(v[0], per)
    with pytest.raises(ValueError, match="positive finite"):
        check_mode_reliability(-v, per)
    with pytest.raises(ValueError, match="max_relative_step"):
        check_mode_reliability(v, per, max_relative_step=0)
    with pytest.raises(ValueError, match="trace_flags"):
        check_mode_reliability(v, per, trace_flags=np.zeros((1, 1), np.uint32))
