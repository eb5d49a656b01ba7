"""Continuity screen (host, C ABI) and the front end's FP64 review.

CPU: the screen's flag rules on hand-built phase arrays (reference
cpp/cpu/continuity_screen.hpp:8-67): endpoints and wide period gaps have no
bilateral neighbours (8), a mode outside [0.95, 1.05] of its neighbours'
envelope or missing between them is an envelope suspect (16), FP32 phases of
adjacent modes within 16 ulp are close (32), a continuation review request
is a prediction suspect (64); unsorted periods are screened in ascending order.

GPU: a result with a planted error (one mode scaled by 1.1 at one period) is
caught by the screen and repaired by the review to the verifier's root, and an
accuracy="auto" solve carries continuity flags whose reviewed traces equal the
accuracy="high" result.
"""
import numpy as np
import pytest

import modisp_amd
from modisp_amd import api, workloads


def smooth(B=2, K=3, T=12):
    per = np.geomspace(1., 3., T)                      # log step 0.1
    base = 3. + np.arange(K)[None, :, None] * .5 + .02 * np.log(per)[None, None, :]
    return per, np.broadcast_to(base, (B, K, T)).copy()


def test_smooth_phases_flag_only_the_ends():
    per, ph = smooth()
    flags, suspect = modisp_amd.continuity_screen(per, ph, None, True)
    assert not suspect
    assert np.all(flags[:, 1:-1] == 0) and np.all(flags[:, [0, -1]] == 8)


def test_envelope_jump_and_missing_mode_are_suspects():
    per, ph = smooth()
    ph[0, 1, 5] *= 1.06                                # 6% above both neighbours
    ph[1, 2, 7] = np.nan                               # vanishes between finite neighbours
    flags, suspect = modisp_amd.continuity_screen(per, ph, None, True)
    assert suspect
    assert flags[0, 5] == 16 and flags[1, 7] == 16
    assert flags[0, 4] == 0 and flags[0, 6] == 0       # the jump stays inside its neighbours' envelope
    ph[0, 1, 5] /= 1.06 / 1.04                         # 4%: inside the envelope
    flags, _ = modisp_amd.continuity_screen(per, ph, None, True)
    assert flags[0, 5] == 0


def test_wide_period_gap_has_no_bilateral_neighbours():
    per = np.array([1., 1.1, 1.2, 2., 2.1])            # log(2/1.1) and log(2.1/1.2) exceed 0.25
    ph = np.full((1, 1, 5), 3.)
    ph[0, 0, 2] = 5.                                   # a jump that is not checked there
    flags, suspect = modisp_amd.continuity_screen(per, ph, None, True)
    assert not suspect
    assert list(flags[0]) == [8, 0, 8, 8, 8]


def test_unsorted_periods_are_screened_in_ascending_order():
    per, ph = smooth(B=1, K=2)
    ph[0, 0, 5] *= 1.1
    perm = np.random.default_rng(0).permutation(per.size)
    flags, _ = modisp_amd.continuity_screen(per, ph, None, True)
    flags_p, _ = modisp_amd.continuity_screen(per[perm], ph[:, :, perm], None, True)
    np.testing.assert_array_equal(flags_p, flags[:, perm])


def test_close_fp32_modes_and_prediction_review():
    per, ph = smooth(B=1, K=2)
    ph32 = ph.astype(np.float32)
    ph32[0, 1, 3] = np.nextafter(ph32[0, 0, 3], np.float32(10.))   # adjacent modes 1 ulp apart
    for i in range(15):
        ph32[0, 1, 3] = np.nextafter(ph32[0, 1, 3], np.float32(10.))  # 16 ulp
    pred = np.zeros((1, per.size), dtype=np.uint32)
    pred[0, 8] = 4096 | 4                                          # continuation asked for review
    flags, suspect = modisp_amd.continuity_screen(per, ph32, pred, True)
    assert suspect and flags[0, 3] & 32 and flags[0, 8] & 64
    ph32[0, 1, 3] = np.nextafter(ph32[0, 1, 3], np.float32(10.))   # 17 ulp: not close
    flags, _ = modisp_amd.continuity_screen(per, ph32, None, True)
    assert not flags[0, 3] & 32
    flags, suspect = modisp_amd.continuity_screen(per, ph32, pred, False)   # unilateral: only the review request
    assert suspect and np.all(flags[0, :8] == 0) and flags[0, 8] == 64


def test_screen_validates_shapes():
    with pytest.raises(ValueError):
        modisp_amd.continuity_screen(np.ones(3), np.ones((1, 2, 4)))
    with pytest.raises(ValueError):
        modisp_amd.continuity_screen(np.ones(3), np.ones((1, 2, 3)), np.zeros((2, 3), np.uint32))


@pytest.fixture(scope="module")
def gpu():
    if not modisp_amd.available(0):
        pytest.fail("no gfx950 device visible to the gpu tests")
    return modisp_amd


@pytest.mark.gpu
def test_review_repairs_a_planted_error(gpu):
    wl = workloads.precision_batch(B=6, seed=21)
    per = np.geomspace(1., 20., 40)                   # neighbours 0.15 apart in log period
    with api.Solver(accuracy="high", c_tol_rel=1e-10) as s:
        res = s(wl.thickness, wl.vp, wl.vs, wl.rho, per, 3)
    vel = res.phase_velocity.copy()
    assert np.isfinite(vel[2, 1, 10])
    vel[2, 1, 10] *= 1.1                               # the planted error
    flags, suspect = modisp_amd.continuity_screen(per, vel, None, True)
    assert suspect and flags[2, 10] & int(api.ContinuityFlag.ENVELOPE)
    stats = {}
    trace, levels, counts = res.trace_flags.copy(), res.search_level.copy(), res.root_count.copy()
    with api.Solver(accuracy="balanced", c_tol_rel=1e-10) as s:
        api._review(s, (wl.thickness, wl.vp, wl.vs, wl.rho), per, "rayleigh", res.search_bounds, vel, trace, levels,
                    counts, None, None, flags, dict(stats, continuity_local_searches=0, continuity_full_searches=0,
                                                     continuity_repaired=0, continuity_unresolved=0,
                                                     continuity_evaluations=0))
    assert flags[2, 10] & int(api.ContinuityFlag.REVIEWED) and flags[2, 10] & int(api.ContinuityFlag.REPAIRED)
    assert not flags[2, 10] & int(api.ContinuityFlag.UNRESOLVED)
    np.testing.assert_allclose(vel[2, :, 10], res.phase_velocity[2, :, 10], rtol=1e-9)
    reviewed = (flags & int(api.ContinuityFlag.REVIEWED)) != 0
    assert reviewed.sum() == np.count_nonzero(flags & (16 | 32 | 64 | 128))


@pytest.mark.gpu
def test_auto_accuracy_carries_continuity_flags(gpu):
    wl = workloads.precision_batch(B=32, seed=22)
    per = np.geomspace(.5, 30., 60)
    with api.Solver(c_tol_rel=1e-10, group_velocity=True) as s:     # accuracy="auto", continuity_check=True
        auto = s(wl.thickness, wl.vp, wl.vs, wl.rho, per, 4)
        st = s.stats()
    with api.Solver(accuracy="high", c_tol_rel=1e-10, group_velocity=True) as s:
        high = s(wl.thickness, wl.vp, wl.vs, wl.rho, per, 4)
    assert high.continuity_flags is None
    cf = auto.continuity_flags
    assert cf.shape == (32, per.size) and cf.dtype == np.uint8
    assert np.all(cf[:, [0, -1]] & int(api.ContinuityFlag.NO_BILATERAL))
    for key in ("continuity_suspected", "continuity_repaired", "continuity_unresolved", "continuity_local_searches",
                "continuity_full_searches"):
        assert key in st
    assert st["continuity_suspected"] == int(np.count_nonzero(cf & int(api.ContinuityFlag.REVIEWED)))
    np.testing.assert_array_equal(auto.mode_mask, high.mode_mask)
    ok = high.mode_mask
    assert np.max(np.abs(auto.phase_velocity[ok] - high.phase_velocity[ok]) / high.phase_velocity[ok]) <= 1e-9
