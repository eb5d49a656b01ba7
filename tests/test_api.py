"""Front end (modisp_amd.api): the reference's Solver / dispersion_batch /
dispersion surface on the ROCm backend.

CPU tests: argument validation and backend selection (no device calls).
GPU tests: results against the oracle's restatement of the reference
front end (oracle.reference_result: bounds, flags, counts, status codes)."""
import numpy as np
import pytest

import modisp_amd
from modisp_amd import api
from modisp_amd import workloads


def _one(B=2, L=5, seed=0):
    return workloads.monotone_models(np.random.default_rng(seed), B, L)


def test_model_validation_messages_follow_the_reference():
    h, vp, vs, rho = _one()
    with pytest.raises(ValueError, match="shapes must agree"):
        api._models(h, vp[:, :-1], vs, rho)
    with pytest.raises(ValueError, match="finite nonempty"):
        api._models(h * np.nan, vp, vs, rho)
    bad = vs.copy()
    bad[:, -1] = 0.
    with pytest.raises(ValueError, match="half-space must be solid"):
        api._models(h, vp, bad, rho)
    with pytest.raises(ValueError, match="sqrt\\(4/3\\)"):
        api._models(h, vs * 1.1, vs, rho)
    with pytest.raises(ValueError, match="must be real"):
        api._models(h + 0j, vp, vs, rho)


@pytest.mark.parametrize("kw,msg", [
    (dict(backend="gpu"), "backend"), (dict(precision="float16"), "precision"), (dict(accuracy="max"), "accuracy"),
    (dict(cmin=2., cmax=1.), "cmax"), (dict(c_tol_rel=-1.), "c_tol_rel"), (dict(group_rtol=2.), "group_rtol"),
    (dict(anchor_stride=1), "anchor_stride"), (dict(frequency_strategy="x"), "frequency_strategy"),
    (dict(nc=2), "nc"), (dict(love_strategy="x"), "love_strategy")])
def test_solver_option_validation(kw, msg):
    with pytest.raises(ValueError, match=msg):
        api.Solver(**kw)


def test_reference_default_tolerances_by_accuracy():
    for acc, tol in (("auto", 1e-5), ("high", 1e-6), ("balanced", 1e-5), ("fast", 1e-4)):
        assert api.Solver(accuracy=acc).c_tol_rel == tol


def test_other_backends_are_not_part_of_this_package():
    h, vp, vs, rho = _one()
    for backend in ("cpu", "cuda"):
        with pytest.raises(modisp_amd.BackendUnavailableError):
            api.Solver(backend=backend)(h, vp, vs, rho, [1., 2.])


@pytest.mark.skipif(modisp_amd.available(0), reason="a gfx950 device is present")
def test_auto_without_a_device_raises_instead_of_falling_back():
    h, vp, vs, rho = _one()
    with pytest.raises(modisp_amd.BackendUnavailableError):
        api.dispersion_batch(h, vp, vs, rho, [1., 2.])


def test_dispersion_rejects_malformed_models():
    with pytest.raises(ValueError, match="columns h/vp/vs/rho"):
        api.dispersion(np.ones((3, 3)), [1.])


def test_dispersion_batch_rejects_options_conflicting_with_a_solver():
    s = api.Solver(accuracy="high")
    h, vp, vs, rho = _one()
    with pytest.raises(ValueError, match="conflict"):
        api.dispersion_batch(h, vp, vs, rho, [1.], solver=s, accuracy="fast")


# --- GPU --------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def gpu():
    if not modisp_amd.available(0):
        pytest.fail("no gfx950 device visible to the gpu tests")
    return modisp_amd


@pytest.mark.gpu
@pytest.mark.parametrize("wave", ["rayleigh", "love"])
@pytest.mark.parametrize("window", ["automatic", "cropped"])
def test_front_end_matches_reference_result(gpu, wave, window):
    import oracle
    wl = workloads.plain_batch(B=12, seed=4)
    per = wl.periods[::6]
    cmin, cmax = (None, None) if window == "automatic" else (1.5, 4.)
    with api.Solver(accuracy="high", c_tol_rel=1e-12, cmin=cmin, cmax=cmax) as s:
        got = s(wl.thickness, wl.vp, wl.vs, wl.rho, per, 3, wave)
    ref = oracle.reference_result(wl.thickness, wl.vp, wl.vs, wl.rho, per, 3, wave, cmin, cmax, tol=1e-12,
                                  refinement=1, love_phase=1)
    np.testing.assert_array_equal(got.search_bounds.view(np.uint64), ref["search_bounds"].view(np.uint64))
    resolved = ref["root_count"] >= 0
    np.testing.assert_array_equal(got.root_count[resolved], ref["root_count"][resolved])
    cells = np.broadcast_to(resolved[:, None, :], got.mode_mask.shape)
    np.testing.assert_array_equal(got.mode_mask[cells], ref["mode_mask"][cells])
    np.testing.assert_array_equal(got.status[cells], ref["status"][cells])
    keep = np.uint32(api.TRACE_NEAR_CUTOFF | api.TRACE_RANGE_TRUNCATED | api.TRACE_CLOSE_ROOT_RISK)
    np.testing.assert_array_equal((got.trace_flags & keep)[resolved], (ref["trace_flags"] & keep)[resolved])
    ok = ref["mode_mask"] & cells
    np.testing.assert_allclose(got.phase_velocity[ok], ref["phase_velocity"][ok], rtol=1e-9)
    assert got.backend_used == "rocm" and got.phase_velocity.dtype == np.float64


@pytest.mark.gpu
def test_unsorted_repeated_periods_are_solved_once_and_restored(gpu):
    wl = workloads.plain_batch(B=4, seed=5)
    per = np.array([5., 1., 5., 2., 20.])
    with api.Solver(accuracy="high", c_tol_rel=1e-12, group_velocity=True) as s:
        a = s(wl.thickness, wl.vp, wl.vs, wl.rho, per, 3)
        assert s.stats()["periods_computed"] == 4
        b = s(wl.thickness, wl.vp, wl.vs, wl.rho, np.array([1., 2., 5., 20.]), 3)
    order = [2, 0, 2, 1, 3]
    np.testing.assert_array_equal(a.phase_velocity, b.phase_velocity[:, :, order])
    np.testing.assert_array_equal(a.group_velocity, b.group_velocity[:, :, order])
    np.testing.assert_array_equal(a.period, per)


@pytest.mark.gpu
def test_float32_precision_and_single_model_dispersion(gpu):
    wl = workloads.plain_batch(B=1, seed=6)
    model = np.stack([wl.thickness[0], wl.vp[0], wl.vs[0], wl.rho[0]], axis=1)
    r64 = api.dispersion(model, [2., 8.], 3, accuracy="high", c_tol_rel=1e-12)
    r32 = api.dispersion(model, [2., 8.], 3, accuracy="high", c_tol_rel=1e-12, precision="float32")
    assert r32.phase_velocity.dtype == np.float32
    np.testing.assert_array_equal(r32.mode_mask, r64.mode_mask)
    np.testing.assert_allclose(r32.phase_velocity[r32.mode_mask], r64.phase_velocity[r64.mode_mask], rtol=1e-6)


@pytest.mark.gpu
def test_anchor_continuation_front_end_equals_independent_search(gpu):
    """frequency_strategy="anchor_continuation" follows the reference's
    conditions (python/modisp/__init__.py:324-341): it runs on sorted unique
    periods, reports prediction flags in the caller's period order, and every
    root equals the independent search's."""
    wl = workloads.precision_batch(B=16, seed=7)
    per = np.geomspace(1., 40., 40)   # log step 0.095: interior targets within the 0.25 gap limit
    shuffled = per[np.random.default_rng(3).permutation(per.size)]
    common = dict(accuracy="balanced", c_tol_rel=1e-12, group_velocity=True)
    with api.Solver(**common) as s:
        ind = s(wl.thickness, wl.vp, wl.vs, wl.rho, shuffled, 4)
    with api.Solver(frequency_strategy="anchor_continuation", **common) as s:
        con = s(wl.thickness, wl.vp, wl.vs, wl.rho, shuffled, 4)
        stats = s.stats()
    assert ind.frequency_strategy_used == "independent" and ind.prediction_flags is None
    assert con.frequency_strategy_used == "anchor_continuation", con.frequency_strategy_reason
    assert con.frequency_strategy_reason == "ROCM anchor continuation on sorted unique periods"
    np.testing.assert_array_equal(con.period, shuffled)
    pred = con.prediction_flags
    assert pred.shape == (16, per.size)
    kinds = sum(((pred & int(bit)) != 0).astype(int) for bit in
                (api.PredictionFlag.ANCHOR, api.PredictionFlag.ACCEPTED, api.PredictionFlag.FALLBACK))
    assert np.all(kinds == 1)
    assert stats["continuation_accepted"] == int(np.count_nonzero(pred & int(api.PredictionFlag.ACCEPTED)))
    np.testing.assert_array_equal(con.mode_mask, ind.mode_mask)
    ok = ind.mode_mask
    assert np.max(np.abs(con.phase_velocity[ok] - ind.phase_velocity[ok]) / ind.phase_velocity[ok]) <= 1e-9
    with api.Solver(frequency_strategy="anchor_continuation", accuracy="high", c_tol_rel=1e-12) as s:
        high = s(wl.thickness, wl.vp, wl.vs, wl.rho, per, 4)
    assert high.frequency_strategy_used == "independent"
    assert high.frequency_strategy_reason == "high accuracy retains independent search"
    assert np.all(high.prediction_flags == int(api.PredictionFlag.BYPASSED))


@pytest.mark.gpu
def test_kernel_report_has_the_reference_keys(gpu):
    wl = workloads.plain_batch(B=2, seed=8)
    with api.Solver(accuracy="high") as s:
        s(wl.thickness, wl.vp, wl.vs, wl.rho, np.array([2., 5.]), 2, "love")
        rep = s.kernel_report()
    for key in ("registers_per_thread", "static_shared_bytes", "local_bytes_per_thread", "search_block_threads",
                "max_active_blocks_per_sm", "configured_blocks_per_sm", "dynamic_shared_bytes",
                "cache_points_per_trace", "table_capacity_per_model", "scan_kernel_report"):
        assert key in rep, key
    assert rep["search_kernel"] == "love_phase_with_scan_fallback"
    scan = rep["scan_kernel_report"]
    assert scan["local_bytes_per_thread"] == 0 and scan["max_active_blocks_per_sm"] >= 1
    deep = gpu.qe_kernel_report(200, single_precision=True)
    assert deep["table_capacity_per_model"] == 256 and deep["search_block_threads"] == 64
    with pytest.raises(ValueError):
        gpu.qe_kernel_report(0)
