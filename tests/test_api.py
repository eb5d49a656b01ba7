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


def _numpy_masks(velocity, flags, counts, rtol, group=None, gflags=None):
    """The reference DispersionBatch's numpy expressions for the per-cell masks
    (python/modisp/__init__.py:100-123)."""
    nmodes = velocity.shape[1]
    close_mask = np.zeros(velocity.shape, dtype=bool)
    if nmodes > 1:
        gaps = np.diff(velocity, axis=1)
        close = np.isfinite(gaps) & (gaps <= rtol * velocity[:, 1:, :])
        close_mask[:, :-1, :] |= close
        close_mask[:, 1:, :] |= close
    risk = ((flags & api.TRACE_CLOSE_ROOT_RISK) != 0) | np.any(close_mask, axis=1)
    mode_mask = np.isfinite(velocity)
    status = np.where(mode_mask, api.SolveStatus.VALID,
                      np.where((flags[:, None, :] & (api.TRACE_SEARCH_UNCONVERGED | api.TRACE_UNRESOLVED_BRACKET)) != 0,
                               api.SolveStatus.UNRESOLVED, api.SolveStatus.NOT_FOUND)).astype(np.uint8)
    rejected = (np.isnan(velocity) & ((flags[:, None, :] & api.TRACE_F_REJECTED) != 0) &
                (np.arange(nmodes)[None, :, None] < counts[:, None, :]))
    status[rejected] = api.SolveStatus.REJECTED
    gmask = None if group is None else np.isfinite(group) & (gflags == 0)
    return status, mode_mask, close_mask, risk, gmask


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("modes", [1, 5])
@pytest.mark.parametrize("rtol", [.005, .007])
def test_result_masks_equal_the_reference_numpy_expressions(dtype, modes, rtol):
    """The native one-pass mask assembly (md_rocm_result_masks) against the
    reference's numpy expressions on cells with missing roots, gaps at and
    around the close-root threshold (in the velocities' own precision),
    rejected / unresolved / close-risk traces and flagged group velocities.
    With FP32 velocities and rtol = 0.007 some threshold gaps get the opposite
    verdict in FP64 arithmetic; with rtol = 0.005 none can (f32(0.005) v
    rounds to the FP64 threshold at every representable gap in [2, 4.5))."""
    rng = np.random.default_rng(modes)
    B, T = 700, 37
    v = np.sort(rng.uniform(2., 4.5, (B, modes, T)), axis=1).astype(dtype)
    split = 0
    if modes > 1:
        # Gaps within a few ulp of the threshold, and a share of them where the
        # verdict differs between the array's precision and FP64 (the reference
        # evaluates in the array's precision): such pairs are rare, so they are
        # drawn from a pool found by search over many upper velocities.
        up = rng.uniform(2., 4.5, 400_000).astype(dtype)
        cand = [(up * (1 - rtol)).astype(dtype)]
        for _ in range(4):
            cand.append(np.nextafter(cand[-1], dtype(np.inf)))
            cand.insert(0, np.nextafter(cand[0], dtype(0)))
        cand = np.stack(cand)
        own = (up - cand) <= dtype(rtol) * up
        wide = (up.astype(np.float64) - cand.astype(np.float64)) <= rtol * up.astype(np.float64)
        i, j = np.nonzero(own != wide)
        pool_up, pool_low = up[j], cand[i, j]
        near_up, near_low = up[:B * T], cand[rng.integers(0, 9, B * T), np.arange(B * T)]
        assert (pool_up.size > 0) == (dtype == np.float32 and rtol == .007), pool_up.size
        pick = rng.random(B * T) < .5
        take = rng.integers(0, max(1, pool_up.size), B * T)
        hi = np.where(pick & (pool_up.size > 0), pool_up[take] if pool_up.size else near_up, near_up)
        lo = np.where(pick & (pool_up.size > 0), pool_low[take] if pool_up.size else near_low, near_low)
        k = rng.random((B, T)) < .3
        v[:, 2, :] = np.where(k, hi.reshape(B, T), v[:, 2, :])
        v[:, 1, :] = np.where(k, lo.reshape(B, T), v[:, 1, :])
        g = v[:, 2, :] - v[:, 1, :]
        split = int((k & ((g <= dtype(rtol) * v[:, 2, :]) !=
                          (v[:, 2, :].astype(np.float64) - v[:, 1, :].astype(np.float64)
                           <= rtol * v[:, 2, :].astype(np.float64)))).sum())
    v[rng.random(v.shape) < .15] = np.nan
    flags = rng.choice(np.array([0, 0, 0, api.TRACE_F_REJECTED, api.TRACE_SEARCH_UNCONVERGED,
                                 api.TRACE_UNRESOLVED_BRACKET, api.TRACE_CLOSE_ROOT_RISK,
                                 api.TRACE_F_REJECTED | api.TRACE_UNRESOLVED_BRACKET], dtype=np.uint32), (B, T))
    counts = rng.integers(-1, modes + 2, (B, T)).astype(np.int32)
    group = np.where(rng.random(v.shape) < .1, np.nan, v.astype(np.float64) * .9)
    gflags = np.where(rng.random(v.shape) < .1, 2, 0).astype(np.uint32)
    for g, gf in ((group, gflags), (None, None)):
        got = modisp_amd.result_masks(v, flags, counts, rtol, g, gf)
        want = _numpy_masks(v, flags, counts, rtol, g, gf)
        for a, b, name in zip(got, want, ("status", "mode_mask", "close_root_mask", "close_root_risk", "group")):
            if b is None:
                assert a is None, name
                continue
            assert a.dtype == b.dtype and a.shape == b.shape, name
            np.testing.assert_array_equal(a, b, err_msg=name)
    if modes > 1:
        assert want[2].any() and not want[2].all()      # the threshold cases are exercised
        assert split > 100 or not (dtype == np.float32 and rtol == .007)  # FP32 cells FP64 would flip
    assert (want[0] == api.SolveStatus.REJECTED).any() and (want[0] == api.SolveStatus.UNRESOLVED).any()


# --- options that select methods (CPU: validation; GPU: effect) -------------------------------
def test_uniform_strategy_follows_the_reference_restrictions():
    with pytest.raises(NotImplementedError, match="legacy refinement"):
        api.Solver(search_strategy="uniform", refinement="brent")
    with pytest.raises(NotImplementedError, match="adaptive solver"):
        api.Solver(search_strategy="uniform", group_velocity=True)
    with pytest.raises(ValueError, match="float32"):
        api.Solver(search_strategy="uniform", precision="float32")
    s = api.Solver(search_strategy="uniform")
    assert s.accuracy == "uniform" and s.refinement == "legacy" and s.c_tol_rel == 1e-6


def test_engine_options_follow_the_front_end_options():
    s = api.Solver(accuracy="balanced", search_strategy="qe_adaptive", love_strategy="scan", precision="float32",
                   refinement="chandrupatla")
    o = s._engine_options(False)
    assert (o["accuracy"], o["shape"], o["love_phase"], o["single_precision"], o["refinement"]) == (1, False, 0, True, 2)
    o = api.Solver()._engine_options(False)           # the reference defaults: auto, shape_adaptive, phase, FP64
    assert (o["accuracy"], o["shape"], o["love_phase"], o["single_precision"], o["sampling_rtol"]) == (4, True, 1,
                                                                                                      False, -1.)
    assert api.Solver(secular="dunkin")._engine_options(False)["secular"] == 0    # refused by the engine


def test_review_groups_validates_its_trace():
    with pytest.raises(ValueError, match="invalid review trace"):
        modisp_amd._review_groups(np.ones(7), 2, 3, 0, 1, 1., 1., 2., 1e-3, .005)   # fewer than 4 layers of data
    with pytest.raises(ValueError, match="invalid review trace"):
        modisp_amd._review_groups(np.ones(8), 2, 3, 0, 1, -1., 1., 2., 1e-3, .005)  # period


@pytest.mark.gpu
def test_precision_float64_runs_the_plain_fp64_scan(gpu):
    wl = workloads.plain_batch(B=4, seed=9)
    with api.Solver(accuracy="high", c_tol_rel=1e-12) as s:
        r64 = s(wl.thickness, wl.vp, wl.vs, wl.rho, wl.periods[::10], 3)
        assert s._engine.scan == "float64" and s._engine.config.single_precision == 0
        assert s.stats()["fp64_rescans"] == 0 and s.stats()["fp64_rechecks"] == 0
    with api.Solver(accuracy="high", c_tol_rel=1e-12, precision="float32") as s:
        r32 = s(wl.thickness, wl.vp, wl.vs, wl.rho, wl.periods[::10], 3)
        assert s._engine.scan == "float32"
    assert r64.phase_velocity.dtype == np.float64 and r32.phase_velocity.dtype == np.float32
    np.testing.assert_array_equal(r64.mode_mask, r32.mode_mask)


@pytest.mark.gpu
@pytest.mark.parametrize("accept", [1e-30, 1e-3])
def test_f_accept_rejects_like_the_reference(gpu, accept):
    """f_accept < 1 rejects roots whose |F| exceeds it: NaN, TRACE_F_REJECTED,
    status REJECTED, group flag MISSING_ROOT (reference cpp/cpu/solver.cpp:333).
    1e-30 rejects every root (|F| at a root is ~1e-16 or more), 1e-3 none."""
    import oracle
    wl = workloads.plain_batch(B=6, seed=10)
    per = wl.periods[::8]
    with api.Solver(accuracy="high", c_tol_rel=1e-12, f_accept=accept, group_velocity=True) as s:
        got = s(wl.thickness, wl.vp, wl.vs, wl.rho, per, 3)
    b = oracle.bounds(wl.thickness, wl.vp, wl.vs, wl.rho, 0)
    ref = oracle.solve(wl.thickness, wl.vp, wl.vs, wl.rho, per, b, 3, 0, automatic=True, tol=1e-12, accept=accept)
    # Where the two evaluators' |F| at a root falls on either side of the
    # threshold only by rounding (|F| of the oracle at the device root within
    # 1e-14 of zero: one of them evaluates exactly 0), the verdicts may differ.
    differ = got.mode_mask != np.isfinite(ref["phase_velocity"])
    for i, m, f in np.argwhere(differ):
        c = got.phase_velocity[i, m, f] if got.mode_mask[i, m, f] else ref["phase_velocity"][i, m, f]
        model = [getattr(wl, k)[i] for k in ("thickness", "vp", "vs", "rho")]
        assert abs(oracle.secular(*model, per[f], c)[0]) <= 1e-14, (i, m, f)
    same = ~differ.any(axis=1)
    np.testing.assert_array_equal((got.trace_flags & api.TRACE_F_REJECTED)[same],
                                  (ref["trace_flags"] & api.TRACE_F_REJECTED)[same])
    assert differ.sum() <= 2
    if accept < 1e-20:
        assert got.mode_mask.sum() <= differ.sum()
        found = np.arange(3)[None, :, None] < got.root_count[:, None, :]
        assert np.all((got.status == api.SolveStatus.REJECTED) == (found & ~got.mode_mask))
        assert np.all(got.group_velocity_flags[got.status == api.SolveStatus.REJECTED] == 1)
    else:
        assert not (got.trace_flags & api.TRACE_F_REJECTED).any()


@pytest.mark.gpu
def test_presets_report_the_reference_levels_and_approximate_flags(gpu):
    wl = workloads.plain_batch(B=3, seed=11)
    per = wl.periods[::12]
    cases = (("high", "shape_adaptive", 8, 0), ("balanced", "shape_adaptive", 9, api.TRACE_APPROXIMATE_SEARCH),
             ("fast", "shape_adaptive", 10, api.TRACE_APPROXIMATE_SEARCH), ("high", "qe_adaptive", 3, 0),
             ("balanced", "qe_adaptive", 6, api.TRACE_APPROXIMATE_SEARCH),
             ("fast", "qe_adaptive", 7, api.TRACE_APPROXIMATE_SEARCH))
    for accuracy, strategy, level, flag in cases:
        with api.Solver(accuracy=accuracy, search_strategy=strategy, continuity_check=False) as s:
            r = s(wl.thickness, wl.vp, wl.vs, wl.rho, per, 2)
        assert np.all(r.search_level == level), (accuracy, strategy)
        assert np.all((r.trace_flags & api.TRACE_APPROXIMATE_SEARCH) == flag), (accuracy, strategy)
    with api.Solver(accuracy="high", love_strategy="scan") as s:
        love = s(wl.thickness, wl.vp, wl.vs, wl.rho, per, 2, "love")
    assert not np.any(love.search_level == 11)
    with api.Solver(accuracy="high") as s:
        love = s(wl.thickness, wl.vp, wl.vs, wl.rho, per, 2, "love")
    # The phase count accepts a trace (11) unless the unwrapped phase is lost
    # at a window end (an opaque evanescent stack at the automatic window's
    # low end); then the scan's roots stand with TRACE_LOVE_PHASE_FALLBACK, as
    # in the reference (cpp/cpu/solver.cpp:285-291).
    fallback = (love.trace_flags & api.TRACE_LOVE_PHASE_FALLBACK) != 0
    assert np.all(love.search_level[~fallback] == 11) and np.all(love.search_level[fallback] == 8)
    assert (~fallback).any()


@pytest.mark.gpu
def test_uniform_strategy_runs_the_uniform_grid(gpu):
    import oracle
    wl = workloads.plain_batch(B=4, seed=12)
    per = wl.periods[::10]
    with api.Solver(search_strategy="uniform", nc=1024, c_tol_rel=1e-12) as s:
        r = s(wl.thickness, wl.vp, wl.vs, wl.rho, per, 3)
        assert s.stats()["search_grid"] == "uniform"
    assert np.all(r.search_level == -2) and np.all(r.trace_flags & api.TRACE_RANGE_TRUNCATED)
    assert np.all(r.search_bounds[:, 0] == .2)
    ref = oracle.solve(wl.thickness, wl.vp, wl.vs, wl.rho, per, r.search_bounds, 3, 0, tol=1e-12)
    ok = r.mode_mask & np.isfinite(ref["phase_velocity"])
    assert ok.sum() > .9 * r.mode_mask.sum()
    np.testing.assert_allclose(r.phase_velocity[ok], ref["phase_velocity"][ok], rtol=1e-9)


@pytest.mark.gpu
def test_review_groups_equal_the_oracle_analytic_group(gpu):
    import oracle
    wl = workloads.precision_batch(B=2, seed=13)
    h, vp, vs, rho = (getattr(wl, k)[0] for k in ("thickness", "vp", "vs", "rho"))
    win = workloads.scan_window(wl.vs)[0]
    period = 3.
    ref = oracle.solve(h, vp, vs, rho, [period], win[None], 6, 0, tol=1e-13, group=2)
    roots = ref["phase_velocity"][0, :, 0]
    roots = roots[np.isfinite(roots)]
    packed = np.concatenate([h, vp, vs, rho, roots])
    u, f, ev = modisp_amd._review_groups(packed, h.size, 4, 0, 1, period, win[0], win[1], 1e-3, .005)
    np.testing.assert_array_equal(f, ref["group_velocity_flags"][0, :4, 0])
    ok = f == 0
    np.testing.assert_allclose(u[ok], ref["group_velocity"][0, :4, 0][ok], rtol=1e-9)
    assert ev == int(ok.sum())
    u2, f2, _ = modisp_amd._review_groups(packed[:4 * h.size + 2], h.size, 4, 0, 1, period, win[0], win[1], 1e-3, .005)
    assert np.all(f2[2:] == 1) and np.all(np.isnan(u2[2:]))
