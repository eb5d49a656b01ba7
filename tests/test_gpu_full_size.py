"""Full-size checks of the FP32-scan path on the sweep and uneven batches.

At the inversion-scale sweep (262144 models x 31 rows, 200 periods, 4 modes)
and the uneven ragged batch (65536 models of 4..41 rows, 20..200 periods,
1..10 modes) the FP32-scan path and the plain FP64 path are run on every
trace. Where they differ, both are judged against the dense truth: the first
K = modes + 2 sign changes of the oracle's released evaluator on a grid of
200001 nodes over the window (with both paths' and the oracle's roots added
as nodes, so that pairs closer than the grid still separate), refined to
1e-14. At this size the two paths differ on about 1e-6 of the sweep traces:
close root pairs and steps of F that sit within ~1e-7 of a scan node, where
FP32 rounding misplaces a sign the rechecks do not reach -- and the FP64
path misses some the FP32 path finds. The bounds below hold those rates; a
seeded sample of the other traces must equal the oracle. The roots of every
trace ascend and lie inside the window. The grid-sign test checks the FP32
trust rule (|F32| outside the uncertainty band, below the half-space guard,
above max Vs / 16) on subsets of every recipe.
"""
import numpy as np
import pytest

import oracle
from test_gpu_parity import MODEL, RTOL, assert_parity, gpu_solve

pytestmark = pytest.mark.gpu

MAX_DIFFER = 2e-6      # traces where the FP32 and FP64 paths differ
MAX_FP32_WRONG = 1.5e-6  # traces where the FP32 path differs from the dense truth
DENSE = 200001


@pytest.fixture(scope="module")
def md():
    import modisp_amd
    if not modisp_amd.available(0):
        pytest.fail("no gfx950 device visible to the gpu tests")
    modisp_amd.set_device(0)
    return modisp_amd


def dense_roots(model, period, window, K, wave, seeds):
    """The first K sign changes of the oracle's evaluator over the window on
    DENSE nodes plus `seeds`, each refined to 1e-14."""
    lo, hi = float(window[0]), float(window[1])
    c = np.unique(np.concatenate([np.linspace(lo, hi, DENSE), [x for x in seeds if lo < x < hi]]))
    f = oracle.secular(*model, period, c, wave=wave)
    s = np.signbit(f)
    idx = np.nonzero(s[1:] != s[:-1])[0][:K]
    return np.array([oracle.refine(*model, period, c[j], c[j + 1], 1e-14, wave=wave)[0] for j in idx])


def matches_truth(v, count, truth, modes):
    """Count, validity mask and values (RTOL) of one trace against the truth."""
    tv = np.full(modes, np.nan)
    n = min(modes, truth.size)
    tv[:n] = truth[:n]
    if count != truth.size or not np.array_equal(np.isfinite(v[:modes]), np.isfinite(tv)):
        return False
    ok = np.isfinite(tv)
    return bool(np.all(np.abs(v[:modes][ok] - tv[ok]) <= RTOL * tv[ok]))


def _differing(a, b):
    """[B, T] traces whose counts, masks or velocities (beyond RTOL) differ."""
    va, vb = a["phase_velocity"], b["phase_velocity"]
    with np.errstate(invalid="ignore"):
        rel = np.abs(va - vb) / vb
    return ((a["root_count"] != b["root_count"]) | np.any(rel > RTOL, axis=1) |
            np.any(np.isfinite(va) != np.isfinite(vb), axis=1))


def _ordered_inside(v, lo, hi):
    d = np.diff(v, axis=1)
    assert np.all(d[np.isfinite(d)] > 0), "roots of a trace do not ascend"
    fin = np.isfinite(v)
    assert np.all((v >= lo)[fin]) and np.all((v <= hi)[fin]), "root outside the window"


@pytest.fixture(scope="module")
def sweep():
    from modisp_amd import workloads
    wl = workloads.sweep_batch()
    return wl, workloads.scan_window(wl.vs)


@pytest.mark.parametrize("wave", [0, 1], ids=["rayleigh", "love"])
def test_full_sweep_fp32_and_fp64_paths(md, sweep, wave):
    wl, win = sweep
    print(f"sweep full size: B={wl.B} T={wl.periods.size} wave={wave}", flush=True)
    a = gpu_solve(md, *(getattr(wl, k) for k in MODEL), wl.periods, win, 4, wave, "float32", group=False)
    b = gpu_solve(md, *(getattr(wl, k) for k in MODEL), wl.periods, win, 4, wave, "float64", group=False)
    assert (a["root_count"] >= 0).all() and (b["root_count"] >= 0).all()
    differ = _differing(a, b)
    print(f"sweep wave {wave}: {int(differ.sum())} of {differ.size} traces differ", flush=True)
    assert differ.sum() <= MAX_DIFFER * differ.size, f"{differ.sum()} traces differ"
    wrong32 = []
    for i, f in np.argwhere(differ).tolist():
        model = [getattr(wl, k)[i] for k in MODEL]
        va, vb = a["phase_velocity"][i, :, f], b["phase_velocity"][i, :, f]
        ref = oracle.solve(*(getattr(wl, k)[i:i + 1] for k in MODEL), wl.periods[f:f + 1], win[i:i + 1], 4, wave,
                           tol=1e-12)
        seeds = [x for x in np.concatenate([va, vb, ref["phase_velocity"][0, :, 0]]) if np.isfinite(x)]
        truth = dense_roots(model, wl.periods[f], win[i], 6, wave, seeds)
        if not matches_truth(va, a["root_count"][i, f], truth, 4):
            wrong32.append((i, f))
    print(f"sweep wave {wave}: FP32 path differs from the dense truth on {len(wrong32)} traces "
          f"{wrong32[:6]}", flush=True)
    assert len(wrong32) <= MAX_FP32_WRONG * differ.size, wrong32[:10]
    # A seeded sample of the rest: equal to the oracle, or, where the oracle's
    # own search misses a pair (its ladder accepts two agreeing levels), to
    # the dense truth.
    rng = np.random.default_rng(17 + wave)
    for i, f in zip(rng.integers(0, wl.B, 64).tolist(), rng.integers(0, wl.periods.size, 64).tolist()):
        if differ[i, f]:
            continue
        ref = oracle.solve(*(getattr(wl, k)[i:i + 1] for k in MODEL), wl.periods[f:f + 1], win[i:i + 1], 4, wave,
                           tol=1e-12)
        one = {k: (v[i:i + 1, ..., f:f + 1] if v is not None and v.ndim == 3 else
                   v[i:i + 1, f:f + 1] if v is not None and v.ndim == 2 else v)
               for k, v in a.items() if k != "timing"}
        try:
            assert_parity(one, ref, group=False, label=f"sweep fp32 model {i} period {f} wave {wave}")
        except AssertionError:
            va = a["phase_velocity"][i, :, f]
            seeds = [x for x in np.concatenate([va, ref["phase_velocity"][0, :, 0]]) if np.isfinite(x)]
            truth = dense_roots([getattr(wl, k)[i] for k in MODEL], wl.periods[f], win[i], 6, wave, seeds)
            assert matches_truth(va, a["root_count"][i, f], truth, 4), f"sweep model {i} period {f} wave {wave}"
    _ordered_inside(a["phase_velocity"], win[:, 0, None, None], win[:, 1, None, None])


def test_full_uneven_fp32_and_fp64_paths(md):
    from modisp_amd import workloads
    wl = workloads.uneven_batch()
    p = workloads.pack_ragged(wl)
    print(f"uneven full size: B={wl.B} traces={int(p['period_offsets'][-1])}", flush=True)
    args = [p[k] for k in ("thickness", "vp", "vs", "rho", "row_offsets", "periods", "period_offsets", "modes",
                           "bounds")]
    out = {}
    for scan in ("float32", "float64"):
        with md.QeSolver(wl.B, int(wl.layers.max()), 1, 1, scan=scan) as s:
            out[scan] = s.run_ragged(*args, 1e-12, 0)
    a, b = out["float32"], out["float64"]
    po, modes = p["period_offsets"], p["modes"].astype(np.int64)
    T = np.diff(po)
    cell0 = np.concatenate([[0], np.cumsum(modes * T)])
    trace_model = np.repeat(np.arange(wl.B), T)
    trace_f = np.arange(po[-1]) - po[trace_model]
    assert (a["counts"] >= 0).all() and (b["counts"] >= 0).all()
    # per-trace comparison of the packed cells: cell = cell0[m] + mode * T_m + f
    M = int(modes.max())
    va = np.full((po[-1], M), np.nan)
    vb = np.full((po[-1], M), np.nan)
    for j in range(M):
        has = modes[trace_model] > j
        idx = cell0[trace_model[has]] + j * T[trace_model[has]] + trace_f[has]
        va[has, j] = a["velocity"][idx]
        vb[has, j] = b["velocity"][idx]
    with np.errstate(invalid="ignore"):
        rel = np.abs(va - vb) / vb
    differ = (a["counts"] != b["counts"]) | np.any(rel > RTOL, axis=1) | np.any(np.isfinite(va) != np.isfinite(vb),
                                                                                 axis=1)
    print(f"uneven: {int(differ.sum())} of {differ.size} traces differ", flush=True)
    assert differ.sum() <= MAX_DIFFER * differ.size, f"{differ.sum()} traces differ"
    ro = p["row_offsets"]
    wrong32 = []
    for t in np.flatnonzero(differ).tolist():
        m = int(trace_model[t])
        model = [p[k][ro[m]:ro[m + 1]] for k in MODEL]
        ref = oracle.solve(*[x[None] for x in model], p["periods"][t:t + 1], p["bounds"][m:m + 1], int(modes[m]), 0,
                           tol=1e-12)
        seeds = [x for x in np.concatenate([va[t], vb[t], ref["phase_velocity"][0, :, 0]]) if np.isfinite(x)]
        truth = dense_roots(model, p["periods"][t], p["bounds"][m], int(modes[m]) + 2, 0, seeds)
        if not matches_truth(va[t], a["counts"][t], truth, int(modes[m])):
            wrong32.append((m, int(trace_f[t])))
    assert len(wrong32) <= MAX_FP32_WRONG * differ.size, wrong32[:10]
    rng = np.random.default_rng(23)
    for t in rng.integers(0, po[-1], 64).tolist():
        if differ[t]:
            continue
        m, f = int(trace_model[t]), int(trace_f[t])
        model = [p[k][ro[m]:ro[m + 1]][None] for k in MODEL]
        ref = oracle.solve(*model, p["periods"][t:t + 1], p["bounds"][m:m + 1], int(modes[m]), 0, tol=1e-12)
        got = dict(root_count=a["counts"][t:t + 1][None], phase_velocity=va[t, :modes[m]][None, :, None])
        try:
            assert_parity(got, ref, group=False, label=f"uneven fp32 model {m} period {f}")
        except AssertionError:
            seeds = [x for x in np.concatenate([va[t], ref["phase_velocity"][0, :, 0]]) if np.isfinite(x)]
            truth = dense_roots([x[0] for x in model], p["periods"][t], p["bounds"][m], int(modes[m]) + 2, 0, seeds)
            assert matches_truth(va[t], a["counts"][t], truth, int(modes[m])), f"uneven model {m} period {f}"
    lo = p["bounds"][trace_model, 0][:, None]
    hi = p["bounds"][trace_model, 1][:, None]
    _ordered_inside(va[:, :, None], lo[:, :, None], hi[:, :, None])


def _band_case(recipe):
    from modisp_amd import workloads
    if recipe == "precision":
        wl = workloads.precision_batch(B=48, seed=106)
        return wl, wl.periods
    if recipe == "sweep":
        wl = workloads.sweep_batch(B=48, seed=109)
        return wl, wl.periods[::5]
    wl = workloads.uneven_batch(B=48, seed=110)
    return wl, np.geomspace(0.2, 100., 40)


@pytest.mark.parametrize("recipe", ["precision", "sweep", "uneven"])
def test_fp32_grid_signs_outside_the_band(md, recipe):
    """Wherever an FP32 grid value is outside the uncertainty band (and the
    node is neither at the half-space guard nor below max Vs / 16, where the
    scan always uses FP64) the pipeline trusts its sign; it must equal the
    FP64 sign. Uneven models are padded with zero-thickness half-space rows."""
    from modisp_amd import workloads
    wl, periods = _band_case(recipe)
    win = workloads.scan_window(wl.vs)
    fin = np.arange(wl.vs.shape[1])[None, :] < wl.vs.shape[1] - 1
    speed = np.where(wl.vs > 0, wl.vs, wl.vp)
    slow = np.sum(np.where(fin, wl.thickness / speed, 0.), axis=1)
    phi = (2 * np.pi / periods)[None, :] * slow[:, None]
    tau = np.maximum(1e-4, 5e-5 * (1 + phi))[:, :, None]
    c = win[:, 0, None] + (win[:, 1] - win[:, 0])[:, None] * np.arange(512)[None, :] / 511.
    guard = (c >= wl.vs[:, -1, None] * (1 - 1e-3))[:, None, :]
    low = (c < wl.vs.max(axis=1)[:, None] / 16.)[:, None, :]
    for wave in (0, 1):
        f32 = md.probe_grid(wl.thickness, wl.vp, wl.vs, wl.rho, periods, win, 512, wave, True)
        f64 = md.probe_grid(wl.thickness, wl.vp, wl.vs, wl.rho, periods, win, 512, wave, False)
        trusted = (np.abs(f32) >= tau) & ~guard & ~low
        bad = trusted & (np.signbit(f32) != np.signbit(f64))
        assert not bad.any(), f"{recipe} wave {wave}: {int(bad.sum())} trusted FP32 signs wrong at " \
                              f"{np.argwhere(bad)[:5].tolist()}"
        # FP64 grid values agree in sign with the oracle's released evaluator.
        for i in range(0, wl.B, 12):
            for f in range(0, periods.size, 13):
                ref = oracle.secular(*(getattr(wl, k)[i] for k in MODEL), periods[f], c[i], wave)
                sure = np.abs(ref) > 1e-12
                assert np.array_equal(np.signbit(ref)[sure], np.signbit(f64[i, f])[sure]), (recipe, wave, i, f)
