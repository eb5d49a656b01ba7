"""GPU parity of the gfx950 root pipeline against the oracle.

Bar: identical root counts and validity masks; phase velocities within 1e-9
relative; analytic group velocities within 1e-9 relative of the oracle's
analytic implicit derivative. Every call goes through the C ABI
(modisp_amd.QeSolver -> libmodisp_rocm.so). Sizes are chosen so the oracle
finishes in seconds; full-size runs are checked through size-independent
properties (FP32 path == FP64 path, ascending roots inside the window).
"""
from pathlib import Path

import numpy as np
import pytest

import oracle
import reference_vectors as refvec

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).parent / "golden"
MODEL = ("thickness", "vp", "vs", "rho")
RTOL = 1e-9


@pytest.fixture(scope="module")
def md():
    import modisp_amd
    if not modisp_amd.available(0):
        pytest.fail("no gfx950 device visible to the gpu tests")
    modisp_amd.set_device(0)
    return modisp_amd


def gpu_solve(md, h, vp, vs, rho, periods, bounds, nmodes, wave, scan="float32", group=True, nc=512, refinement=1):
    h, vp, vs, rho = (np.atleast_2d(np.asarray(x, dtype=np.float64)) for x in (h, vp, vs, rho))
    B, L = h.shape
    per = np.asarray(periods, dtype=np.float64)
    nf = per.shape[-1]
    with md.QeSolver(B, L, nf, nmodes, group_velocity=group, scan=scan, nc=nc, refinement=refinement,
                     shape=False) as s:
        vel, flags, st, lev, cnt, ex = s.run(h, vp, vs, rho, per, bounds, 1e-12, 1., wave, False)
        return dict(phase_velocity=vel, trace_flags=flags, search_level=lev, root_count=cnt,
                    group_velocity=ex["group_velocity"], group_velocity_flags=ex["group_velocity_flags"],
                    timing=s.last_timing())


def assert_parity(got, ref, group=True, label=""):
    """Traces the oracle itself leaves unresolved (count -1: its six-level
    search never agreed) have no reference values and are not compared.
    Trace flags and search levels follow tests/reference_vectors.py: the
    result-defined flag bits, and the level where the oracle's ladder accepted
    the trace at its adaptive level (3), by the Love phase count (11) or as
    empty (-3/-4)."""
    resolved = ref["root_count"] >= 0
    if "trace_flags" in got and "trace_flags" in ref:
        primary = np.isin(ref["search_level"], refvec.PRIMARY_LEVELS["qe"])
        bits = np.where(primary, refvec.FLAG_BITS, refvec.FLAG_BITS & ~4096).astype(np.uint32)
        # The reference's Love phase path narrows each phase bracket to 16 ulp
        # at c_tol_rel = 1e-12 and falls back to its scan (16384) when F does
        # not change sign across it, i.e. when its phase crossing and F's zero
        # differ by more than that; the scan returns the same roots (compared
        # below). The device's exact phase count accepts them (level 11).
        love_fallback = (((ref["trace_flags"] & 16384) != 0) & ((got["trace_flags"] & 16384) == 0)
                         & (got["search_level"] == 11))
        bad = (((got["trace_flags"] ^ ref["trace_flags"]) & bits) != 0) & ~love_fallback
        assert not bad.any(), f"{label} trace flags differ at {np.argwhere(bad)[:5].tolist()}"
        lv = primary & (got["search_level"] != ref["search_level"]) & ~love_fallback
        assert not lv.any(), f"{label} search levels differ at {np.argwhere(lv)[:5].tolist()}"
    cells = np.broadcast_to(resolved[:, None, :], ref["phase_velocity"].shape)
    np.testing.assert_array_equal(got["root_count"][resolved], ref["root_count"][resolved],
                                  err_msg=f"{label} root counts")
    gv, rv = got["phase_velocity"], ref["phase_velocity"]
    np.testing.assert_array_equal(np.isfinite(gv)[cells], np.isfinite(rv)[cells], err_msg=f"{label} validity mask")
    ok = np.isfinite(rv) & cells
    if ok.any():
        rel = np.abs(gv[ok] - rv[ok]) / rv[ok]
        assert rel.max() <= RTOL, f"{label} phase rel {rel.max():.3e}"
    if group:
        gg, rg = got["group_velocity"], ref["group_velocity"]
        np.testing.assert_array_equal(np.isfinite(gg)[cells], np.isfinite(rg)[cells], err_msg=f"{label} group mask")
        gok = np.isfinite(rg) & cells
        if gok.any():
            grel = np.abs(gg[gok] - rg[gok]) / np.abs(rg[gok])
            assert grel.max() <= RTOL, f"{label} group rel {grel.max():.3e}"
        np.testing.assert_array_equal(got["group_velocity_flags"][cells], ref["group_velocity_flags"][cells],
                                      err_msg=f"{label} group flags")


# --- the precision workload (FP32 scan + FP64 refine) --------------------------------------------
@pytest.mark.parametrize("wave", [0, 1], ids=["rayleigh", "love"])
@pytest.mark.parametrize("scan", ["float32", "float64"])
def test_precision_workload_subset(md, wave, scan):
    from modisp_amd import workloads
    wl = workloads.precision_batch(B=24, seed=101)
    win = workloads.scan_window(wl.vs)
    got = gpu_solve(md, *(getattr(wl, k) for k in MODEL), wl.periods, win, 5, wave, scan)
    ref = oracle.solve(*(getattr(wl, k) for k in MODEL), wl.periods, win, 5, wave, tol=1e-12, group=2)
    assert_parity(got, ref, label=f"wave {wave} scan {scan}")


@pytest.mark.parametrize("nc", [256, 1024])
def test_other_grid_sizes(md, nc):
    from modisp_amd import workloads
    wl = workloads.precision_batch(B=8, seed=102)
    win = workloads.scan_window(wl.vs)
    per = wl.periods[::4]
    got = gpu_solve(md, *(getattr(wl, k) for k in MODEL), per, win, 5, 0, "float32", nc=nc)
    ref = oracle.solve(*(getattr(wl, k) for k in MODEL), per, win, 5, 0, tol=1e-12, group=2)
    assert_parity(got, ref, label=f"nc {nc}")


def test_plain_workload_subset_automatic_window(md):
    """The CPU-runnable workload with the reference's automatic search window."""
    from modisp_amd import workloads
    wl = workloads.plain_batch(B=64, seed=0)
    b = md.bounds(wl.thickness, wl.vp, wl.vs, wl.rho, 0)
    got = gpu_solve(md, *(getattr(wl, k) for k in MODEL), wl.periods, b, 3, 0, "float32")
    ref = oracle.solve(*(getattr(wl, k) for k in MODEL), wl.periods, b, 3, 0, automatic=True, tol=1e-12, group=2)
    assert_parity(got, ref, label="plain automatic")


@pytest.mark.parametrize("wave", [0, 1], ids=["rayleigh", "love"])
def test_precision_workload_subset_automatic_window(md, wave):
    """The reference's automatic window reaches below max Vs / 16, where the
    FP32 secular function is not trusted: those nodes take FP64 signs
    (the band is root-free for these solid stacks) instead of the whole trace
    going to the FP64 grid."""
    from modisp_amd import workloads
    wl = workloads.precision_batch(B=24, seed=103)
    b = md.bounds(wl.thickness, wl.vp, wl.vs, wl.rho, wave)
    assert np.all(b[:, 0] < wl.vs.max(axis=1) / 16)
    got = gpu_solve(md, *(getattr(wl, k) for k in MODEL), wl.periods, b, 5, wave, "float32")
    ref = oracle.solve(*(getattr(wl, k) for k in MODEL), wl.periods, b, 5, wave, automatic=True, tol=1e-12, group=2)
    assert_parity(got, ref, label=f"precision automatic wave {wave}")


def test_roots_sit_on_sign_changes(md):
    """Refinement ends on the secant point of its final bracket, which is the
    root to rounding: F (the oracle's evaluator) changes sign within 1e-14
    relative either side of the device root (589 of 590 roots here). The
    bracket midpoint, anywhere within half the 1.25e-13 stop width, passes
    this for 18 of the 590."""
    from modisp_amd import workloads
    wl = workloads.precision_batch(B=8, seed=107)
    per = wl.periods[::10]
    win = workloads.scan_window(wl.vs)
    hits = total = 0
    for wave in (0, 1):
        got = gpu_solve(md, *(getattr(wl, k) for k in MODEL), per, win, 5, wave, "float32", group=False)
        for i in range(wl.thickness.shape[0]):
            model = [np.ascontiguousarray(getattr(wl, k)[i]) for k in MODEL]
            for f in range(per.size):
                for r in got["phase_velocity"][i, :, f]:
                    if not np.isfinite(r):
                        continue
                    fl, fr = oracle.secular(*model, per[f], np.array([r * (1 - 1e-14), r * (1 + 1e-14)]), wave=wave)
                    total += 1
                    hits += np.signbit(fl) != np.signbit(fr)
    assert total > 300
    assert hits >= 0.99 * total, f"{hits} of {total} roots within 1e-14 of a sign change"


def test_deep_stack_refinement(md):
    """Deep random-walk stacks at short periods, where the max-normalized F of
    the evanescent stack is a +-1 step across most brackets: refinement on the
    unnormalized F (end scales from the scan) takes few evaluations (a plain
    Brent on the step takes about 17 per bracket), and the roots are the
    reference's except on the traces named in
    tests/golden/reference/deviations.json (a collapsing pair the reference's
    qe_adaptive search misses, verified there as FP64 sign changes)."""
    z = np.load(refvec.VECTORS / "sweep_b32_s106_rayleigh.npz")
    h, vp, vs, rho = (np.ascontiguousarray(z[k]) for k in MODEL)
    per, win = z["periods"], z["bounds"]
    B = h.shape[0]
    named = refvec.load_deviations().get("sweep_b32_s106_rayleigh|qe", {})
    with md.QeSolver(B, h.shape[1], per.size, 4, group_velocity=True, scan="float32", shape=False) as s:
        vel, flags, st, lev, cnt, ex = s.run(h, vp, vs, rho, per, win, 1e-12, 1., 0, False)
        stats = s.last_stats()
    per_bracket = stats["refine_evaluations"] / max(1, cnt.clip(0).sum())
    assert per_bracket < 8., f"{per_bracket:.1f} evaluations per bracket"
    ref = oracle.solve(h, vp, vs, rho, per, win, 4, 0, tol=1e-12, group=2)
    skip = np.zeros(ref["root_count"].shape, bool)
    for key in named:
        i, t = map(int, key.split(","))
        skip[i, t] = True
        np.testing.assert_allclose(vel[i, :, t], np.array(named[key]["roots"]), rtol=RTOL)
    got = dict(phase_velocity=vel, trace_flags=flags, search_level=lev, root_count=cnt,
               group_velocity=ex["group_velocity"], group_velocity_flags=ex["group_velocity_flags"])
    rest = dict(ref, root_count=np.where(skip, -1, ref["root_count"]))
    assert_parity(got, rest, label="deep stacks")


def test_ragged_uneven_batch(md):
    """Ragged models (3..40 finite layers, low-velocity zones, per-model
    period and mode counts) through the ragged entry point."""
    from modisp_amd import workloads
    wl = workloads.uneven_batch(B=24, seed=105)
    rows, row_off, per, per_off, modes, bounds = [], [0], [], [0], [], []
    for i in range(wl.B):
        n = wl.layers[i]
        rows.append(np.stack([a[i, :n] for a in (wl.thickness, wl.vp, wl.vs, wl.rho)]))
        row_off.append(row_off[-1] + n)
        p = wl.periods[i][::3]
        per.append(p)
        per_off.append(per_off[-1] + p.size)
        modes.append(int(wl.nmodes[i]))
        bounds.append([0.8 * wl.vs[i, :n].min(), np.nextafter(wl.vs[i, n - 1], 0.)])
    packed = np.concatenate(rows, axis=1)
    with md.QeSolver(wl.B, int(max(wl.layers)), 1, 1, group_velocity=True) as s:
        out = s.run_ragged(*packed, row_off, np.concatenate(per), per_off, modes, np.array(bounds), 1e-12, 0)
    cell = 0
    for i in range(wl.B):
        n, T, M = wl.layers[i], per_off[i + 1] - per_off[i], modes[i]
        ref = oracle.solve(*(packed[:, row_off[i]:row_off[i + 1]]), per[i], np.array(bounds[i])[None], M, 0,
                           tol=1e-12, group=2)
        got = dict(root_count=out["counts"][per_off[i]:per_off[i + 1]][None],
                   phase_velocity=out["velocity"][cell:cell + M * T].reshape(1, M, T),
                   group_velocity=out["group_velocity"][cell:cell + M * T].reshape(1, M, T),
                   group_velocity_flags=out["group_velocity_flags"][cell:cell + M * T].reshape(1, M, T))
        assert_parity(got, ref, label=f"ragged model {i} ({n} rows)")
        cell += M * T


# --- reference fixtures and closed forms --------------------------------------------------------
@pytest.mark.parametrize("path", sorted((GOLDEN / "accuracy_high_cpu_linux").glob("*.npz")), ids=lambda p: p.stem)
def test_reference_fixtures(md, path):
    """Root counts and masks equal the reference's frozen accuracy="high"
    results on every converged trace; velocities agree to the fixture's own
    refinement tolerance (c_tol_rel 1e-6, legacy refinement)."""
    f = np.load(path)
    wave = int(str(f["wave"]) == "love")
    cmin = float(f["cmin"]) if np.isfinite(f["cmin"]) else None
    cmax = float(f["cmax"]) if np.isfinite(f["cmax"]) else None
    h, vp, vs, rho = (np.atleast_2d(f[k]) for k in MODEL)
    b = md.bounds(h, vp, vs, rho, wave, cmin, cmax)
    np.testing.assert_array_equal(b, f["search_bounds"])
    got = gpu_solve(md, h, vp, vs, rho, f["periods"], b, int(f["nmodes"]), wave, "float32", group=False)
    converged = f["root_count"] >= 0
    np.testing.assert_array_equal(got["root_count"][converged], f["root_count"][converged])
    ref = f["phase_velocity"]
    np.testing.assert_array_equal(np.isfinite(got["phase_velocity"])[np.broadcast_to(converged[:, None], ref.shape)],
                                  f["mode_mask"][np.broadcast_to(converged[:, None], ref.shape)])
    ok = np.isfinite(ref)
    np.testing.assert_allclose(got["phase_velocity"][ok], ref[ok], rtol=2e-6, atol=3e-6)


@pytest.mark.parametrize("path", sorted((GOLDEN / "witness").glob("*.npz")), ids=lambda p: p.stem)
def test_independent_witnesses(md, path):
    """Hard close-pair cases resolved by independent FP64 references."""
    f = np.load(path)
    cropped = path.name.startswith("B10000")
    expected = (f["reference"] if cropped else f["reference_hi"]).ravel()
    h, vp, vs, rho = (np.atleast_2d(f[k]) for k in MODEL)
    b = md.bounds(h, vp, vs, rho, 0, .2 if cropped else None)
    got = gpu_solve(md, h, vp, vs, rho, f["periods"], b, expected.size, 0, "float32", group=False)
    np.testing.assert_allclose(got["phase_velocity"][0, :, 0], expected, atol=3e-6, rtol=2e-6)


def test_two_layer_love_closed_form(md):
    from test_oracle_pins import _two_layer_love, LOVE_MODEL
    periods = np.array([.1, .5, 1., 2., 10.])
    b = md.bounds(*(np.atleast_2d(x) for x in LOVE_MODEL), 1)
    got = gpu_solve(md, *LOVE_MODEL, periods, b, 3, 1)
    for i, p in enumerate(periods):
        roots, group = _two_layer_love(p)
        n = roots.size
        np.testing.assert_allclose(got["phase_velocity"][0, :n, i], roots, rtol=1e-10)
        np.testing.assert_allclose(got["group_velocity"][0, :n, i], group, rtol=1e-9)
        assert np.isnan(got["phase_velocity"][0, n:, i]).all()


@pytest.mark.parametrize("wave", [0, 1], ids=["rayleigh", "love"])
def test_group_velocity_next_to_a_shear_speed(md, wave):
    """The group kernel at a root 1e-9 below a layer's shear speed (r -> 0 in
    that layer: attenuation held, small-z series of the wave factors) against
    differences of the dispersion curve, 1e-9 relative."""
    h, vp, vs, rho, period, win = oracle.shear_speed_crossing(wave)
    got = gpu_solve(md, h[None], vp[None], vs[None], rho[None], np.array([period]), np.reshape(win, (1, 2)), 1, wave)
    c, u = got["phase_velocity"][0, 0, 0], got["group_velocity"][0, 0, 0]
    assert abs(vs[1] - c) <= 2e-9 * c
    want = oracle.curve_group_velocity(h, vp, vs, rho, period, win, 0, wave)
    assert abs(u - want) <= 1e-9 * want, (u, want)


def test_group_velocity_of_a_precision_model_at_a_shear_speed(md):
    """Model 337 of the precision workload at T = 35.25 s: its fundamental
    Rayleigh root lies 8.8e-8 below the shear speed of row 8, where the
    closed-form derivatives had put the group velocity 3e-8 off."""
    from modisp_amd import workloads
    wl = workloads.precision_batch()
    m, f = 337, 88
    sub = wl.subset(np.array([m]))
    win = workloads.scan_window(sub.vs)
    per = wl.periods[f:f + 1]
    got = gpu_solve(md, sub.thickness, sub.vp, sub.vs, sub.rho, per, win, 1, 0)
    c, u = got["phase_velocity"][0, 0, 0], got["group_velocity"][0, 0, 0]
    assert abs(sub.vs[0, 8] - c) <= 1e-7 * c
    want = oracle.curve_group_velocity(sub.thickness[0], sub.vp[0], sub.vs[0], sub.rho[0], per[0], win[0], 0, 0)
    assert abs(u - want) <= 1e-9 * want, (u, want)


def test_halfspace_rayleigh_and_empty_families(md):
    from scipy.optimize import brentq
    k = (3.5 / 6.) ** 2
    x = brentq(lambda x: x ** 3 - 8 * x * x + (24 - 16 * k) * x - 16 * (1 - k), 0., 1., xtol=1e-15)
    m = [np.array([[0.]]), np.array([[6.]]), np.array([[3.5]]), np.array([[2.7]])]
    b = md.bounds(*m, 0)
    got = gpu_solve(md, *m, [.1, 1., 10.], b, 1, 0)
    np.testing.assert_allclose(got["phase_velocity"][0, 0], 3.5 * np.sqrt(x), rtol=1e-12)
    np.testing.assert_allclose(got["group_velocity"][0, 0], got["phase_velocity"][0, 0], rtol=1e-9)
    # Love on a half-space has no guided modes (reference level -3, NOT_FOUND).
    love = gpu_solve(md, *m, [.1, 1.], md.bounds(*m, 1), 2, 1)
    assert np.isnan(love["phase_velocity"]).all() and (love["root_count"] == 0).all()
    assert (love["search_level"] == -3).all()
    assert (love["group_velocity_flags"] == 1).all()
    # An empty window (hi <= lo) reports level -4.
    empty = gpu_solve(md, *m, [1.], np.array([[3., 2.]]), 1, 0)
    assert (empty["search_level"] == -4).all() and (empty["root_count"] == 0).all()


@pytest.mark.parametrize("wave", [0, 1], ids=["rayleigh", "love"])
def test_refinement_tiles_without_brackets(md, wave):
    """The refinement takes traces in tiles of 64 and stages the rows of a
    tile's models in LDS; tiles where no trace holds a bracket (here the first
    300 traces: three models with empty windows) stage nothing, and the models
    after them equal the oracle."""
    from modisp_amd import workloads
    wl = workloads.precision_batch(B=8, seed=107)
    win = np.array(workloads.scan_window(wl.vs), dtype=np.float64)
    win[:3] = win[:3, ::-1]
    got = gpu_solve(md, *(getattr(wl, k) for k in MODEL), wl.periods, win, 5, wave)
    ref = oracle.solve(*(getattr(wl, k) for k in MODEL), wl.periods, win, 5, wave, tol=1e-12, group=2)
    assert (got["root_count"][:3] == 0).all() and (got["search_level"][:3] == -4).all()
    assert (got["root_count"][3:] > 0).all()
    assert_parity(got, ref, label=f"wave {wave}")


# --- full-size, size-independent properties ------------------------------------------------------
@pytest.mark.parametrize("wave", [0, 1], ids=["rayleigh", "love"])
def test_full_precision_workload_fp32_and_fp64_paths(md, wave):
    """At the full 16384 x 100 size: the FP32-scan path and the plain FP64 path
    agree on all but a handful of traces, and on every trace where they differ
    (plus a seeded sample of the others) both equal the oracle."""
    from modisp_amd import workloads
    wl = workloads.precision_batch()
    win = workloads.scan_window(wl.vs)
    a = gpu_solve(md, *(getattr(wl, k) for k in MODEL), wl.periods, win, 5, wave, "float32")
    b = gpu_solve(md, *(getattr(wl, k) for k in MODEL), wl.periods, win, 5, wave, "float64")
    assert (a["root_count"] >= 0).all() and (b["root_count"] >= 0).all()
    with np.errstate(invalid="ignore"):
        rel = np.abs(a["phase_velocity"] - b["phase_velocity"]) / b["phase_velocity"]
    differ = ((a["root_count"] != b["root_count"]) | np.any(rel > RTOL, axis=1) |
              np.any(np.isfinite(a["phase_velocity"]) != np.isfinite(b["phase_velocity"]), axis=1))
    assert differ.sum() <= 1e-4 * differ.size, f"{differ.sum()} traces differ"
    rng = np.random.default_rng(7 + wave)
    check = np.argwhere(differ).tolist() + [[int(i), int(f)] for i, f in
                                            zip(rng.integers(0, wl.B, 64), rng.integers(0, wl.periods.size, 64))]
    for i, f in check:
        ref = oracle.solve(*(getattr(wl, k)[i:i + 1] for k in MODEL), wl.periods[f:f + 1], win[i:i + 1], 5, wave,
                           tol=1e-12, group=2)
        for path, got in (("fp32", a), ("fp64", b)):
            one = {k: (v[i:i + 1, ..., f:f + 1] if v is not None and v.ndim == 3 else
                       v[i:i + 1, f:f + 1] if v is not None and v.ndim == 2 else v)
                   for k, v in got.items() if k != "timing"}
            assert_parity(one, ref, label=f"{path} model {i} period {f} wave {wave}")
    v = a["phase_velocity"]
    d = np.diff(v, axis=1)
    assert np.all(d[np.isfinite(d)] > 0)
    lo, hi = win[:, 0, None, None], win[:, 1, None, None]
    fin = np.isfinite(v)
    assert np.all((v >= lo)[fin]) and np.all((v <= hi)[fin])


@pytest.mark.parametrize("method", [0, 1, 2], ids=["legacy", "brent", "chandrupatla"])
def test_refinement_methods_match_the_oracle(md, method):
    """Each refinement the reference offers (QeSolver refinement 0 legacy
    regula falsi, 1 Brent-Dekker, 2 Chandrupatla) against the oracle run with
    the same method, on the precision recipe and on deep random-walk models."""
    from modisp_amd import workloads
    for name, wl, per in (("precision", workloads.precision_batch(B=12, seed=61), None),
                          ("sweep", workloads.sweep_batch(B=6, seed=62), None)):
        per = wl.periods[::4]
        win = workloads.scan_window(wl.vs)
        for wave in (0, 1):
            got = gpu_solve(md, *(getattr(wl, k) for k in MODEL), per, win, 4, wave, refinement=method)
            ref = oracle.solve(*(getattr(wl, k) for k in MODEL), per, win, 4, wave, tol=1e-12, group=2,
                               refinement=method)
            assert_parity(got, ref, label=f"{name} wave {wave} refinement {method}")
