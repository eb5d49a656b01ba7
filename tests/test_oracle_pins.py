"""Pins of the oracle (the CPU restatement of the reference FP64 backend).

Every oracle function used as a checker is pinned here by something other
than itself: the reference's frozen bitwise fixtures, the reference's own
headers compiled into oracle/_ref (when built), closed forms and
brute-force sign scans. CPU only.
"""
import ctypes
import json
from pathlib import Path

import numpy as np
import pytest
from scipy.optimize import brentq

import oracle

GOLDEN = Path(__file__).parent / "golden"
FIELDS = ("phase_velocity", "mode_mask", "status", "trace_flags", "search_level", "root_count", "search_bounds")
MODEL = ("thickness", "vp", "vs", "rho")


# --- the reference's bitwise accuracy="high" fixtures (tests/test_solver.py:78-107) --------------
@pytest.mark.parametrize("path", sorted((GOLDEN / "accuracy_high_cpu_linux").glob("*.npz")), ids=lambda p: p.stem)
def test_oracle_matches_reference_fixture_bitwise(path):
    f = np.load(path)
    cmin = float(f["cmin"]) if np.isfinite(f["cmin"]) else None
    cmax = float(f["cmax"]) if np.isfinite(f["cmax"]) else None
    out = oracle.reference_result(*(f[k] for k in MODEL), f["periods"], int(f["nmodes"]), str(f["wave"]), cmin, cmax,
                                  tol=1e-6, refinement=0, love_phase=0, threads=4)
    for name in FIELDS:
        ref, got = np.asarray(f[name]), np.asarray(out[name])
        assert got.shape == ref.shape, name
        if got.dtype.kind == "f":
            np.testing.assert_array_equal(got.view(np.uint64), ref.view(np.uint64), err_msg=name)
        else:
            np.testing.assert_array_equal(got, ref, err_msg=name)


def test_fixture_manifest_covers_every_case():
    manifest = json.loads((GOLDEN / "accuracy_high_cpu_linux" / "manifest.json").read_text())
    names = {c["file"] for c in manifest["cases"]}
    assert names == {p.name for p in (GOLDEN / "accuracy_high_cpu_linux").glob("*.npz")}


# --- independent witnesses (reference tests/test_api.py:71-83, 286-304) ----------------------------
@pytest.mark.parametrize("path", sorted((GOLDEN / "witness").glob("*.npz")), ids=lambda p: p.stem)
def test_oracle_recovers_independent_witness_roots(path):
    f = np.load(path)
    cropped = path.name.startswith("B10000")
    cmin = .2 if cropped else None
    expected = (f["reference"] if cropped else f["reference_hi"]).ravel()[:3]
    out = oracle.reference_result(*(f[k] for k in MODEL), f["periods"], 3, "rayleigh", cmin, None, tol=1e-6,
                                  refinement=1, love_phase=1, threads=2)
    if out["trace_flags"][0, 0] & oracle.TRACE_SEARCH_UNCONVERGED:
        # As in the reference tests, an unresolved trace must report NaN and
        # UNRESOLVED; the B10000 cases and witness 9384 are allowed to be unresolved.
        assert cropped or path.stem == "witness_model9384_period11"
        assert np.isnan(out["phase_velocity"]).all() and np.all(out["status"] == 2)
        return
    np.testing.assert_allclose(out["phase_velocity"][0, :, 0], expected, atol=3e-6, rtol=2e-6)


# --- closed forms -------------------------------------------------------------------------------
def test_halfspace_rayleigh_cubic():
    vp, vs = 6., 3.5
    k = (vs / vp) ** 2
    x = brentq(lambda x: x ** 3 - 8 * x * x + (24 - 16 * k) * x - 16 * (1 - k), 0., 1., xtol=1e-15)
    b = oracle.bounds([0.], [vp], [vs], [2.7])
    out = oracle.solve([0.], [vp], [vs], [2.7], [10., .1, 1.], b, 1, tol=1e-12)
    np.testing.assert_allclose(out["phase_velocity"][0, 0], vs * np.sqrt(x), rtol=1e-12)


def _two_layer_love(period, nmodes=3):
    """Two-layer Love roots and their analytic implicit group velocity
    (the traction equation, independent of the SH recurrence)."""
    h, b1, b2, r1, r2 = 1., 1.5, 3.5, 2., 2.7
    mu1, mu2 = r1 * b1 * b1, r2 * b2 * b2
    w = 2 * np.pi / period

    def sec(c):
        a = np.sqrt(1 / b1 ** 2 - 1 / c ** 2)
        b = np.sqrt(1 / c ** 2 - 1 / b2 ** 2)
        return mu2 * b * np.cos(w * h * a) - mu1 * a * np.sin(w * h * a)

    grid = np.linspace(np.nextafter(b1, b2), np.nextafter(b2, b1), 32769)
    v = sec(grid)
    idx = np.flatnonzero(np.signbit(v[:-1]) != np.signbit(v[1:]))[:nmodes]
    roots = np.array([brentq(sec, grid[i], grid[i + 1], xtol=1e-15, rtol=1e-15) for i in idx])
    a = np.sqrt(1 / b1 ** 2 - 1 / roots ** 2)
    b = np.sqrt(1 / roots ** 2 - 1 / b2 ** 2)
    da, db = 1 / (roots ** 3 * a), -1 / (roots ** 3 * b)
    z = w * h * a
    dc = mu2 * db * np.cos(z) - mu2 * b * np.sin(z) * w * h * da - mu1 * da * np.sin(z) - mu1 * a * np.cos(z) * w * h * da
    dw = -mu2 * b * np.sin(z) * h * a - mu1 * a * np.cos(z) * h * a
    return roots, roots * roots * dc / (roots * dc + w * dw)


LOVE_MODEL = ([1., 0.], [3., 6.], [1.5, 3.5], [2., 2.7])


@pytest.mark.parametrize("love_phase", [0, 1])
@pytest.mark.parametrize("period", [.1, .5, 1., 2., 10.])
def test_two_layer_love_roots_and_analytic_group(period, love_phase):
    roots, group = _two_layer_love(period)
    b = oracle.bounds(*LOVE_MODEL, wave=1)
    out = oracle.solve(*LOVE_MODEL, [period], b, 3, wave=1, tol=1e-12, love_phase=love_phase, group=2)
    n = roots.size
    np.testing.assert_allclose(out["phase_velocity"][0, :n, 0], roots, rtol=1e-11)
    assert np.isnan(out["phase_velocity"][0, n:, 0]).all()
    np.testing.assert_allclose(out["group_velocity"][0, :n, 0], group, rtol=1e-9)
    # The reference's finite-difference group velocity agrees to its own tolerance.
    fd = oracle.solve(*LOVE_MODEL, [period], b, 3, wave=1, tol=1e-12, love_phase=love_phase, group=1,
                      group_rtol=1e-3)
    np.testing.assert_allclose(fd["group_velocity"][0, :n, 0], group, rtol=1e-3)


def test_thick_water_scholte_limit():
    vf, vp, vs, rf, rs = 1.5, 6., 3.5, 1.02, 2.7

    def sec(c):
        x = (c / vs) ** 2
        A, B, C = np.sqrt(1 - (c / vp) ** 2), np.sqrt(1 - x), np.sqrt(1 - (c / vf) ** 2)
        return 4 * (1 + (vs / vp) ** 2 - (vs / vp) ** 2 * x) / (1 + A * B) - 4 + x + (rf / rs) * x * A / C

    root = brentq(sec, 1e-8, vf * (1 - 1e-10), xtol=1e-14)
    m = ([20., 0.], [vf, vp], [0., vs], [rf, rs])
    out = oracle.solve(*m, [1.], oracle.bounds(*m), 1, automatic=True, tol=1e-12)
    assert abs(out["phase_velocity"][0, 0, 0] - root) < 2e-6


def test_halfspace_group_equals_phase():
    m = ([0.], [6.], [3.5], [2.7])
    out = oracle.solve(*m, [1., 7.], oracle.bounds(*m), 1, automatic=True, tol=1e-12, group=2)
    np.testing.assert_allclose(out["group_velocity"], out["phase_velocity"], rtol=1e-12)


def test_love_ignores_water_depth():
    per = np.logspace(-1, 1, 12)
    dry = ([1., 0.], [3., 6.], [1.5, 3.5], [2., 2.7])
    ref = oracle.solve(*dry, per, oracle.bounds(*dry, wave=1), 3, wave=1, tol=1e-12)["phase_velocity"]
    for h in (.025, 1., 5.):
        wet = ([h, 1., 0.], [1.5, 3., 6.], [0., 1.5, 3.5], [1.02, 2., 2.7])
        got = oracle.solve(*wet, per, oracle.bounds(*wet, wave=1), 3, wave=1, tol=1e-12)["phase_velocity"]
        np.testing.assert_allclose(got, ref, rtol=1e-10)


def test_zero_thickness_padding_is_neutral():
    h, vp, vs, rho = [1.2, 2., 0.], [3., 4.5, 6.5], [1.6, 2.5, 3.7], [2.1, 2.4, 2.8]
    pad = ([1.2, 2., 0., 0., 0.], [3., 4.5, 6.5, 6.5, 6.5], [1.6, 2.5, 3.7, 3.7, 3.7], [2.1, 2.4, 2.8, 2.8, 2.8])
    c = np.linspace(1.3, 3.69, 257)
    for wave in (0, 1):
        a, b = oracle.secular(h, vp, vs, rho, 1.3, c, wave), oracle.secular(*pad, 1.3, c, wave)
        # Identity layers only re-round the normalization: same signs, same values to roundoff.
        np.testing.assert_array_equal(np.signbit(a), np.signbit(b))
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-15)


def test_brute_force_sign_scan_brackets_every_root():
    """Dense FP64 sign scan + brentq finds exactly the oracle's first roots."""
    m = ([1.5, 3., 0.], [3.2, 5., 7.], [1.8, 2.9, 4.], [2.2, 2.5, 2.9])
    period = .8
    b = oracle.bounds(*m)
    c = np.linspace(b[0, 0], b[0, 1], 200001)
    f = oracle.secular(*m, period, c)
    idx = np.flatnonzero(np.signbit(f[:-1]) != np.signbit(f[1:]))[:5]
    brute = [brentq(lambda x: oracle.secular(*m, period, [x])[0], c[i], c[i + 1], xtol=1e-14) for i in idx]
    out = oracle.solve(*m, [period], b, 5, automatic=True, tol=1e-12)
    np.testing.assert_allclose(out["phase_velocity"][0, :len(brute), 0], brute, rtol=1e-11)


# --- analytic derivatives against Richardson central differences ---------------------------------
def _central(fun, x, h):
    d1 = (fun(x + h) - fun(x - h)) / (2 * h)
    d2 = (fun(x + h / 2) - fun(x - h / 2)) / h
    return (4 * d2 - d1) / 3


@pytest.mark.parametrize("wave", [0, 1])
def test_secular_derivatives_match_richardson(wave):
    """At a root the per-layer normalization multiplies F = 0, so d(F/s) = dF/s:
    the dual-number derivatives must equal central differences of the
    released evaluator there, for k, omega and every layer parameter."""
    h = np.array([1.1, 2.3, .7, 0.])
    vs = np.array([1.5, 2.2, 2.9, 3.8])
    vp, rho = 1.8 * vs, .32 * 1.8 * vs + .77
    period = 1.7
    w = 2 * np.pi / period
    b = oracle.bounds(h, vp, vs, rho, wave)
    roots = oracle.solve(h, vp, vs, rho, [period], b, 2, wave, tol=1e-15)["phase_velocity"][0, :, 0]
    for c in roots:
        f_k = lambda k: oracle.secular(h, vp, vs, rho, period, [w / k], wave)[0]
        got = oracle.secular_derivative(h, vp, vs, rho, period, c, 0, wave)
        want = _central(f_k, w / c, 1e-5 * w / c)
        assert abs(got - want) <= 1e-6 * abs(want), (c, got, want)
        f_w = lambda ww: oracle.secular(h, vp, vs, rho, 2 * np.pi / ww, [ww / (w / c)], wave)[0]
        got = oracle.secular_derivative(h, vp, vs, rho, period, c, 1, wave)
        want = _central(f_w, w, 1e-5 * w)
        assert abs(got - want) <= 1e-6 * (abs(want) + 1e-9), (c, got, want)
        scale = max(abs(oracle.secular_derivative(h, vp, vs, rho, period, c, 0, wave)), 1e-12)
        for j in range(4):
            for p, arr in enumerate((h, vp, vs, rho)):
                if p == 0 and j == 3:
                    continue  # the half-space thickness is not a parameter

                def f_par(x, j=j, arr=arr):
                    saved = arr[j]
                    arr[j] = x
                    val = oracle.secular(h, vp, vs, rho, period, [c], wave)[0]
                    arr[j] = saved
                    return val

                got = oracle.secular_derivative(h, vp, vs, rho, period, c, 2 + 4 * j + p, wave)
                want = _central(f_par, arr[j], 1e-5 * abs(arr[j]))
                assert abs(got - want) <= 1e-6 * abs(want) + 1e-10 * scale, (j, p, got, want)


@pytest.mark.parametrize("wave", [0, 1], ids=["rayleigh", "love"])
def test_group_velocity_next_to_a_shear_speed(wave):
    """At a root within 1e-9 of a layer's shear speed the layer's radial
    wavenumber r is ~1e-5 of k: the closed-form derivatives of sin(rh)/r and of
    the evanescent attenuation lose up to 1e-5 of U there. The analytic group
    velocity (attenuation held, small-z series) against differences of the
    dispersion curve itself."""
    h, vp, vs, rho, period, win = oracle.shear_speed_crossing(wave)
    r = oracle.solve(h, vp, vs, rho, [period], win, 1, wave, tol=1e-13, group=2)
    c, u = r["phase_velocity"][0, 0, 0], r["group_velocity"][0, 0, 0]
    assert abs(vs[1] - c) <= 2e-9 * c
    want = oracle.curve_group_velocity(h, vp, vs, rho, period, win, 0, wave)
    assert abs(u - want) <= 1e-10 * want, (u, want)


def test_phase_gradient_matches_root_finite_differences():
    h = np.array([1.4, 2.6, 1.1, 0.])
    vs = np.array([1.6, 2.3, 3.1, 4.0])
    vp, rho = 1.732 * vs, .32 * 1.732 * vs + .77
    period = 3.
    b = oracle.bounds(h, vp, vs, rho)

    def root(h, vp, vs, rho, mode=1):
        return oracle.solve(h, vp, vs, rho, [period], b, 3, tol=1e-14)["phase_velocity"][0, mode, 0]

    c = root(h, vp, vs, rho)
    g = oracle.phase_gradient(h, vp, vs, rho, period, c)
    for j in range(3):
        for p in range(4):
            arrs = [a.copy() for a in (h, vp, vs, rho)]
            x = arrs[p][j]
            step = 1e-5 * x
            arrs[p][j] = x + step
            up = root(*arrs)
            arrs[p][j] = x - step
            dn = root(*arrs)
            fd = (up - dn) / (2 * step)
            assert abs(g[j, p] - fd) <= 1e-6 * max(abs(fd), 1e-3), (j, p, g[j, p], fd)


# --- the reference's own headers, compiled into oracle/_ref -----------------------------------
REF = oracle.HERE / "_ref" / "libmdref.so"


@pytest.mark.skipif(not REF.exists(), reason="reference checker not built (reference tree absent)")
def test_oracle_matches_reference_headers():
    R = ctypes.CDLL(str(REF))
    dp = ctypes.POINTER(ctypes.c_double)
    R.mdr_refine.restype = ctypes.c_double
    R.mdr_group.restype = ctypes.c_double
    R.mdr_love_phase.restype = ctypes.c_double
    from modisp_amd import workloads
    wl = workloads.precision_batch(B=4, seed=17)
    P = lambda a: a.ctypes.data_as(dp)
    for i in range(wl.B):
        h, vp, vs, rho = (np.ascontiguousarray(a[i]) for a in (wl.thickness, wl.vp, wl.vs, wl.rho))
        for wave in (0, 1):
            for T in (.6, 4., 40.):
                b = oracle.bounds(h, vp, vs, rho, wave)[0]
                c = np.linspace(b[0], b[1], 64)
                f, g = oracle.secular(h, vp, vs, rho, T, c, wave, shape=True)
                fr, gr = np.empty_like(c), np.empty_like(c)
                R.mdr_secular(P(h), P(vp), P(vs), P(rho), h.size, wave, 0, ctypes.c_double(T), P(c), c.size, P(fr), P(gr))
                np.testing.assert_array_equal(f.view(np.uint64), fr.view(np.uint64))
                np.testing.assert_array_equal(g.view(np.uint64), gr.view(np.uint64))
                out = oracle.solve(h, vp, vs, rho, [T], b[None], 4, wave, tol=1e-12)
                roots = out["phase_velocity"][0, :, 0]
                roots = roots[np.isfinite(roots)]
                for j, r in enumerate(roots):
                    a_, b_ = r * (1 - 2e-3), r * (1 + 2e-3)
                    for method in (0, 1, 2):   # legacy, Brent, Chandrupatla (refinement.hpp:11-65)
                        mine, ev_m = oracle.refine(h, vp, vs, rho, T, a_, b_, 1e-12, wave, method)
                        ev_r = ctypes.c_int(0)
                        ref = R.mdr_refine(P(h), P(vp), P(vs), P(rho), h.size, wave, ctypes.c_double(T),
                                           ctypes.c_double(a_), ctypes.c_double(b_), ctypes.c_double(1e-12), method,
                                           ctypes.byref(ev_r))
                        assert mine == ref or (np.isnan(mine) and np.isnan(ref)), (method, mine, ref)
                        assert ev_m == ev_r.value, (method, ev_m, ev_r.value)
                    lo = b[0] if j == 0 else .5 * (roots[j - 1] + r)
                    hi = b[1] if j + 1 == roots.size else .5 * (r + roots[j + 1])
                    fl = ctypes.c_uint(0)
                    u = R.mdr_group(P(h), P(vp), P(vs), P(rho), h.size, wave, ctypes.c_double(T), ctypes.c_double(r),
                                    ctypes.c_double(lo), ctypes.c_double(hi), ctypes.c_double(1e-3), ctypes.byref(fl))
                    um, flm = oracle.group_fd(h, vp, vs, rho, T, r, lo, hi, 1e-3, wave)
                    assert (um == u or (np.isnan(um) and np.isnan(u))) and flm == fl.value
                if wave == 1:
                    for K, tol in ((4, 1e-12), (7, 1e-6), (12, 1e-9)):   # love_phase.hpp:76-91
                        ok, mine = oracle.love_phase_roots(h, vp, vs, rho, T, b[0], b[1], K, tol)
                        rr, nr = np.empty(K), ctypes.c_int(0)
                        rok = R.mdr_love_phase_roots(P(h), P(vp), P(vs), P(rho), h.size, ctypes.c_double(T),
                                                     ctypes.c_double(b[0]), ctypes.c_double(b[1]), K,
                                                     ctypes.c_double(tol), P(rr), ctypes.byref(nr))
                        assert ok == bool(rok) and mine.size == nr.value
                        np.testing.assert_array_equal(mine.view(np.uint64), rr[:nr.value].view(np.uint64))
                    for cc in np.linspace(b[0], b[1], 7)[1:-1]:
                        mine = oracle.love_phase(h, vp, vs, rho, T, cc)
                        ref = R.mdr_love_phase(P(h), P(vp), P(vs), P(rho), h.size, ctypes.c_double(T), ctypes.c_double(cc))
                        assert mine == ref or (np.isnan(mine) and np.isnan(ref))
