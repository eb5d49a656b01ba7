"""The oracle: TEST INFRASTRUCTURE ONLY.

A CPU FP64 restatement of the reference FP64 CPU backend (YuanYusung/modisp
0.5.0, ``cpp/cpu/solver.cpp`` and ``cpp/common/*.hpp``), built from
``oracle/md_oracle_solver.cpp`` into ``oracle/liboracle.so`` and loaded here
with ctypes. Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py`` may import this package, and only as the checker: the
product path (``modisp_amd``) never calls it.

Parity pins (see tests/test_oracle_*.py):
  * released Rayleigh/Love secular functions, the high/qe_adaptive search
    ladder, legacy refinement, bounds and flags: bit-identical to the frozen
    reference fixtures ``tests/golden/accuracy_high_cpu_linux/*.npz``;
  * Brent refinement, Love phase counting, FD group velocity: against the
    reference's own headers compiled into ``oracle/_ref`` (when built), the
    two-layer Love closed form, the half-space Rayleigh cubic and witness
    fixtures;
  * analytic group velocity and phase gradients: closed forms (half-space
    U=c, two-layer Love implicit derivative) and Richardson central
    differences of the released evaluator.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force: bool = False) -> Path:
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    target = HERE / "liboracle.so"
    if force or not target.exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "-B" if force else "all", "all"], check=True)
    return target


def build_reference_checker() -> Path | None:
    """Compile oracle/_ref/libmdref.so from the reference headers, if present."""
    ref = Path(os.environ.get("MODISP_REFERENCE", "/root/reference"))
    if not (ref / "cpp" / "common" / "dunkin.hpp").exists():
        return None
    subprocess.run(["make", "-s", "-C", str(HERE), "-B", "ref", f"REF={ref}"], check=True)
    return HERE / "_ref" / "libmdref.so"


def lib():
    global _LIB
    if _LIB is None:
        path = HERE / "liboracle.so"
        if not path.exists():
            build()
        L = ctypes.CDLL(str(path))
        L.mdo_bounds.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, _dp]
        L.mdo_bounds.restype = None
        L.mdo_secular.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp,
                                  ctypes.c_int, _dp, _dp]
        L.mdo_secular.restype = None
        L.mdo_love_phase.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.mdo_love_phase.restype = ctypes.c_double
        L.mdo_refine.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, _ip]
        L.mdo_refine.restype = ctypes.c_double
        L.mdo_love_phase_roots.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_int, ctypes.c_double, _dp, _ip]
        L.mdo_love_phase_roots.restype = ctypes.c_int
        L.mdo_group_fd.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.POINTER(ctypes.c_uint)]
        L.mdo_group_fd.restype = ctypes.c_double
        L.mdo_group_analytic.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double]
        L.mdo_group_analytic.restype = ctypes.c_double
        L.mdo_secular_derivative.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_int]
        L.mdo_secular_derivative.restype = ctypes.c_double
        L.mdo_phase_gradient.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double, _dp]
        L.mdo_phase_gradient.restype = None
        L.mdo_solve.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int,
                                ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                ctypes.c_int, ctypes.c_double, _dp, _u32p, _i32p, _i32p, _dp, _u32p]
        L.mdo_solve.restype = ctypes.c_ulonglong
        _LIB = L
    return _LIB


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a, t=_dp):
    return a.ctypes.data_as(t) if a is not None else None


def _model(h, vp, vs, rho):
    arrays = [_f64(x) for x in (h, vp, vs, rho)]
    arrays = [a[None, :] if a.ndim == 1 else a for a in arrays]
    if any(a.shape != arrays[0].shape for a in arrays):
        raise ValueError("model array shapes must agree")
    return arrays


def bounds(h, vp, vs, rho, wave=0, cmin=None, cmax=None):
    """[B,2] search window (reference cpp/cpu/solver.cpp:781-790)."""
    h, vp, vs, rho = _model(h, vp, vs, rho)
    B, L = h.shape
    out = np.empty((B, 2))
    lib().mdo_bounds(_p(h), _p(vp), _p(vs), _p(rho), B, L, int(wave), int(cmin is None),
                     0. if cmin is None else float(cmin), int(cmax is None), 0. if cmax is None else float(cmax),
                     _p(out))
    return out


def secular(h, vp, vs, rho, period, c, wave=0, shape=False):
    """Released secular function F (and shape feature g) of one model."""
    h, vp, vs, rho = (_f64(x).ravel() for x in (h, vp, vs, rho))
    c = _f64(np.atleast_1d(c))
    f = np.empty(c.size)
    g = np.empty(c.size) if shape else None
    lib().mdo_secular(_p(h), _p(vp), _p(vs), _p(rho), h.size, int(wave), float(period), _p(c), c.size, _p(f), _p(g))
    return (f, g) if shape else f


def love_phase(h, vp, vs, rho, period, c):
    h, vp, vs, rho = (_f64(x).ravel() for x in (h, vp, vs, rho))
    return lib().mdo_love_phase(_p(h), _p(vp), _p(vs), _p(rho), h.size, float(period), float(c))


def love_phase_roots(h, vp, vs, rho, period, low, high, K, tol):
    """Love roots by the unwrapped SH phase: (ok, roots); ok False where the
    reference's phase path fails and it falls back to its scan."""
    h, vp, vs, rho = (_f64(x).ravel() for x in (h, vp, vs, rho))
    roots = np.empty(int(K))
    nr = ctypes.c_int(0)
    ok = lib().mdo_love_phase_roots(_p(h), _p(vp), _p(vs), _p(rho), h.size, float(period), float(low), float(high),
                                    int(K), float(tol), _p(roots), ctypes.byref(nr))
    return bool(ok), roots[:nr.value].copy()


def refine(h, vp, vs, rho, period, a, b, tol, wave=0, method=1):
    h, vp, vs, rho = (_f64(x).ravel() for x in (h, vp, vs, rho))
    ev = ctypes.c_int(0)
    r = lib().mdo_refine(_p(h), _p(vp), _p(vs), _p(rho), h.size, int(wave), float(period), float(a), float(b),
                         float(tol), int(method), ctypes.byref(ev))
    return r, ev.value


def group_fd(h, vp, vs, rho, period, root, lower, upper, rtol=1e-3, wave=0):
    h, vp, vs, rho = (_f64(x).ravel() for x in (h, vp, vs, rho))
    flags = ctypes.c_uint(0)
    u = lib().mdo_group_fd(_p(h), _p(vp), _p(vs), _p(rho), h.size, int(wave), float(period), float(root),
                           float(lower), float(upper), float(rtol), ctypes.byref(flags))
    return u, flags.value


def group_analytic(h, vp, vs, rho, period, c, wave=0):
    h, vp, vs, rho = (_f64(x).ravel() for x in (h, vp, vs, rho))
    return lib().mdo_group_analytic(_p(h), _p(vp), _p(vs), _p(rho), h.size, int(wave), float(period), float(c))


def secular_derivative(h, vp, vs, rho, period, c, which, wave=0):
    """dF/dx with x = k (0), omega (1) or layer parameter 2+4j+{h,vp,vs,rho}."""
    h, vp, vs, rho = (_f64(x).ravel() for x in (h, vp, vs, rho))
    return lib().mdo_secular_derivative(_p(h), _p(vp), _p(vs), _p(rho), h.size, int(wave), float(period), float(c),
                                        int(which))


def phase_gradient(h, vp, vs, rho, period, c, wave=0):
    """[L,4] implicit dc/d(h, vp, vs, rho) at a root."""
    h, vp, vs, rho = (_f64(x).ravel() for x in (h, vp, vs, rho))
    out = np.empty((h.size, 4))
    lib().mdo_phase_gradient(_p(h), _p(vp), _p(vs), _p(rho), h.size, int(wave), float(period), float(c), _p(out))
    return out


def solve(h, vp, vs, rho, periods, search_bounds, nmodes, wave=0, automatic=False, tol=1e-12, accept=1.,
          refinement=1, love_phase=1, close_rtol=.005, threads=None, group=0, group_rtol=1e-3):
    """Native multimode solve; returns the CpuSolver.run-shaped arrays as a dict.

    periods: [nf] shared, or [B, nf] per-model rows. group: 0 none, 1 reference
    finite-difference group velocity, 2 analytic implicit group velocity.
    """
    h, vp, vs, rho = _model(h, vp, vs, rho)
    B, L = h.shape
    per = _f64(periods)
    stride = 0
    if per.ndim == 2:
        if per.shape[0] != B:
            raise ValueError("row periods must be [B, nf]")
        stride = per.shape[1]
        nf = per.shape[1]
    else:
        nf = per.size
    bnd = _f64(search_bounds).reshape(B, 2)
    velocity = np.empty((B, nmodes, nf))
    flags = np.empty((B, nf), dtype=np.uint32)
    levels = np.empty((B, nf), dtype=np.int32)
    counts = np.empty((B, nf), dtype=np.int32)
    gv = np.empty((B, nmodes, nf)) if group else None
    gf = np.empty((B, nmodes, nf), dtype=np.uint32) if group else None
    if threads is None:
        threads = len(os.sched_getaffinity(0))
    ev = lib().mdo_solve(_p(h), _p(vp), _p(vs), _p(rho), B, L, _p(per), nf, stride, _p(bnd), int(nmodes), int(wave),
                         int(bool(automatic)), float(tol), float(accept), int(refinement), int(love_phase),
                         float(close_rtol), int(threads), int(group), float(group_rtol), _p(velocity),
                         _p(flags, _u32p), _p(levels, _i32p), _p(counts, _i32p), _p(gv), _p(gf, _u32p))
    return dict(phase_velocity=velocity, trace_flags=flags, search_level=levels, root_count=counts,
                group_velocity=gv, group_velocity_flags=gf, evaluations=int(ev), search_bounds=bnd)


TRACE_UNRESOLVED_BRACKET = 1
TRACE_SEARCH_UNCONVERGED = 4
TRACE_F_REJECTED = 128
TRACE_RANGE_TRUNCATED = 256


def reference_result(h, vp, vs, rho, periods, nmodes, wave="rayleigh", cmin=None, cmax=None, tol=1e-6,
                     refinement=0, love_phase=0, threads=None):
    """The reference Solver(accuracy="high")'s result arrays for one call.

    Applies the reference Python layer around CpuSolver.run
    (python/modisp/__init__.py:343, 476-477, 108-123): automatic bounds, the
    range-truncated flag, unconverged counts and status codes.
    """
    w = int(wave == "love")
    h, vp, vs, rho = _model(h, vp, vs, rho)
    b = bounds(h, vp, vs, rho, w, cmin, cmax)
    out = solve(h, vp, vs, rho, periods, b, nmodes, w, automatic=cmin is None, tol=tol, refinement=refinement,
                love_phase=love_phase, threads=threads)
    flags = out["trace_flags"]
    if cmin is not None or cmax is not None:
        flags |= np.uint32(TRACE_RANGE_TRUNCATED)
    counts = out["root_count"]
    counts[(flags & TRACE_SEARCH_UNCONVERGED) != 0] = -1
    velocity = out["phase_velocity"]
    mask = np.isfinite(velocity)
    status = np.where(mask, 0, np.where((flags[:, None, :] & (TRACE_SEARCH_UNCONVERGED | TRACE_UNRESOLVED_BRACKET)) != 0,
                                        2, 1)).astype(np.uint8)
    rejected = (np.isnan(velocity) & ((flags[:, None, :] & TRACE_F_REJECTED) != 0)
                & (np.arange(nmodes)[None, :, None] < counts[:, None, :]))
    status[rejected] = 3
    return dict(phase_velocity=velocity, mode_mask=mask, status=status, trace_flags=flags,
                search_level=out["search_level"], root_count=counts, search_bounds=b)


# --- references for derivative checks --------------------------------------------------------
def curve_group_velocity(h, vp, vs, rho, period, window, mode, wave):
    """U = d(omega)/dk from phase roots at neighbouring frequencies (central
    differences of k(omega) = omega / c, Richardson-extrapolated)."""
    w0 = 2 * np.pi / period

    def k_of(w):
        c = solve(h, vp, vs, rho, [2 * np.pi / w], window, mode + 1, wave, tol=1e-14)["phase_velocity"]
        return w / c[0, mode, 0]
    d = 5e-4 * w0
    d1 = (k_of(w0 + d) - k_of(w0 - d)) / (2 * d)
    d2 = (k_of(w0 + 2 * d) - k_of(w0 - 2 * d)) / (4 * d)
    return 3 / (4 * d1 - d2)


def shear_speed_crossing(wave):
    """A four-row model and the period at which its fundamental mode's phase
    velocity sits 1e-9 below the second layer's shear speed (bisection in
    log T): there r -> 0 in that layer."""
    h = np.array([1.1, 2.3, 4.7, 0.])
    vs = np.array([1.5, 2.2, 2.9, 3.8])
    vp, rho = 1.8 * vs, .32 * 1.8 * vs + .77
    win = bounds(h, vp, vs, rho, wave)
    target = vs[1] * (1 - 1e-9)
    lo, hi = np.log(.5), np.log(50.)
    for _ in range(60):
        mid = .5 * (lo + hi)
        c = solve(h, vp, vs, rho, [np.exp(mid)], win, 1, wave, tol=1e-15)["phase_velocity"][0, 0, 0]
        lo, hi = (mid, hi) if c < target else (lo, mid)
    return h, vp, vs, rho, np.exp(.5 * (lo + hi)), win
