"""FP64 building blocks of the device evaluators, checked on the host from the
same source (secular.hpp): the phase sine/cosine (Cody-Waite reduction by
pi/2, Taylor polynomials) against long-double libm, and the Newton
reciprocal square root from an estimate worse than the hardware's (22 of
24 bits); and the root adjoint's reverse-mode Rayleigh layer derivative
against forward-mode dual numbers and long-double differences."""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_phase_sincos_and_newton_rsq_accuracy(tmp_path):
    exe = tmp_path / "device_math_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", str(HERE / "device_math_check.cpp"), "-o",
                    str(exe)], check=True)
    es, ec, ey, er = (float(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                                        text=True).stdout.split())
    eps = 2.220446049250313e-16
    assert es <= 2 * eps and ec <= 2 * eps, (es, ec)     # absolute, |x| <= 1e4 and near multiples of pi/2
    assert ey <= 2 * eps and er <= 2 * eps, (ey, er)     # relative


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_reverse_mode_layer_derivative(tmp_path):
    """The adjoint's hand-derived reverse mode of one Rayleigh layer
    (rayleigh_layer_adjoint) against forward-mode dual numbers through the
    compound matrix and long-double central differences. The gk and rho
    derivatives lose digits to cancellation in either double evaluation, so
    reverse mode is held to forward mode's own distance from the long-double
    differences; a wrong or dropped factor in any coefficient shows up as an
    O(1) difference."""
    exe = tmp_path / "adjoint_layer_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", str(HERE / "adjoint_layer_check.cpp"), "-o",
                    str(exe)], check=True)
    hk, gr, lt, fd_rev, fd_fwd = (float(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                                                    text=True).stdout.split())
    assert hk <= 1e-7, hk          # h, ka, kb: reverse vs forward
    assert gr <= 5e-6, gr          # gk, rho: reverse vs forward
    assert lt == 0.0, lt           # the carried row l^T C is the same contraction
    assert fd_fwd <= 1e-5 and fd_rev <= 3 * fd_fwd + 1e-9, (fd_rev, fd_fwd)


def _adjoint_stack_cases():
    """Roots of whole stacks (the oracle's, refined to 1e-12): the
    differentiation recipe (21 rows), deep random-walk stacks (31 rows), water
    over an 8-row solid stack, and a root 1e-9 below a layer's shear speed."""
    import numpy as np
    import oracle
    from modisp_amd import workloads
    rng = np.random.default_rng(23)
    cases = []

    def add(h, vp, vs, rho, periods, modes, wave, every=1):
        for i in range(h.shape[0]):
            lo = 0.8 * vs[i][vs[i] > 0].min()
            win = np.array([[lo, vs[i, -1] * (1 - 1e-9)]])
            ref = oracle.solve(h[i:i + 1], vp[i:i + 1], vs[i:i + 1], rho[i:i + 1], periods, win, modes, wave,
                               tol=1e-12)
            for m in range(modes):
                for f in range(0, periods.size, every):
                    c = ref["phase_velocity"][0, m, f]
                    if np.isfinite(c):
                        cases.append((wave, h[i], vp[i], vs[i], rho[i], float(periods[f]), float(c)))

    g = workloads.gradient_batch(B=3, seed=7)
    for wave in (0, 1):
        add(g.thickness, g.vp, g.vs, g.rho, g.periods[::12], 3, wave)
    s = workloads.sweep_batch(B=3, seed=9)
    for wave in (0, 1):
        add(s.thickness, s.vp, s.vs, s.rho, s.periods[::50], 3, wave)
    h, vp, vs, rho = workloads.monotone_models(rng, 3, 8)
    vs[:, 0], vp[:, 0], rho[:, 0], h[:, 0] = 0., 1.5, 1.03, rng.uniform(.5, 3., 3)
    add(h, vp, vs, rho, np.array([1., 4., 12.]), 3, 0)
    for wave in (0, 1):
        hh, pp, ss, rr, period, win = oracle.shear_speed_crossing(wave)
        c = oracle.solve(hh, pp, ss, rr, [period], win, 1, wave, tol=1e-15)["phase_velocity"][0, 0, 0]
        cases.append((wave, hh, pp, ss, rr, float(period), float(c)))
    return cases


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_root_adjoint_whole_stacks_under_sanitizers(tmp_path):
    """The device adjoint's per-root sweeps (adjoint_root.hpp), built for the
    host with AddressSanitizer and UndefinedBehaviorSanitizer, over whole
    stacks of both wave types (water top and a shear-speed crossing
    included): no out-of-bounds row, tape or accumulator access, no undefined
    behaviour, every derivative the oracle has written by the sweep, and each
    column equal to the oracle's implicit phase gradient to 1e-8 of the
    root's largest entry."""
    import numpy as np
    import oracle
    exe = tmp_path / "adjoint_stack_check"
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-ffp-contract=off", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", str(HERE / "adjoint_stack_check.cpp"),
                    "-o", str(exe)], check=True)
    cases = _adjoint_stack_cases()
    lines = [str(len(cases))]
    for wave, h, vp, vs, rho, period, c in cases:
        top = 1 if vs[0] == 0. else 0
        lines.append(f"{wave} {h.size} {top} {2 * np.pi / period!r} {c!r}")
        lines += [" ".join(repr(float(x)) for x in col) for col in (h, vp, vs, rho)]
    run = subprocess.run([str(exe)], input="\n".join(lines) + "\n", capture_output=True, text=True)
    assert run.returncode == 0, run.stderr[-3000:]
    out = [np.array(l.split(), dtype=float) for l in run.stdout.strip().split("\n")]
    assert len(out) == len(cases)
    worst = 0.
    for (wave, h, vp, vs, rho, period, c), got in zip(cases, out):
        L = h.size
        got = got.reshape(4, L).T                                   # [L, 4]
        want = oracle.phase_gradient(h, vp, vs, rho, period, c, wave)
        if vs[0] == 0.:
            want[0, 2] = 0.  # the water layer has no shear speed to differentiate
        untouched = np.isnan(got)
        assert np.all(want[untouched] == 0.), (wave, L, np.argwhere(untouched & (want != 0.)).tolist())
        scale = np.abs(want).max()
        err = np.abs(np.where(untouched, 0., got) - want).max() / scale
        worst = max(worst, err)
        assert err <= 1e-8, (wave, L, period, c, err)
        assert np.all(np.abs(got[:-1, 3][~untouched[:-1, 3]]) > 0.), "a rho derivative is exactly zero"
    print(f"{len(cases)} roots, worst column error {worst:.2e} of the largest entry")
