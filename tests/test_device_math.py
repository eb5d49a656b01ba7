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
