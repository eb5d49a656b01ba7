"""FP64 building blocks of the device evaluators, checked on the host from the
same source (secular.hpp): the phase sine/cosine (Cody-Waite reduction by
pi/2, Taylor polynomials) against long-double libm, and the Newton
reciprocal square root from an estimate worse than the hardware's (22 of
24 bits)."""
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
