"""Packaging of the plugin: the distribution `modisp-rocm` builds offline into
a wheel whose import package is `modisp_rocm` (the reference's plugin shape,
cuda-plugin/pyproject.toml:1-19 and python/modisp_cuda/__init__.py:1-7: a
package whose load() returns the native module), with the gfx950 library
inside. The installed package loads on a host without a GPU (no device call)."""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "modisp_amd" / "libmodisp_rocm.so"


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def test_wheel_installs_modisp_rocm(tmp_path):
    wheels, site = tmp_path / "wheels", tmp_path / "site"
    build = tmp_path / "build"
    subprocess.run([sys.executable, "-m", "pip", "wheel", "--no-build-isolation", "--no-deps", "-q", "-w", str(wheels),
                    str(ROOT)], check=True, cwd=tmp_path, env={"PIP_NO_INDEX": "1", "PATH": "/usr/bin:/bin",
                                                               "HOME": str(tmp_path), "TMPDIR": str(build.parent)})
    wheel = sorted(wheels.glob("modisp_rocm-0.5.0-*.whl"))
    assert len(wheel) == 1, list(wheels.iterdir())
    subprocess.run([sys.executable, "-m", "pip", "install", "-q", "--no-deps", "--no-index", "--target", str(site),
                    str(wheel[0])], check=True)
    assert (site / "modisp_rocm" / "libmodisp_rocm.so").exists()
    entry = next(site.glob("modisp_rocm-0.5.0.dist-info")) / "entry_points.txt"
    scripts = entry.read_text()
    assert "modisp-rocm = modisp_rocm.cli:main" in scripts
    assert "modisp-rocm-benchmark = modisp_rocm.benchmark:main" in scripts
    probe = ("import sys; sys.path.insert(0, sys.argv[1]); import modisp_rocm; m = modisp_rocm.load(); "
             "assert m is modisp_rocm and m.VERSION == '0.5.0' and m.ABI_VERSION == 9; "
             "assert m.SUPPORTS_FP32 and m.SUPPORTS_GROUP_VELOCITY and m.SUPPORTS_CONTINUATION; "
             "assert callable(m._review_groups) and hasattr(m, 'QeSolver'); "
             "import modisp_rocm.api as api; assert api.Solver().accuracy == 'auto'; print(m.__file__)")
    out = subprocess.run([sys.executable, "-c", probe, str(site)], check=True, capture_output=True, text=True,
                         cwd=tmp_path)
    assert str(site) in out.stdout


@pytest.fixture(autouse=True)
def _no_build_leftovers():
    """setuptools builds in the source tree; remove what the wheel build leaves there."""
    import shutil
    before = {p for p in (ROOT / "build", ROOT / "modisp_rocm.egg-info") if p.exists()}
    yield
    for p in (ROOT / "build", ROOT / "modisp_rocm.egg-info"):
        if p.exists() and p not in before:
            shutil.rmtree(p)
