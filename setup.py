"""Distribution `modisp-rocm`: the import package `modisp_rocm` (directory
`modisp_amd/`) with its prebuilt gfx950 library, `python -m modisp_amd.build`."""
from pathlib import Path

from setuptools import setup

HERE = Path(__file__).resolve().parent
if not (HERE / "modisp_amd" / "libmodisp_rocm.so").exists():
    raise SystemExit("build the gfx950 library first: python -m modisp_amd.build")

setup(
    name="modisp-rocm",
    version="0.5.0",
    description="MI355X (gfx950) ROCm backend for modisp: multimode Rayleigh/Love root search, group velocity, "
                "anchor continuation and adjoint gradients",
    license="BSD-3-Clause",
    python_requires=">=3.10",
    install_requires=["numpy>=1.23"],
    extras_require={"autograd": ["torch"]},
    package_dir={"modisp_rocm": "modisp_amd"},
    packages=["modisp_rocm"],
    package_data={"modisp_rocm": ["libmodisp_rocm.so", "csrc/*.hip", "csrc/*.hpp", "csrc/*.cpp"]},
    entry_points={"console_scripts": ["modisp-rocm=modisp_rocm.cli:main",
                                      "modisp-rocm-benchmark=modisp_rocm.benchmark:main"]},
    zip_safe=False,
)
