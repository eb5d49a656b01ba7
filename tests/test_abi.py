"""The C-ABI library: loads, exports every declared symbol, validates
arguments on the host, and its host-only bounds match the oracle bit for bit.
No GPU compute is issued here."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import modisp_amd
import oracle
from modisp_amd import _native

HEADER = Path(__file__).resolve().parents[1] / "include" / "modisp_rocm.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(md_rocm_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _native.lib()
    names = declared_symbols()
    assert len(names) >= 14
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_native.EXPORTS)


def test_plugin_attributes_match_reference_loader_contract():
    module = modisp_amd.load()
    assert module.ABI_VERSION == 9 and module.VERSION == "0.5.0" and modisp_amd.__version__ == "0.5.0"
    for attr in ("SUPPORTS_FP32", "SUPPORTS_AUTO_ACCURACY", "SUPPORTS_CANONICAL_LOWER", "SUPPORTS_ROW_PERIODS",
                 "SUPPORTS_GROUP_VELOCITY", "SUPPORTS_CONTINUATION", "STAT_NAMES"):
        assert hasattr(module, attr)
    for fn in ("available", "set_device", "memory_info", "device_name", "QeSolver", "qe_kernel_report",
               "love_phase_kernel_report"):
        assert callable(getattr(module, fn))


def test_kernel_reports_validate_the_row_count_on_the_host():
    # reference qe.cu:1132: "L must be 1..256"
    for bad in (0, 257):
        with pytest.raises(ValueError, match="L must be 1..256"):
            modisp_amd.qe_kernel_report(bad)
        with pytest.raises(ValueError, match="L must be 1..256"):
            modisp_amd.love_phase_kernel_report(bad)


def test_invalid_configuration_is_a_value_error_before_any_device_call():
    with pytest.raises(ValueError, match="nc"):
        modisp_amd.QeSolver(4, 5, 3, 2, nc=2)
    with pytest.raises(ValueError, match="rho"):
        modisp_amd.QeSolver(4, 5, 3, 2, rho=.5)
    with pytest.raises(ValueError):
        modisp_amd.QeSolver(4, 300, 3, 2)   # more than 256 rows
    with pytest.raises(ValueError):
        modisp_amd.QeSolver(4, 5, 3, 2, scan="float16")


def test_last_error_is_thread_local_text():
    lib = _native.lib()
    cfg = _native.Config()
    lib.md_rocm_default_config(ctypes.byref(cfg))
    cfg.nc = 2
    assert not lib.md_rocm_create(1, 2, 1, 1, ctypes.byref(cfg))
    assert "nc" in _native.last_error()


def _random_models(rng, B, L, water=False):
    h = rng.uniform(.3, 4., (B, L))
    h[:, -1] = 0.
    vs = np.sort(rng.uniform(.8, 4.5, (B, L)), axis=1)
    vp = vs * rng.uniform(1.6, 2.1, (B, L))
    rho = .32 * vp + .77
    if water:
        vs[:, 0] = 0.
        vp[:, 0] = 1.5
        rho[:, 0] = 1.03
    return h, vp, vs, rho


@pytest.mark.parametrize("wave", [0, 1])
@pytest.mark.parametrize("water", [False, True])
def test_host_bounds_match_oracle_bitwise(wave, water):
    rng = np.random.default_rng(7 + wave + 2 * water)
    h, vp, vs, rho = _random_models(rng, 64, 6, water)
    for low, high in ((None, None), (.3, None), (None, 3.), (.25, 9.)):
        got = modisp_amd.bounds(h, vp, vs, rho, wave, low, high)
        want = oracle.bounds(h, vp, vs, rho, wave, low, high)
        np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))


def test_available_is_false_or_gfx950():
    ok = modisp_amd.available(0)
    assert ok in (True, False)
    if ok:
        assert "gfx950" in modisp_amd.device_name(0)
