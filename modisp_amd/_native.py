"""ctypes binding of the modisp-rocm C ABI (include/modisp_rocm.h).

The library is modisp_amd/libmodisp_rocm.so, built in-tree by
modisp_amd/build.py. There is no fallback: if the library or a HIP device is
missing, calls raise instead of computing anything elsewhere.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIBRARY = HERE / "libmodisp_rocm.so"


class BackendUnavailableError(RuntimeError):
    """The ROCm plugin, its library or its device cannot be used
    (reference python/modisp/_loader.py:8-12)."""


class Config(ctypes.Structure):
    _fields_ = [("accuracy", ctypes.c_int), ("close_root_rtol", ctypes.c_double),
                ("single_precision", ctypes.c_int), ("group_velocity", ctypes.c_int),
                ("group_rtol", ctypes.c_double), ("continuation", ctypes.c_int), ("anchor_stride", ctypes.c_int),
                ("nc", ctypes.c_int), ("tau_floor", ctypes.c_double), ("tau_phase", ctypes.c_double),
                ("chunk_traces", ctypes.c_int), ("grid", ctypes.c_int), ("rho", ctypes.c_double),
                ("graph", ctypes.c_int), ("scan_waves", ctypes.c_int), ("consumer_blocks", ctypes.c_int),
                ("refinement", ctypes.c_int), ("shape", ctypes.c_int), ("love_phase", ctypes.c_int),
                ("secular", ctypes.c_int), ("sampling_rtol", ctypes.c_double), ("period_threshold", ctypes.c_double)]


class KernelReport(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "registers_per_thread", "local_bytes_per_thread", "static_shared_bytes", "dynamic_shared_bytes",
        "search_block_threads", "max_active_blocks_per_cu", "configured_blocks_per_cu", "table_capacity_per_model",
        "cache_points_per_trace")]


_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p
_lib = None

# Every symbol declared in include/modisp_rocm.h.
EXPORTS = ("md_rocm_default_config", "md_rocm_create", "md_rocm_free", "md_rocm_run", "md_rocm_run_device",
           "md_rocm_run_ragged", "md_rocm_run_ragged_device", "md_rocm_autotune_ragged_device", "md_rocm_synchronize", "md_rocm_last_timing", "md_rocm_bounds",
           "md_rocm_available", "md_rocm_set_device", "md_rocm_memory_info", "md_rocm_device_name",
           "md_rocm_last_error", "md_rocm_probe_grid", "md_rocm_last_stats", "md_rocm_adjoint",
           "md_rocm_set_stream", "md_rocm_autotune", "md_rocm_qe_kernel_report",
           "md_rocm_love_phase_kernel_report", "md_rocm_continuity_screen", "md_rocm_result_masks",
           "md_rocm_run_device_static", "md_rocm_review_groups")


def lib():
    """Load the library and declare every C-ABI signature (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    # PyTorch ships its own HIP runtime under the same soname. Loading torch
    # first makes the process use that one runtime for both; loading this
    # library first would leave torch unable to initialise its devices.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not LIBRARY.exists():
        raise BackendUnavailableError(
            f"{LIBRARY} is not built; run `python -m modisp_amd.build` (hipcc, gfx950)")
    try:
        L = ctypes.CDLL(str(LIBRARY))
    except OSError as exc:
        raise BackendUnavailableError(f"cannot load {LIBRARY}: {exc}") from exc
    L.md_rocm_default_config.argtypes = [ctypes.POINTER(Config)]
    L.md_rocm_default_config.restype = None
    L.md_rocm_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Config)]
    L.md_rocm_create.restype = _vp
    L.md_rocm_free.argtypes = [_vp]
    L.md_rocm_free.restype = None
    L.md_rocm_run.argtypes = [_vp, _dp, _dp, _dp, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_double,
                              ctypes.c_double, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _u32p, _i32p, _i32p,
                              _dp, _u32p, _u32p, _u64p, ctypes.c_int]
    L.md_rocm_run.restype = ctypes.c_int
    L.md_rocm_run_device.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp,
                                     ctypes.c_double, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp]
    L.md_rocm_run_device.restype = ctypes.c_int
    L.md_rocm_run_device_static.argtypes = L.md_rocm_run_device.argtypes
    L.md_rocm_run_device_static.restype = ctypes.c_int
    L.md_rocm_review_groups.argtypes = [_dp, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, _dp, _u32p, ctypes.POINTER(ctypes.c_int)]
    L.md_rocm_review_groups.restype = ctypes.c_int
    L.md_rocm_run_ragged.argtypes = [_vp, _dp, _dp, _dp, _dp, _i64p, ctypes.c_int, _dp, _i64p, _i32p, _dp,
                                     ctypes.c_double, ctypes.c_int, _dp, _u32p, _i32p, _i32p, _dp, _u32p, _u32p]
    L.md_rocm_run_ragged.restype = ctypes.c_int
    ll = ctypes.c_longlong
    L.md_rocm_run_ragged_device.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, ll, ctypes.c_int, _vp, _vp, ll, _vp, _vp,
                                            ll, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _vp,
                                            _vp, _vp, _vp, _vp, _vp]
    L.md_rocm_run_ragged_device.restype = ctypes.c_int
    L.md_rocm_autotune_ragged_device.argtypes = L.md_rocm_run_ragged_device.argtypes + [
        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.md_rocm_autotune_ragged_device.restype = ctypes.c_int
    L.md_rocm_synchronize.argtypes = [_vp]
    L.md_rocm_synchronize.restype = ctypes.c_int
    L.md_rocm_last_timing.argtypes = [_vp, _dp, ctypes.c_int]
    L.md_rocm_last_timing.restype = ctypes.c_int
    L.md_rocm_adjoint.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp,
                                  _vp, _vp, _vp, _vp]
    L.md_rocm_adjoint.restype = ctypes.c_int
    L.md_rocm_set_stream.argtypes = [_vp, _vp, ctypes.c_int]
    L.md_rocm_autotune.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.c_double,
                                   ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_int)]
    L.md_rocm_autotune.restype = ctypes.c_int
    L.md_rocm_set_stream.restype = ctypes.c_int
    L.md_rocm_last_stats.argtypes = [_vp, _u64p, ctypes.c_int]
    L.md_rocm_last_stats.restype = ctypes.c_int
    L.md_rocm_bounds.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_double, ctypes.c_int, ctypes.c_double, _dp]
    L.md_rocm_bounds.restype = ctypes.c_int
    L.md_rocm_available.argtypes = [ctypes.c_int]
    L.md_rocm_available.restype = ctypes.c_int
    L.md_rocm_set_device.argtypes = [ctypes.c_int]
    L.md_rocm_set_device.restype = ctypes.c_int
    L.md_rocm_memory_info.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    L.md_rocm_memory_info.restype = ctypes.c_int
    L.md_rocm_device_name.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    L.md_rocm_device_name.restype = ctypes.c_int
    L.md_rocm_probe_grid.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
    L.md_rocm_probe_grid.restype = ctypes.c_int
    L.md_rocm_qe_kernel_report.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(Config), ctypes.POINTER(KernelReport)]
    L.md_rocm_qe_kernel_report.restype = ctypes.c_int
    L.md_rocm_love_phase_kernel_report.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(KernelReport)]
    L.md_rocm_love_phase_kernel_report.restype = ctypes.c_int
    L.md_rocm_continuity_screen.argtypes = [_dp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _u32p,
                                            ctypes.c_int, ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_int)]
    L.md_rocm_continuity_screen.restype = ctypes.c_int
    _u8p = ctypes.POINTER(ctypes.c_uint8)
    L.md_rocm_result_masks.argtypes = [_vp, ctypes.c_int, _u32p, _i32p, _dp, _u32p, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_double, _u8p, _u8p, _u8p, _u8p, _u8p]
    L.md_rocm_result_masks.restype = ctypes.c_int
    L.md_rocm_last_error.argtypes = []
    L.md_rocm_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def last_error() -> str:
    msg = lib().md_rocm_last_error()
    return msg.decode() if msg else "modisp-rocm: unknown error"


def check(code: int) -> None:
    """Map C-ABI return codes to the reference's exception types
    (ValueError for arguments, MemoryError for allocation, RuntimeError otherwise)."""
    if code == 0:
        return
    msg = last_error()
    if code == 2:
        raise ValueError(msg)
    if code == 3:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def ptr(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)
