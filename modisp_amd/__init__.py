"""modisp-rocm: the MI355X (gfx950) backend plugin for modisp-style multimode
surface-wave dispersion.

This package is the `modisp-rocm` distribution's import package. Like the
reference CUDA plugin (cuda-plugin/python/modisp_cuda/__init__.py:1-7) it
exposes ``load()``, returning the native module object the reference's loader
expects (python/modisp/_loader.py:15-28): ``ABI_VERSION``, ``VERSION``, the
``SUPPORTS_*`` capability attributes, ``available``/``set_device``/
``memory_info``/``device_name`` and the ``QeSolver`` engine class. The engine
calls gfx950 HIP kernels through the C ABI in include/modisp_rocm.h.

``modisp_amd.api`` adds the ``Solver`` / ``dispersion_batch`` /
``dispersion`` front end with a ``rocm`` backend entry (auto selection
included), and ``modisp_amd.autograd`` the torch.autograd binding.
"""
from __future__ import annotations

import ctypes
import sys

import numpy as np

from . import _native
from ._native import BackendUnavailableError

__version__ = "0.5.0"
VERSION = "0.5.0"
ABI_VERSION = 9
SUPPORTS_FP32 = True
SUPPORTS_AUTO_ACCURACY = True
SUPPORTS_CANONICAL_LOWER = True
SUPPORTS_ROW_PERIODS = True
SUPPORTS_GROUP_VELOCITY = True
SUPPORTS_CONTINUATION = True
STAT_NAMES = ("traces", "fp64_rechecks", "pair_probes", "fp64_rescans", "love_phase_searches", "scan_nodes",
              "refine_evaluations", "probe_evaluations", "probe_evaluations_fp64", "invalid_models",
              "continuation_targets", "continuation_accepted", "continuation_evaluations")

__all__ = ["load", "QeSolver", "ragged_geometry", "available", "set_device", "memory_info", "device_name", "bounds",
           "BackendUnavailableError", "ABI_VERSION", "VERSION"]


def load():
    """Return the native module object (this package) after loading the library."""
    try:
        _native.lib()
    except BackendUnavailableError as exc:
        raise ImportError(str(exc)) from exc
    return sys.modules[__name__]


def available(device: int = 0) -> bool:
    """True when `device` exists and is a gfx950 (MI355X) GPU."""
    try:
        return bool(_native.lib().md_rocm_available(int(device)))
    except BackendUnavailableError:
        return False


def set_device(device: int) -> None:
    _native.check(_native.lib().md_rocm_set_device(int(device)))


def memory_info():
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    _native.check(_native.lib().md_rocm_memory_info(ctypes.byref(free), ctypes.byref(total)))
    return int(free.value), int(total.value)


def device_name(device: int = 0) -> str:
    buf = ctypes.create_string_buffer(256)
    _native.check(_native.lib().md_rocm_device_name(int(device), buf, 256))
    return buf.value.decode()


def qe_kernel_report(L, accuracy=0, shape=False, single_precision=False, config=None):
    """Compiled resources of the scan kernel that serves models of L rows
    (reference qe_kernel_report, cpp/cuda/qe.cu:1143-1164: same keys; "per
    SM" figures are per CU here). ``accuracy`` and ``shape`` pick among the
    reference's CUDA search kernels; this backend has one scan kernel, FP32
    (``single_precision``) or FP64. ``config`` is a solver's ``config`` (scan
    shape), default the library's."""
    r = _native.KernelReport()
    cfg = ctypes.byref(config) if config is not None else None
    _native.check(_native.lib().md_rocm_qe_kernel_report(int(L), int(accuracy), int(bool(shape)),
                                                         int(bool(single_precision)), cfg, ctypes.byref(r)))
    return dict(registers_per_thread=r.registers_per_thread, static_shared_bytes=r.static_shared_bytes,
                local_bytes_per_thread=r.local_bytes_per_thread, search_block_threads=r.search_block_threads,
                max_active_blocks_per_sm=r.max_active_blocks_per_cu,
                configured_blocks_per_sm=r.configured_blocks_per_cu, dynamic_shared_bytes=r.dynamic_shared_bytes,
                cache_points_per_trace=r.cache_points_per_trace,
                table_capacity_per_model=r.table_capacity_per_model,
                search_kernel="adaptive_scan_fp32" if single_precision else "adaptive_scan_fp64")


def love_phase_kernel_report(L, single_precision=False):
    """Compiled resources of the Love phase search kernel (reference
    love_phase_kernel_report, cpp/cuda/qe.cu:1131-1139: same keys). The phase
    search is FP64 for both scan precisions."""
    r = _native.KernelReport()
    _native.check(_native.lib().md_rocm_love_phase_kernel_report(int(L), int(bool(single_precision)),
                                                                 ctypes.byref(r)))
    return dict(registers_per_thread=r.registers_per_thread, local_bytes_per_thread=r.local_bytes_per_thread,
                static_shared_bytes=r.static_shared_bytes, dynamic_shared_bytes=r.dynamic_shared_bytes,
                search_block_threads=r.search_block_threads)


def continuity_screen(periods, phase, prediction=None, bilateral=True):
    """Continuity flags [B, T] (uint8) of completed phases [B, K, T] and
    whether any trace is suspect (reference _cpu.continuity_screen,
    cpp/cpu/continuity_screen.hpp:8-67): 8 no bilateral neighbours, 16 a mode
    outside [0.95, 1.05] of its neighbouring periods' envelope, 32 FP32 phases
    of adjacent modes within 16 ulp, 64 continuation prediction flagged for
    review. Host computation on the returned phases."""
    per = np.ascontiguousarray(periods, dtype=np.float64)
    ph = np.asarray(phase)
    if per.ndim != 1 or ph.ndim != 3 or ph.shape[2] != per.size:
        raise ValueError("continuity screening requires periods[T] and phases[B,K,T]")
    f32 = ph.dtype == np.float32
    ph = np.ascontiguousarray(ph, dtype=np.float32 if f32 else np.float64)
    B, K, T = ph.shape
    pred = None
    if prediction is not None:
        pred = np.ascontiguousarray(prediction, dtype=np.uint32)
        if pred.shape != (B, T):
            raise ValueError("prediction flags must have shape [B,T]")
    out = np.zeros((B, T), dtype=np.uint8)
    suspect = ctypes.c_int(0)
    _native.check(_native.lib().md_rocm_continuity_screen(
        _native.ptr(per, _native._dp), T, ph.ctypes.data_as(ctypes.c_void_p), int(f32), B, K,
        None if pred is None else _native.ptr(pred, _native._u32p), int(bool(bilateral)),
        out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(suspect)))
    return out, bool(suspect.value)


def result_masks(velocity, flags, counts, close_root_rtol, group=None, group_flags=None):
    """Per-cell masks of a solve (reference DispersionBatch,
    python/modisp/__init__.py:100-123): status [B, modes, T] (uint8: 0 valid,
    1 not found, 2 unresolved, 3 rejected), mode_mask, close_root_mask,
    close_root_risk [B, T] and, with group velocities, the group mask (finite
    with zero flags; else None). One threaded host pass; the same values as the
    reference's numpy expressions."""
    v = np.asarray(velocity)
    f32 = v.dtype == np.float32
    v = np.ascontiguousarray(v, dtype=np.float32 if f32 else np.float64)
    B, M, T = v.shape
    fl = np.ascontiguousarray(flags, dtype=np.uint32)
    ct = np.ascontiguousarray(counts, dtype=np.int32)
    if fl.shape != (B, T) or ct.shape != (B, T):
        raise ValueError("flags and counts must have shape [B,T]")
    g = gf = gm = None
    if group is not None:
        g = np.ascontiguousarray(group, dtype=np.float64)
        gf = np.ascontiguousarray(group_flags, dtype=np.uint32)
        if g.shape != v.shape or gf.shape != v.shape:
            raise ValueError("group velocity and flags must have the velocity's shape")
        gm = np.empty(v.shape, dtype=bool)
    status = np.empty(v.shape, dtype=np.uint8)
    mode_mask = np.empty(v.shape, dtype=bool)
    close = np.empty(v.shape, dtype=bool)
    risk = np.empty((B, T), dtype=bool)
    u8 = ctypes.POINTER(ctypes.c_uint8)
    code = _native.lib().md_rocm_result_masks(
        v.ctypes.data_as(ctypes.c_void_p), int(f32), _native.ptr(fl, _native._u32p), _native.ptr(ct, _native._i32p),
        _native.ptr(g), _native.ptr(gf, _native._u32p), B, M, T, float(close_root_rtol), _native.ptr(status, u8),
        _native.ptr(mode_mask, u8), _native.ptr(close, u8), _native.ptr(risk, u8), _native.ptr(gm, u8))
    if code:
        raise ValueError("invalid result mask arguments")
    return status, mode_mask, close, risk, gm


def _review_groups(packed, layers, modes, wave, secular, period, lo, hi, rtol, close_rtol):
    """Group velocities [modes], their flags and the evaluation count of one
    confirmed trace (packed: the model's h, vp, vs, rho then its ordered
    roots), on the device. Reference plugin function _review_groups
    (cpp/common/review_groups.hpp:39-63)."""
    p = np.asarray(packed)
    if p.ndim != 1:
        raise ValueError("invalid review trace")
    p = np.ascontiguousarray(p, dtype=np.float64)
    group = np.empty(max(1, int(modes)))
    flags = np.empty(max(1, int(modes)), dtype=np.uint32)
    ev = ctypes.c_int(0)
    _native.check(_native.lib().md_rocm_review_groups(
        _native.ptr(p), p.size, int(layers), int(modes), int(wave), int(secular), float(period), float(lo),
        float(hi), float(rtol), float(close_rtol), _native.ptr(group), _native.ptr(flags, _native._u32p),
        ctypes.byref(ev)))
    return group[:modes], flags[:modes], int(ev.value)


def _rows(a, name):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != 2:
        raise ValueError(f"{name} must be [B, L]")
    return a


def bounds(thickness, vp, vs, rho, wave=0, low=None, high=None):
    """Search window [B, 2] (reference _cpu.bounds, cpp/cpu/solver.cpp:781-790)."""
    h, a, b, r = (_rows(x, n) for x, n in zip((thickness, vp, vs, rho), ("thickness", "vp", "vs", "rho")))
    B, L = h.shape
    out = np.empty((B, 2))
    code = _native.lib().md_rocm_bounds(_native.ptr(h), _native.ptr(a), _native.ptr(b), _native.ptr(r), B, L,
                                        int(wave), int(low is None), 0. if low is None else float(low),
                                        int(high is None), 0. if high is None else float(high), _native.ptr(out))
    if code:
        raise ValueError("invalid bounds geometry")
    return out


def probe_grid(thickness, vp, vs, rho, periods, bounds, nc=512, wave=0, fp32=True):
    """F on each (model, period) scan grid with the device evaluators: [B, nf, nc]."""
    h, a, b, r = (_rows(x, n) for x, n in zip((thickness, vp, vs, rho), ("thickness", "vp", "vs", "rho")))
    B, L = h.shape
    per = np.ascontiguousarray(periods, dtype=np.float64).ravel()
    bnd = np.ascontiguousarray(bounds, dtype=np.float64).reshape(B, 2)
    out = np.empty((B, per.size, int(nc)))
    _native.check(_native.lib().md_rocm_probe_grid(_native.ptr(h), _native.ptr(a), _native.ptr(b), _native.ptr(r), B,
                                                   L, _native.ptr(per), per.size, _native.ptr(bnd), int(nc),
                                                   int(wave), int(bool(fp32)), _native.ptr(out)))
    return out


def ragged_geometry(row_offsets, period_offsets, modes):
    """Checks the host offsets of a ragged batch and returns (geometry,
    cell_offsets): geometry holds B, nrows, ntraces, ncells, max_rows and
    max_modes for run_ragged_device_batch, cell_offsets [B+1] (int64) the
    packed velocity cells of each model (modes[m] x periods of m)."""
    ro = np.asarray(row_offsets, dtype=np.int64)
    po = np.asarray(period_offsets, dtype=np.int64)
    md = np.asarray(modes, dtype=np.int64)
    B = md.size
    if B < 1 or ro.shape != (B + 1,) or po.shape != (B + 1,) or ro[0] != 0 or po[0] != 0:
        raise ValueError("row_offsets and period_offsets need B + 1 entries starting at 0")
    rows, T = np.diff(ro), np.diff(po)
    if rows.min() < 1 or T.min() < 0 or md.min() < 1 or md.max() > 256:
        raise ValueError("every model needs >= 1 row, >= 0 periods and 1..256 modes")
    cell = np.concatenate([[0], np.cumsum(md * T)]).astype(np.int64)
    geometry = dict(B=int(B), nrows=int(ro[-1]), ntraces=int(po[-1]), ncells=int(cell[-1]), max_rows=int(rows.max()),
                    max_modes=int(md.max()))
    return geometry, cell


class QeSolver:
    """Persistent native engine; the constructor and run() mirror the reference
    QeSolver (cpp/cuda/qe.cu:1166 signature, run at qe.cu:1060-1118).

    Reference arguments and what they do here:

    * ``single_precision``: the evaluation precision of the grid scan and the
      dtype of returned velocities, as in the reference (True: FP32 scan with
      FP64 recheck, refinement and group velocity, float32 velocities;
      False: the plain FP64 path, float64 velocities). ``scan`` overrides the
      scan precision alone (``scan="float32"`` with float64 velocities is the
      FP32-scan + FP64 product path).
    * ``accuracy`` (0 high, 1 balanced, 2 fast, 4 auto) and ``shape``: every
      preset runs the same exhaustive grid scan; together they set the
      reported search_level codes and TRACE_APPROXIMATE_SEARCH as the
      reference's searches report them (see include/modisp_rocm.h).
    * ``refinement`` (0 legacy, 1 Brent, 2 Chandrupatla) and ``love_phase``
      (1 exact phase count, 0 scan only) select those methods.
    * ``secular``: 1 (the contracted form, the only one built); 0 raises.
    * ``group_rtol``: analytic group velocities whose rounding bound exceeds
      0.25 group_rtol max(c, |U|) are flagged DERIVATIVE_UNSTABLE.
    * ``sampling_rtol`` and ``period_threshold``: validated and reported; the
      reference's shape sampling they tune is not this backend's scan.

    Extra keyword-only arguments select this backend's grid: ``grid``
    ("adaptive": nodes uniform in the WKB mode count with ``rho`` nodes per
    mode over a 64-point floor, the default; "uniform": ``nc`` equally spaced
    nodes) and the FP32 uncertainty band (``tau_floor``, ``tau_phase``).
    """

    def __init__(self, B, L, nf, nmodes, accuracy=0, period_threshold=0., close_root_rtol=.005, shape=True,
                 sampling_rtol=.001, secular=1, refinement=1, love_phase=1, single_precision=False,
                 group_velocity=False, group_rtol=.001, continuation=False, anchor_stride=4, *, scan=None,
                 grid="adaptive", rho=8., nc=512, tau_floor=None, tau_phase=None, chunk_traces=0, graph=False,
                 scan_waves=0, consumer_blocks=0):
        if grid not in ("adaptive", "uniform"):
            raise ValueError("grid must be adaptive or uniform")
        if scan is None:
            scan = "float32" if single_precision else "float64"
        if scan not in ("float32", "float64"):
            raise ValueError("scan must be float32 or float64")
        cfg = _native.Config()
        L_ = _native.lib()
        L_.md_rocm_default_config(ctypes.byref(cfg))
        cfg.accuracy = int(accuracy)
        cfg.close_root_rtol = float(close_root_rtol)
        cfg.single_precision = int(scan == "float32")
        cfg.group_velocity = int(bool(group_velocity))
        cfg.group_rtol = float(group_rtol)
        cfg.continuation = int(bool(continuation))
        cfg.anchor_stride = int(anchor_stride)
        cfg.nc = int(nc)
        if tau_floor is not None:
            cfg.tau_floor = float(tau_floor)
        if tau_phase is not None:
            cfg.tau_phase = float(tau_phase)
        cfg.chunk_traces = int(chunk_traces)
        cfg.grid = int(grid == "adaptive")
        cfg.rho = float(rho)
        cfg.graph = int(bool(graph))
        cfg.scan_waves = int(scan_waves)
        cfg.consumer_blocks = int(consumer_blocks)
        cfg.refinement = int(refinement)
        cfg.shape = int(bool(shape))
        cfg.love_phase = int(love_phase)
        cfg.secular = int(secular)
        cfg.sampling_rtol = float(sampling_rtol)
        cfg.period_threshold = float(period_threshold)
        self.capacity, self.L, self.nf, self.nmodes = int(B), int(L), int(nf), int(nmodes)
        self.single_precision = bool(single_precision)
        self.scan = scan
        self.group_velocity = bool(group_velocity)
        self.continuation = bool(continuation)
        self.config = cfg
        handle = L_.md_rocm_create(self.capacity, self.L, self.nf, self.nmodes, ctypes.byref(cfg))
        if not handle:
            msg = _native.last_error()
            raise (ValueError if "invalid" in msg or "must" in msg else RuntimeError)(msg)
        self._handle = ctypes.c_void_p(handle)

    @property
    def handle(self):
        if self._handle is None:
            raise RuntimeError("solver is closed")
        return self._handle

    def run(self, thickness, vp, vs, rho, periods, bounds, c_tol_rel, f_accept, wave, automatic, profile=False,
            canonical_lower=False):
        h, a, b, r = (_rows(x, n) for x, n in zip((thickness, vp, vs, rho), ("thickness", "vp", "vs", "rho")))
        B, L = h.shape
        if any(x.shape != (B, L) for x in (a, b, r)):
            raise ValueError("model shapes differ")
        if B > self.capacity or L != self.L:
            raise ValueError(f"model batch [{B},{L}] exceeds solver geometry [{self.capacity},{self.L}]")
        per = np.ascontiguousarray(periods, dtype=np.float64)
        row = 0
        if per.ndim == 2:
            if per.shape[0] != B or per.shape[1] != self.nf:
                raise ValueError("row periods must be [B, nf]")
            row = 1
        elif per.ndim != 1 or per.size != self.nf:
            raise ValueError("periods must have nf entries")
        bnd = np.ascontiguousarray(bounds, dtype=np.float64).reshape(B, 2)
        M, T = self.nmodes, self.nf
        vel = np.empty((B, M, T), dtype=np.float32 if self.single_precision else np.float64)
        flags = np.empty((B, T), dtype=np.uint32)
        levels = np.empty((B, T), dtype=np.int32)
        counts = np.empty((B, T), dtype=np.int32)
        gv = np.empty((B, M, T)) if self.group_velocity else None
        gf = np.empty((B, M, T), dtype=np.uint32) if self.group_velocity else None
        stats = np.zeros(len(STAT_NAMES), dtype=np.uint64)
        pred = np.empty((B, T), dtype=np.uint32) if self.continuation else None
        code = _native.lib().md_rocm_run(
            self.handle, _native.ptr(h), _native.ptr(a), _native.ptr(b), _native.ptr(r), B, _native.ptr(per), row,
            _native.ptr(bnd), float(c_tol_rel), float(f_accept), int(wave), int(bool(automatic)),
            vel.ctypes.data_as(ctypes.c_void_p), int(self.single_precision), _native.ptr(flags, _native._u32p),
            _native.ptr(levels, _native._i32p), _native.ptr(counts, _native._i32p), _native.ptr(gv),
            _native.ptr(gf, _native._u32p), _native.ptr(pred, _native._u32p), _native.ptr(stats, _native._u64p),
            len(STAT_NAMES))
        _native.check(code)
        st = {k: int(v) for k, v in zip(STAT_NAMES, stats)}
        if profile:
            st["timings_s"] = {"device_total": self.last_timing()[0] * 1e-3}
        extras = {"group_velocity": gv, "group_velocity_flags": gf, "prediction_flags": pred}
        return vel, flags, st, levels, counts, extras

    def run_ragged(self, thickness, vp, vs, rho, row_offsets, periods, period_offsets, modes, bounds, c_tol_rel,
                   wave):
        """Ragged batch: packed rows/periods with offsets; outputs packed per model."""
        h, a, b, r = (np.ascontiguousarray(x, dtype=np.float64) for x in (thickness, vp, vs, rho))
        ro = np.ascontiguousarray(row_offsets, dtype=np.int64)
        po = np.ascontiguousarray(period_offsets, dtype=np.int64)
        md = np.ascontiguousarray(modes, dtype=np.int32)
        per = np.ascontiguousarray(periods, dtype=np.float64)
        B = md.size
        bnd = np.ascontiguousarray(bounds, dtype=np.float64).reshape(B, 2)
        T = np.diff(po)
        ncell = int(np.sum(md.astype(np.int64) * T))
        ntr = int(po[-1])
        vel = np.empty(ncell)
        flags = np.empty(ntr, dtype=np.uint32)
        levels = np.empty(ntr, dtype=np.int32)
        counts = np.empty(ntr, dtype=np.int32)
        gv = np.empty(ncell) if self.group_velocity else None
        gf = np.empty(ncell, dtype=np.uint32) if self.group_velocity else None
        pred = np.empty(ntr, dtype=np.uint32)
        code = _native.lib().md_rocm_run_ragged(
            self.handle, _native.ptr(h), _native.ptr(a), _native.ptr(b), _native.ptr(r), _native.ptr(ro, _native._i64p),
            B, _native.ptr(per), _native.ptr(po, _native._i64p), _native.ptr(md, _native._i32p), _native.ptr(bnd),
            float(c_tol_rel), int(wave), _native.ptr(vel), _native.ptr(flags, _native._u32p),
            _native.ptr(levels, _native._i32p), _native.ptr(counts, _native._i32p), _native.ptr(gv),
            _native.ptr(gf, _native._u32p), _native.ptr(pred, _native._u32p))
        _native.check(code)
        return dict(velocity=vel, flags=flags, levels=levels, counts=counts, group_velocity=gv,
                    group_velocity_flags=gf, prediction_flags=pred)

    _RAGGED_KEYS = ("thickness", "vp", "vs", "rho", "row_offsets", "periods", "period_offsets", "modes",
                    "cell_offsets", "bounds", "velocity", "flags", "levels", "counts")

    def _ragged_args(self, geometry, pointers, c_tol_rel, wave):
        g, p = geometry, pointers
        missing = [k for k in self._RAGGED_KEYS if not p.get(k)]
        if missing:
            raise ValueError(f"device pointers missing: {missing}")
        return (self.handle, p["thickness"], p["vp"], p["vs"], p["rho"], p["row_offsets"], int(g["nrows"]), int(g["B"]),
                p["periods"], p["period_offsets"], int(g["ntraces"]), p["modes"], p["cell_offsets"], int(g["ncells"]),
                p["bounds"], int(g["max_rows"]), int(g["max_modes"]), float(c_tol_rel), int(wave), p["velocity"],
                p["flags"], p["levels"], p["counts"], p.get("group"), p.get("group_flags"))

    def run_ragged_device_batch(self, geometry, pointers, c_tol_rel, wave):
        """Ragged batch resident on the device (md_rocm_run_ragged_device).
        geometry: B, nrows, ntraces, ncells, max_rows, max_modes (see
        ragged_geometry); pointers: device addresses of thickness, vp, vs, rho
        [nrows], row_offsets, period_offsets, cell_offsets [B+1] (int64), modes
        [B] (int32), periods [ntraces], bounds [B, 2], velocity [ncells],
        flags, levels, counts [ntraces][, group, group_flags [ncells]].
        With graph=True repeated calls replay one captured hipGraph."""
        _native.check(_native.lib().md_rocm_run_ragged_device(*self._ragged_args(geometry, pointers, c_tol_rel, wave)))

    def autotune_ragged_device_batch(self, geometry, pointers, c_tol_rel, wave):
        """autotune_device_batch for a device ragged batch; returns (traces per
        scan block, consumer blocks per CU)."""
        w, b = ctypes.c_int(0), ctypes.c_int(0)
        _native.check(_native.lib().md_rocm_autotune_ragged_device(
            *self._ragged_args(geometry, pointers, c_tol_rel, wave), ctypes.byref(w), ctypes.byref(b)))
        return w.value, b.value

    def run_device_batch(self, B, pointers, c_tol_rel, wave, row_periods=False):
        """pointers: dict of device addresses (int) for thickness, vp, vs, rho,
        periods, bounds, velocity, flags, levels, counts[, group, group_flags]."""
        p = pointers
        code = _native.lib().md_rocm_run_device(
            self.handle, p["thickness"], p["vp"], p["vs"], p["rho"], int(B), p["periods"], int(bool(row_periods)),
            p["bounds"], float(c_tol_rel), int(wave), p["velocity"], p["flags"], p["levels"], p["counts"],
            p.get("group"), p.get("group_flags"))
        _native.check(code)

    def run_device_static_batch(self, B, pointers, c_tol_rel, wave, row_periods=False):
        """run_device_batch through md_rocm_run_device_static: no allocation,
        every work buffer in the library's fixed per-GPU device array; calls
        that overlap on one GPU race."""
        p = pointers
        code = _native.lib().md_rocm_run_device_static(
            self.handle, p["thickness"], p["vp"], p["vs"], p["rho"], int(B), p["periods"], int(bool(row_periods)),
            p["bounds"], float(c_tol_rel), int(wave), p["velocity"], p["flags"], p["levels"], p["counts"],
            p.get("group"), p.get("group_flags"))
        _native.check(code)

    def autotune_device_batch(self, B, pointers, c_tol_rel, wave, row_periods=False):
        """Time the launch-shape candidates on this device batch (arguments as
        run_device_batch; outputs are overwritten) and keep the fastest.
        Returns (traces per scan block, consumer blocks per CU)."""
        p = pointers
        w, b = ctypes.c_int(0), ctypes.c_int(0)
        _native.check(_native.lib().md_rocm_autotune(
            self.handle, p["thickness"], p["vp"], p["vs"], p["rho"], int(B), p["periods"], int(bool(row_periods)),
            p["bounds"], float(c_tol_rel), int(wave), p["velocity"], p["flags"], p["levels"], p["counts"],
            p.get("group"), p.get("group_flags"), ctypes.byref(w), ctypes.byref(b)))
        return w.value, b.value

    def adjoint_device_batch(self, B, pointers, wave, row_periods=False):
        """Weighted root gradients on the device: pointers holds device addresses
        of thickness, vp, vs, rho, periods, roots, weight (dLoss/dc, layout of
        roots) and grad_thickness, grad_vp, grad_vs, grad_rho ([B, L] outputs):
        grad_p = sum over roots of weight * dc/dp (implicit rule, FP64)."""
        p = pointers
        _native.check(_native.lib().md_rocm_adjoint(
            self.handle, p["thickness"], p["vp"], p["vs"], p["rho"], int(B), p["periods"], int(bool(row_periods)),
            int(wave), p["roots"], p["weight"], p["grad_thickness"], p["grad_vp"], p["grad_vs"], p["grad_rho"]))

    def set_stream(self, stream=None):
        """Issue this solver's work on a HIP stream given as an integer handle
        (e.g. torch.cuda.current_stream().cuda_stream, 0 for the default
        stream); None returns to the solver's own stream."""
        own = stream is None
        _native.check(_native.lib().md_rocm_set_stream(self.handle, None if own else int(stream), int(own)))

    def synchronize(self):
        _native.check(_native.lib().md_rocm_synchronize(self.handle))

    STAGES = ("total", "scan", "rescan", "recheck", "assemble", "refine", "group")

    def last_timing(self):
        """Device milliseconds of the last run per stage (see STAGES)."""
        ms = np.zeros(len(self.STAGES))
        _native.lib().md_rocm_last_timing(self.handle, _native.ptr(ms), len(self.STAGES))
        return ms

    def last_stats(self):
        """Work counters of the last run as a dict keyed by STAT_NAMES."""
        st = np.zeros(len(STAT_NAMES), dtype=np.uint64)
        if _native.lib().md_rocm_last_stats(self.handle, _native.ptr(st, _native._u64p), len(STAT_NAMES)) < 0:
            raise RuntimeError(_native.last_error())
        return {k: int(v) for k, v in zip(STAT_NAMES, st)}

    def close(self):
        if getattr(self, "_handle", None) is not None:
            _native.lib().md_rocm_free(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
