"""Front end of the ROCm backend: ``Solver``, ``dispersion_batch``,
``dispersion`` and the ``DispersionBatch`` result, with the reference's
argument names, validation messages, result fields and flag values
(python/modisp/__init__.py:15-139 result types, 142-157 model validation,
160-271 Solver options and backend selection, 277-507 the solve,
572-597 the convenience functions).

Only the ``rocm`` backend lives here: ``backend="auto"`` resolves to it when
a gfx950 device is available and raises ``BackendUnavailableError``
otherwise -- this package has no CPU path. In the reference the same
selection falls through to its CPU backend.

The engine is this package's ``QeSolver`` (FP32 density-adaptive grid scan,
FP64 recheck, FP64 bracket refinement, analytic FP64 group velocity) for every
``accuracy`` setting; ``accuracy`` selects the default refinement tolerance
as in the reference. ``frequency_strategy="anchor_continuation"`` runs the
engine's anchor continuation under the reference's conditions (not at
``accuracy="high"``, at least 8 unique periods, an interior target within
the anchor log-gap limit; python/modisp/__init__.py:324-341) on the sorted
unique periods, and reports prediction flags.
"""
from __future__ import annotations

import threading
import time
from enum import IntEnum, IntFlag

import numpy as np

from . import QeSolver, BackendUnavailableError, available, bounds as _auto_bounds, set_device, STAT_NAMES
from . import continuity_screen, result_masks

TRACE_UNRESOLVED_BRACKET = 1
TRACE_RELOCATED = 2
TRACE_SEARCH_UNCONVERGED = 4
TRACE_WORKSPACE_LIMIT = 8
TRACE_NEAR_CUTOFF = 16
TRACE_NONFINITE = 32
TRACE_REFINEMENT_FAILED = 64
TRACE_F_REJECTED = 128
TRACE_RANGE_TRUNCATED = 256
TRACE_SAMPLING_UNCERTAIN = 512
TRACE_ROOT_CONFLICT = 1024
TRACE_APPROXIMATE_SEARCH = 2048
TRACE_CLOSE_ROOT_RISK = 4096
TRACE_ADAPTIVE_FALLBACK = 8192
TRACE_LOVE_PHASE_FALLBACK = 16384
TRACE_FP64_REPAIR = 32768
TRACE_CONTINUATION_APPROXIMATE = 65536
TRACE_AUTO_RELAXED = 1 << 17
DEFAULT_ACCURACY = "auto"
_APPROXIMATE_ACCURACIES = ("auto", "balanced", "fast")
_REFINEMENT = {"legacy": 0, "brent": 1, "chandrupatla": 2}   # QeSolver codes (reference qe.cu:983)
_C_TOL = {"auto": 1e-5, "high": 1e-6, "balanced": 1e-5, "fast": 1e-4}


class ContinuityFlag(IntFlag):
    """Continuity review bits (reference python/modisp/_continuity.py:6-14)."""
    REVIEWED = 1
    REPAIRED = 2
    UNRESOLVED = 4
    NO_BILATERAL = 8
    ENVELOPE = 16
    CLOSE_FP32 = 32
    PREDICTION = 64
    SEARCH = 128


class SolveStatus(IntEnum):
    VALID = 0
    NOT_FOUND = 1
    UNRESOLVED = 2
    REJECTED = 3


class GroupVelocityFlag(IntFlag):
    MISSING_ROOT = 1
    POLISH_FAILED = 2
    DERIVATIVE_UNSTABLE = 4
    ILL_CONDITIONED = 8
    CLOSE_ROOT = 16
    NONFINITE = 32


class PredictionFlag(IntFlag):
    ANCHOR = 1
    ACCEPTED = 2
    FALLBACK = 4
    BYPASSED = 8
    AUDIT = 16
    AUDIT_FAILED = 32
    SOURCE_SEARCH_UNCERTAIN = 64
    ROOT_COUNT_UNSAFE = 128
    ANCHOR_UNSAFE = 256
    WINDOW_UNSAFE = 512
    BRACKET_FAILED = 1024
    CORRECTION_FAILED = 2048
    REVIEW = 4096


class DispersionBatch:
    """Phase (and group) velocities [B, nmodes, T] with the reference's masks,
    status codes and per-trace diagnostics."""

    def __init__(self, velocity, periods, nmodes, wave, flags, levels, counts, bounds, backend, reason,
                 accuracy=None, close_root_rtol=.005, group_velocity=None, group_velocity_flags=None,
                 prediction_flags=None, group_rtol=None, frequency_strategy="independent",
                 frequency_strategy_used="independent",
                 frequency_strategy_reason="independent frequency search requested", anchor_stride=4,
                 continuity_flags=None, continuity_unresolved=False):
        self.phase_velocity = velocity
        self.precision = velocity.dtype.name
        self.period = periods.copy()
        self.mode = np.arange(nmodes)
        self.wave = wave
        self.accuracy = accuracy
        self.close_root_rtol = float(close_root_rtol)
        self.group_velocity = group_velocity
        self.group_velocity_flags = group_velocity_flags
        (self.status, self.mode_mask, self.close_root_mask, self.close_root_risk,
         self.group_velocity_mask) = result_masks(velocity, flags, counts, close_root_rtol, group_velocity,
                                                  group_velocity_flags)
        self.group_velocity_valid = self.group_velocity_mask
        self.prediction_flags = prediction_flags
        self.continuity_flags = continuity_flags
        self.group_rtol = group_rtol
        self.frequency_strategy = frequency_strategy
        self.frequency_strategy_used = frequency_strategy_used
        self.frequency_strategy_reason = frequency_strategy_reason
        self.anchor_stride = int(anchor_stride)
        self.approximate_search = (flags & (TRACE_APPROXIMATE_SEARCH | TRACE_CONTINUATION_APPROXIMATE)) != 0
        self.mode_index_uncertain = self.approximate_search.copy()
        if continuity_unresolved:
            self.mode_index_uncertain |= (continuity_flags & ContinuityFlag.UNRESOLVED) != 0
        self.potentially_ambiguous = self.mode_index_uncertain | self.close_root_risk
        self.trace_flags = flags
        self.numerical_flags = flags & np.uint32(0xFFFFFFFF ^ (TRACE_RANGE_TRUNCATED | TRACE_AUTO_RELAXED))
        self.search_level = levels
        self.root_count = counts
        self.search_bounds = bounds.copy()
        self.backend_used = backend
        self.backend_reason = reason
        self.mode_convention = ("ascending recovered roots; close-cluster omissions may shift indices"
                                if np.any(self.approximate_search) else "ascending roots in the reported search interval")

    @property
    def n_models(self):
        return self.phase_velocity.shape[0]

    @property
    def n_modes(self):
        return self.phase_velocity.shape[1]

    def __len__(self):
        return self.n_models

    def __repr__(self):
        return (f"<DispersionBatch {self.wave} backend={self.backend_used} B={self.n_models} "
                f"nmodes={self.n_modes} T={self.period.size}>")


def _models(thickness, vp, vs, rho):
    """Reference model validation (python/modisp/__init__.py:142-157)."""
    arrays = []
    for name, value in zip(("thickness", "vp", "vs", "rho"), (thickness, vp, vs, rho)):
        if np.iscomplexobj(value):
            raise ValueError(f"{name} must be real")
        a = np.ascontiguousarray(value, dtype=np.float64)
        if a.ndim == 1:
            a = a[None, :]
        if a.ndim != 2 or not a.size or not np.isfinite(a).all():
            raise ValueError(f"{name} must be a finite nonempty [B,L] or [L] array")
        arrays.append(a)
    h, a, b, r = arrays
    if any(v.shape != h.shape for v in arrays):
        raise ValueError("model array shapes must agree")
    if h.shape[1] > 256:
        raise ValueError("at most 256 layers are supported")
    if np.any(h < 0) or np.any(a <= 0) or np.any(r <= 0) or np.any(b < 0):
        raise ValueError("thickness/vs must be non-negative and vp/rho positive")
    if np.any(b[:, -1] <= 0) or np.any(b[:, 1:] == 0):
        raise ValueError("only a single first water layer may have vs=0; the half-space must be solid")
    solid = b > 0
    if np.any(a[solid] / b[solid] <= np.sqrt(4. / 3.)):
        raise ValueError("solid vp/vs must exceed sqrt(4/3) for positive bulk modulus")
    return arrays


class Solver:
    """Reusable phase dispersion solver on the ROCm backend; use one instance
    per concurrent caller (python/modisp/__init__.py:160 for the option set).

    ``refinement`` selects the bracket refinement (Brent-Dekker, Chandrupatla or
    the legacy regula falsi, as in the reference). The options the reference
    uses to pick among its CPU/CUDA search strategies (search_strategy,
    sampling_rtol, secular, love_strategy, adaptive_period_threshold, threads,
    block, nc) are validated and recorded; this backend has one device search
    for all of them."""

    def __init__(self, nc=512, cmin=None, cmax=None, c_tol_rel=None, f_accept=1., block=256,
                 search_strategy="shape_adaptive", workspace_bytes=256 * 1024 ** 2, chunk_size=None, *,
                 backend="auto", device=0, threads=None, profile=False, accuracy=None,
                 adaptive_period_threshold=None, close_root_rtol=.005, sampling_rtol=None, secular=None,
                 refinement=None, love_strategy="phase", precision="float64", group_velocity=False,
                 group_rtol=.001, frequency_strategy="independent", anchor_stride=4, continuity_check=True):
        if backend not in ("auto", "cpu", "cuda", "rocm"):
            raise ValueError("backend must be auto, cpu, cuda or rocm")
        if search_strategy not in ("shape_adaptive", "qe_adaptive", "uniform"):
            raise ValueError("invalid search_strategy")
        if not isinstance(continuity_check, (bool, np.bool_)):
            raise ValueError("continuity_check must be boolean")
        if not isinstance(group_velocity, (bool, np.bool_)):
            raise ValueError("group_velocity must be boolean")
        if (not np.isscalar(group_rtol) or isinstance(group_rtol, (str, bytes, bool, np.bool_))
                or np.iscomplexobj(group_rtol) or not np.isfinite(group_rtol) or not 0 < group_rtol < 1):
            raise ValueError("group_rtol must be finite and between 0 and 1")
        if frequency_strategy not in ("independent", "anchor_continuation"):
            raise ValueError("frequency_strategy must be independent or anchor_continuation")
        if (not isinstance(anchor_stride, (int, np.integer)) or isinstance(anchor_stride, (bool, np.bool_))
                or anchor_stride < 2):
            raise ValueError("anchor_stride must be an integer of at least 2")
        if not isinstance(precision, str) or precision not in ("float64", "float32"):
            raise ValueError("precision must be float64 or float32")
        if secular not in (None, "dunkin", "fast"):
            raise ValueError("secular must be None, 'dunkin' or 'fast'")
        if refinement is None:
            refinement = "brent"
        if refinement not in ("legacy", "brent", "chandrupatla"):
            raise ValueError("refinement must be legacy, brent or chandrupatla")
        if love_strategy not in ("phase", "scan"):
            raise ValueError("love_strategy must be phase or scan")
        if accuracy is None:
            accuracy = DEFAULT_ACCURACY
        elif accuracy not in ("auto", "high", "balanced", "fast"):
            raise ValueError("accuracy must be auto, high, balanced or fast")
        if not np.isscalar(close_root_rtol) or np.iscomplexobj(close_root_rtol) or not np.isfinite(close_root_rtol) \
                or not 0 < close_root_rtol < 1:
            raise ValueError("close_root_rtol must be finite and between 0 and 1")
        if c_tol_rel is None:
            c_tol_rel = _C_TOL[accuracy]
        for name, v in (("cmin", cmin), ("cmax", cmax)):
            if v is not None and (not np.isfinite(v) or v <= 0):
                raise ValueError(f"{name} must be finite and positive")
        if cmin is not None and cmax is not None and cmax <= cmin:
            raise ValueError("cmax must exceed cmin")
        if not np.isfinite(c_tol_rel) or c_tol_rel <= 0:
            raise ValueError("c_tol_rel must be positive")
        if not np.isfinite(f_accept) or f_accept < 0:
            raise ValueError("f_accept must be non-negative")
        if nc < 3:
            raise ValueError("nc must be at least 3")
        if not isinstance(workspace_bytes, (int, np.integer)) or workspace_bytes < 1:
            raise ValueError("workspace_bytes must be a positive integer")
        if chunk_size is not None and (not isinstance(chunk_size, (int, np.integer)) or chunk_size < 1):
            raise ValueError("chunk_size must be positive")
        if not isinstance(device, (int, np.integer)) or device < 0:
            raise ValueError("device must be a non-negative integer")
        if not isinstance(profile, (bool, np.bool_)):
            raise ValueError("profile must be boolean")
        self.nc, self.cmin, self.cmax = int(nc), cmin, cmax
        self.c_tol_rel, self.f_accept = float(c_tol_rel), float(f_accept)
        self.search_strategy, self.workspace_bytes, self.chunk_size = search_strategy, int(workspace_bytes), chunk_size
        self.backend, self.device, self.threads, self.profile = backend, int(device), threads, bool(profile)
        self.accuracy, self.close_root_rtol = accuracy, float(close_root_rtol)
        self.secular, self.refinement, self.love_strategy = secular, refinement, love_strategy
        self.precision, self.group_velocity, self.group_rtol = precision, bool(group_velocity), float(group_rtol)
        self.frequency_strategy, self.anchor_stride = frequency_strategy, int(anchor_stride)
        self.continuity_check = bool(continuity_check)
        self._engine = None
        self._engine_key = None
        self._calls = 0
        self._last_stats = None
        self.last_result = None
        self._key = None
        self._lock = threading.RLock()

    # -- backend selection (reference __init__.py:255-271) --------------------------------------
    def _select(self):
        if self.backend in ("cpu", "cuda"):
            raise BackendUnavailableError(f"backend '{self.backend}' is not part of the ROCm plugin; "
                                          "use the reference package for it")
        if not available(self.device):
            raise BackendUnavailableError(f"ROCm device {self.device} is unavailable (a gfx950 GPU is required)")
        try:
            set_device(self.device)
        except RuntimeError as exc:
            raise BackendUnavailableError(f"ROCm device {self.device} initialization failed: {exc}") from exc
        return "rocm", "ROCm plugin and device available"

    def close(self):
        if self._engine is not None:
            self._engine.close()
        self._engine = None
        self._engine_key = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __call__(self, thickness, vp, vs, rho, periods, nmodes=5, wave="rayleigh", names=None):
        if not self._lock.acquire(blocking=False):
            raise RuntimeError("a Solver instance cannot be used concurrently; create one instance per caller")
        try:
            return self._solve(thickness, vp, vs, rho, periods, nmodes, wave)
        finally:
            self._lock.release()

    def _solve(self, thickness, vp, vs, rho, periods, nmodes, wave, _bounds=None, _row_periods=None):
        """One solve. The private options serve the continuity review: explicit
        search bounds [B, 2] and one period per model (`periods` then holds
        that list's first entry, as in the reference's verifier call)."""
        started = time.perf_counter()
        if wave not in ("rayleigh", "love"):
            raise ValueError("wave must be rayleigh or love")
        if not isinstance(nmodes, (int, np.integer)) or not 1 <= nmodes <= 256:
            raise ValueError("nmodes must be 1..256")
        h, a, b, r = _models(thickness, vp, vs, rho)
        B, L = h.shape
        if np.iscomplexobj(periods):
            raise ValueError("periods must be real")
        per = np.ascontiguousarray(periods, dtype=np.float64)
        if per.ndim != 1 or not per.size or not np.isfinite(per).all() or np.any(per <= 0):
            raise ValueError("periods must be a finite positive 1-D array")
        backend, reason = self._select()
        input_periods = per
        unique_periods = per
        inverse = None
        # Unsorted or repeated periods are solved once each (reference 314-341, restored at 463-475);
        # continuation needs them sorted.
        if per.size > 1 and not np.all(per[1:] > per[:-1]):
            needs_sorted = self.frequency_strategy == "anchor_continuation" and self.accuracy != "high" \
                and per.size >= 8
            if needs_sorted or not np.all(per[1:] < per[:-1]):
                unique_periods, inverse = np.unique(per, return_inverse=True)
        continuation = False
        frequency_reason = "independent frequency search requested"
        if self.frequency_strategy == "anchor_continuation":
            if self.accuracy == "high":
                frequency_reason = "high accuracy retains independent search"
            elif unique_periods.size < 8:
                frequency_reason = "fewer than 8 unique periods; using independent search"
            elif not np.any(np.log(unique_periods[2:]) - np.log(unique_periods[:-2]) <= .25):
                frequency_reason = ("period grid has no interior target within the anchor log-gap limit; "
                                    "using independent search")
            else:
                continuation = True
                frequency_reason = "ROCM anchor continuation on sorted unique periods"
        if continuation or unique_periods.size < per.size:
            per = unique_periods
        else:
            inverse = None  # keep the caller's order for independent, all-unique grids
        frequency_used = "anchor_continuation" if continuation else "independent"
        wave_code = int(wave == "love")
        bnd = (_auto_bounds(h, a, b, r, wave_code, self.cmin, self.cmax) if _bounds is None
               else np.ascontiguousarray(_bounds, dtype=np.float64))
        if not np.isfinite(bnd).all() or np.any(bnd[:, 0] <= 0):
            raise ValueError("automatic search range could not be constructed")
        row_periods = None
        if _row_periods is not None:
            row_periods = np.ascontiguousarray(_row_periods, dtype=np.float64).reshape(B, 1)
            per, inverse, continuation = per[:1], None, False
            frequency_used, frequency_reason = "independent", "independent frequency search requested"
        capacity = min(B, self.chunk_size or B)
        key = (capacity, L, per.size, int(nmodes), self.precision, self.group_velocity, self.group_rtol,
               self.close_root_rtol, continuation, self.anchor_stride, self.refinement)
        allocated = self._engine_key is None or self._engine_key[1:] != key[1:] or self._engine_key[0] < capacity
        if allocated:
            self.close()
            self._engine = QeSolver(capacity, L, per.size, int(nmodes), close_root_rtol=self.close_root_rtol,
                                    single_precision=self.precision == "float32",
                                    group_velocity=self.group_velocity, group_rtol=self.group_rtol,
                                    continuation=continuation, anchor_stride=self.anchor_stride,
                                    refinement=_REFINEMENT[self.refinement])
            self._engine_key = key
        else:
            capacity = self._engine_key[0]
        # One chunk: the engine's output arrays become the result as they are.
        single = capacity >= B
        if not single:
            velocity = np.empty((B, nmodes, per.size), dtype=self.precision)
            flags = np.empty((B, per.size), dtype=np.uint32)
            levels = np.empty((B, per.size), dtype=np.int32)
            counts = np.empty_like(levels)
        group = gflags = None
        if self.group_velocity and not single:
            group = np.empty(velocity.shape, dtype=np.float64)
            gflags = np.empty(velocity.shape, dtype=np.uint32)
        prediction = (np.full((B, per.size), int(PredictionFlag.BYPASSED), dtype=np.uint32)
                      if self.frequency_strategy == "anchor_continuation" else None)
        stats = {k: 0 for k in STAT_NAMES}
        chunks = 0
        for start in range(0, B, capacity):
            end = min(B, start + capacity)
            chunk_periods = per if row_periods is None else row_periods[start:end]
            out = self._engine.run(h[start:end], a[start:end], b[start:end], r[start:end], chunk_periods,
                                   bnd[start:end], self.c_tol_rel, self.f_accept, wave_code, self.cmin is None,
                                   self.profile)
            if single:
                velocity, flags, part, levels, counts, extras = out
                if self.group_velocity:
                    group, gflags = extras["group_velocity"], extras["group_velocity_flags"]
            else:
                velocity[start:end], flags[start:end], part, levels[start:end], counts[start:end], extras = out
                if self.group_velocity:
                    group[start:end] = extras["group_velocity"]
                    gflags[start:end] = extras["group_velocity_flags"]
            if continuation:
                chunk_prediction = extras["prediction_flags"]
                prediction[start:end] = chunk_prediction
                for name, bit in (("root_count_unsafe", 128), ("anchor_unsafe", 256), ("window_unsafe", 512),
                                  ("bracket_failed", 1024), ("correction_failed", 2048), ("audit_failed", 32)):
                    counter = "continuation_fallback_" + name
                    stats[counter] = stats.get(counter, 0) + int(np.count_nonzero(chunk_prediction & bit))
                stats["continuation_accepted"] = stats.get("continuation_accepted", 0) + int(
                    np.count_nonzero(chunk_prediction & int(PredictionFlag.ACCEPTED)))
                stats["continuation_fallbacks"] = stats.get("continuation_fallbacks", 0) + int(
                    np.count_nonzero(chunk_prediction & int(PredictionFlag.FALLBACK)))
            for k, v in part.items():
                stats[k] = stats.get(k, 0) + int(v)
            chunks += 1
        # Continuity screen and FP64 review of suspect traces (reference 445-460).
        continuity = None
        if self.accuracy in _APPROXIMATE_ACCURACIES and (self.continuity_check or continuation
                                                         or self.accuracy == "auto"):
            continuity, suspect = continuity_screen(per, velocity, prediction, self.continuity_check)
            if self.accuracy == "auto":
                failure = (flags & (TRACE_UNRESOLVED_BRACKET | TRACE_SEARCH_UNCONVERGED | TRACE_ROOT_CONFLICT)) != 0
                if np.any(failure):
                    continuity[failure] |= int(ContinuityFlag.SEARCH)
                    suspect = True
            stats.update(continuity_suspected=0, continuity_local_searches=0, continuity_full_searches=0,
                         continuity_repaired=0, continuity_unresolved=0, continuity_evaluations=0)
            if suspect:
                _review(self, (h, a, b, r), per, wave, bnd, velocity, flags, levels, counts, group, gflags,
                        continuity, stats)
        computed_periods = per.size
        if inverse is not None:
            velocity, flags, levels, counts = (x.take(inverse, axis=-1) for x in (velocity, flags, levels, counts))
            if group is not None:
                group, gflags = group.take(inverse, axis=-1), gflags.take(inverse, axis=-1)
            if prediction is not None:
                prediction = prediction.take(inverse, axis=-1)
            if continuity is not None:
                continuity = continuity.take(inverse, axis=-1)
        if self.cmin is not None or self.cmax is not None:
            flags |= TRACE_RANGE_TRUNCATED
        counts[(flags & TRACE_SEARCH_UNCONVERGED) != 0] = -1
        stats.update(backend=backend, precision=self.precision, accuracy=self.accuracy,
                     strategy="rocm_adaptive_scan", requested_strategy=self.search_strategy,
                     refinement=self.refinement, love_strategy=self.love_strategy if wave == "love" else None,
                     close_root_rtol=self.close_root_rtol, refinement_tolerance=self.c_tol_rel, chunks=chunks,
                     capacity_models=int(capacity), engine_allocations=int(allocated),
                     group_velocity=self.group_velocity, group_rtol=self.group_rtol,
                     frequency_strategy=self.frequency_strategy, frequency_strategy_used=frequency_used,
                     frequency_strategy_reason=frequency_reason, anchor_stride=self.anchor_stride,
                     periods_requested=int(input_periods.size), periods_computed=int(computed_periods),
                     close_root_risk_traces=int(np.count_nonzero(flags & TRACE_CLOSE_ROOT_RISK)))
        stats["evaluations"] = sum(int(stats.get(k, 0)) for k in ("scan_nodes", "fp64_rechecks", "refine_evaluations",
                                                                  "probe_evaluations", "probe_evaluations_fp64"))
        attempts = stats.get("continuation_accepted", 0) + stats.get("continuation_fallbacks", 0)
        stats["continuation_local_success_rate"] = (stats.get("continuation_accepted", 0) / attempts
                                                    if attempts else None)
        if self.profile:
            stats["timings_s"] = {"solver_total": time.perf_counter() - started}
        self._last_stats = stats
        self._calls += 1
        result = DispersionBatch(velocity, input_periods, nmodes, wave, flags, levels, counts, bnd, backend, reason,
                                 accuracy=self.accuracy, close_root_rtol=self.close_root_rtol,
                                 group_velocity=group, group_velocity_flags=gflags,
                                 prediction_flags=prediction,
                                 group_rtol=self.group_rtol if self.group_velocity else None,
                                 frequency_strategy=self.frequency_strategy, frequency_strategy_used=frequency_used,
                                 frequency_strategy_reason=frequency_reason, anchor_stride=self.anchor_stride,
                                 continuity_flags=continuity,
                                 continuity_unresolved=stats.get("continuity_unresolved", 0) > 0)
        self._key = (B, L, per.size, self.nc, int(nmodes))
        self.last_result = result
        return result

    @property
    def calls(self):
        return self._calls

    def kernel_report(self):
        """Settings of the last solve (or of this solver) and the compiled
        resources of the scan kernel for its model depth, plus the Love phase
        kernel after a Love solve (reference Solver.kernel_report,
        python/modisp/__init__.py:517-557)."""
        from . import qe_kernel_report, love_phase_kernel_report
        last = self.last_result
        precision = last.precision if last is not None else self.precision
        report = dict(backend="rocm", precision=precision, accuracy=self.accuracy, secular=self.secular,
                      secular_kernel=("sh" if last is not None and last.wave == "love" else "dunkin"),
                      close_root_rtol=self.close_root_rtol, refinement_tolerance=self.c_tol_rel,
                      refinement=self.refinement, love_strategy=self.love_strategy,
                      group_velocity=self.group_velocity, group_rtol=self.group_rtol,
                      frequency_strategy=self.frequency_strategy, anchor_stride=self.anchor_stride)
        if last is not None:
            report.update(frequency_strategy_used=last.frequency_strategy_used,
                          frequency_strategy_reason=last.frequency_strategy_reason)
        if self._key is not None:
            cfg = self._engine.config if self._engine is not None else None
            # The engine scans in FP32 for both output precisions.
            report.update(qe_kernel_report(self._key[1], int(self.accuracy != "high"),
                                           self.search_strategy == "shape_adaptive", True, config=cfg))
            if last is not None and last.wave == "love":
                report["scan_kernel_report"] = dict(report)
                report.update(love_phase_kernel_report(self._key[1], precision == "float32"))
                report["scan_kernel"] = report.get("search_kernel")
                report["search_kernel"] = "love_phase_with_scan_fallback"
        return report

    def stats(self):
        if self._last_stats is None:
            raise RuntimeError("call the solver first")
        return dict(self._last_stats)


def _review(solver, models, periods, wave, bounds, velocity, trace, levels, counts, group, group_flags, flags,
            stats):
    """FP64 review of the traces the continuity screen flagged (reference
    python/modisp/_continuity.py:17-205, its row-period path): every suspect
    (model, period) is re-solved alone by an accuracy="high" solver with the
    requested modes plus two guard roots -- first in a window cut just above
    the observed roots, then, where that is not conclusive, in the full window
    -- and adopted when the verifier's search is clean; a changed result
    replaces the trace (REPAIRED), one the verifier cannot confirm is marked
    UNRESOLVED, and every suspect is REVIEWED."""
    reason_bits = int(ContinuityFlag.ENVELOPE | ContinuityFlag.CLOSE_FP32 | ContinuityFlag.PREDICTION
                      | ContinuityFlag.SEARCH)
    model_ids, freq_ids = np.nonzero(flags & reason_bits)
    n = len(model_ids)
    stats["continuity_suspected"] = n
    if not n:
        return
    modes = velocity.shape[1]
    guards = min(256, modes + 2)
    settings = dict(backend="rocm", device=solver.device, accuracy="high", cmin=solver.cmin, cmax=solver.cmax,
                    c_tol_rel=min(solver.c_tol_rel, 1e-6), f_accept=solver.f_accept, secular=solver.secular,
                    refinement=solver.refinement, group_velocity=solver.group_velocity, group_rtol=solver.group_rtol,
                    close_root_rtol=solver.close_root_rtol, continuity_check=False,
                    workspace_bytes=solver.workspace_bytes, chunk_size=solver.chunk_size)
    with Solver(**settings) as verifier:
        def solve(rows, frequencies, limits, local):
            selected = periods[frequencies]
            stats["continuity_" + ("local_searches" if local else "full_searches")] += len(rows)
            try:
                result = verifier._solve(*(v[rows] for v in models), selected[:1], guards, wave, _bounds=limits,
                                         _row_periods=selected)
            except MemoryError:
                return None
            counters = verifier.stats()
            stats["continuity_rocm_batches"] = stats.get("continuity_rocm_batches", 0) + counters["chunks"]
            stats["continuity_evaluations"] += counters.get("evaluations", 0)
            return result

        def adopt(result, i, row, f, local):
            if result is None:
                return False
            count = int(result.root_count[i, 0])
            roots = result.phase_velocity[i, :, 0]
            known = min(guards, count)
            if (count < 0 or result.search_level[i, 0] < 0 or result.trace_flags[i, 0]
                    & (TRACE_SEARCH_UNCONVERGED | TRACE_UNRESOLVED_BRACKET | TRACE_F_REJECTED)):
                return False
            if local and count < modes + 2:
                return False
            recovered = roots[:known].astype(velocity.dtype)
            if not np.isfinite(recovered).all() or np.any(recovered[1:] <= recovered[:-1]):
                return False
            if local and (not known or roots[known - 1] >= result.search_bounds[i, 1] * (1 - 4 * verifier.c_tol_rel)):
                return False
            old = velocity[row, :, f]
            new = roots[:modes].astype(old.dtype)
            count = min(count, modes + 2)
            alias_changed = bool(flags[row, f] & int(ContinuityFlag.CLOSE_FP32)) and not np.array_equal(
                old, new, equal_nan=True)
            search_failed = bool(flags[row, f] & int(ContinuityFlag.SEARCH))
            if (search_failed or counts[row, f] != count or alias_changed
                    or not np.allclose(old, new, rtol=solver.c_tol_rel, atol=0, equal_nan=True)):
                old[:] = new
                counts[row, f] = count
                word = int(result.trace_flags[i, 0])
                cutoff = .99 * models[2][row, -1]
                if (word & TRACE_NEAR_CUTOFF and np.any(roots[modes:] >= cutoff)
                        and not np.any(roots[:modes] >= cutoff)):
                    word &= ~TRACE_NEAR_CUTOFF
                trace[row, f] = word
                if old.dtype == np.float32:
                    trace[row, f] |= TRACE_FP64_REPAIR
                levels[row, f] = result.search_level[i, 0]
                if group is not None:
                    group[row, :, f] = result.group_velocity[i, :modes, 0]
                    group_flags[row, :, f] = result.group_velocity_flags[i, :modes, 0]
                flags[row, f] |= int(ContinuityFlag.REPAIRED)
                stats["continuity_repaired"] += 1
            return True

        # Local window: keep the original lower bound (rank) and cut the upper
        # bound just above the observed roots when two guards were found.
        roots = velocity[model_ids, :, freq_ids][:, :, None]
        observed = np.where(np.isfinite(roots), roots, 0.).max(axis=(1, 2))
        if modes > 1:
            gaps = np.diff(roots, axis=1)
            margin = np.where(np.isfinite(gaps), gaps, 0.).max(axis=(1, 2))
        else:
            margin = .1 * observed
        limits = bounds[model_ids]
        upper = observed + np.maximum(2.5 * margin, .02 * observed)
        have_guards = counts[model_ids, freq_ids] >= modes + 2
        eligible = have_guards & (guards == modes + 2) & (upper > limits[:, 0]) & (upper < limits[:, 1] * .98)
        pending = np.ones(n, dtype=bool)
        for local in (True, False):
            ids = np.flatnonzero(eligible if local else pending)
            if not ids.size:
                continue
            cut = limits[ids].copy()
            if local:
                cut[:, 1] = upper[ids]
            result = solve(model_ids[ids], freq_ids[ids], cut, local)
            for i, k in enumerate(ids):
                pending[k] = not adopt(result, i, model_ids[k], freq_ids[k], local)
        flags[model_ids[pending], freq_ids[pending]] |= int(ContinuityFlag.UNRESOLVED)
        stats["continuity_unresolved"] += int(np.count_nonzero(pending))
    flags[model_ids, freq_ids] |= int(ContinuityFlag.REVIEWED)


def dispersion_batch(thickness, vp, vs, rho, periods, nmodes=5, wave="rayleigh", solver=None, names=None,
                     search_strategy=None, accuracy=None, **options):
    """One call on a batch [B, L] (reference __init__.py:572-589)."""
    if solver is not None:
        if (options or (search_strategy is not None and search_strategy != solver.search_strategy)
                or (accuracy is not None and accuracy != solver.accuracy)):
            raise ValueError("options conflict with the supplied Solver")
        return solver(thickness, vp, vs, rho, periods, nmodes, wave, names)
    with Solver(search_strategy="shape_adaptive" if search_strategy is None else search_strategy,
                accuracy=accuracy, **options) as s:
        return s(thickness, vp, vs, rho, periods, nmodes, wave, names)


def dispersion(model, periods, nmodes=5, wave="rayleigh", **options):
    """One model given as an [L, 4] array of columns h, vp, vs, rho or a dict
    of the four arrays (reference __init__.py:591-597)."""
    batch = model
    if not isinstance(batch, dict):
        arr = np.asarray(batch)
        if arr.ndim != 2 or arr.shape[1] != 4:
            raise ValueError("model must be an array with columns h/vp/vs/rho, or a batch dict")
        batch = dict(zip(("thickness", "vp", "vs", "rho"), arr.T))
    return dispersion_batch(**batch, periods=periods, nmodes=nmodes, wave=wave, **options)
