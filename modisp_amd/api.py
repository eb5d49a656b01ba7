"""Front end of the ROCm backend: ``Solver``, ``dispersion_batch``,
``dispersion`` and the ``DispersionBatch`` result.

The public surface is the reference's (python/modisp/__init__.py): argument
names and defaults, validation messages, result fields, status codes and
flag values. Those are held here as tables; the machinery around them is
this package's own:

* a declarative check list validates models and options;
* a period plan decides once which periods the device solves, in which
  order, and how results return to the caller's order;
* one device engine (``QeSolver``) is kept per geometry and reused for
  smaller calls; batches larger than the memory budget run in chunks;
* result masks come from one native pass (``md_rocm_result_masks``);
* the continuity review of approximate presets is array-oriented: all
  suspects of a pass go to one FP64 verifier call with one period per row,
  and acceptance, change detection and write-back are vector operations.

Only the ``rocm`` backend lives here; ``backend="auto"`` resolves to it when a
gfx950 device is present and raises ``BackendUnavailableError`` otherwise.
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass
from enum import IntEnum, IntFlag

import numpy as np

from . import QeSolver, BackendUnavailableError, available, bounds as _auto_bounds, set_device, STAT_NAMES
from . import continuity_screen, result_masks

# --- interface tables (values fixed by the reference's API) -------------------------------
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


class ContinuityFlag(IntFlag):
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


@dataclass(frozen=True)
class _Preset:
    """One accuracy preset: its engine code, default refinement tolerance and
    shape sampling tolerance, and whether its searches are approximate."""
    code: int
    c_tol: float
    sampling: float
    approximate: bool


_PRESETS = {"auto": _Preset(4, 1e-5, 1e-2, True), "high": _Preset(0, 1e-6, 1e-3, False),
            "balanced": _Preset(1, 1e-5, 1e-2, True), "fast": _Preset(2, 1e-4, 5e-2, True),
            "uniform": _Preset(0, 1e-6, 1e-3, False)}
_REFINEMENT_CODE = {"legacy": 0, "brent": 1, "chandrupatla": 2}
_STRATEGIES = ("shape_adaptive", "qe_adaptive", "uniform")
_CONTINUATION_COUNTERS = (("root_count_unsafe", PredictionFlag.ROOT_COUNT_UNSAFE),
                          ("anchor_unsafe", PredictionFlag.ANCHOR_UNSAFE),
                          ("window_unsafe", PredictionFlag.WINDOW_UNSAFE),
                          ("bracket_failed", PredictionFlag.BRACKET_FAILED),
                          ("correction_failed", PredictionFlag.CORRECTION_FAILED),
                          ("audit_failed", PredictionFlag.AUDIT_FAILED))
# Device bytes per (model, period) of the engine's trace work buffers, and per
# model row; used to turn workspace_bytes into models per chunk.
_TRACE_BYTES, _ROW_BYTES, _FIXED_BYTES = 3600, 64, 1 << 20


# --- validation --------------------------------------------------------------------------
def _real_scalar(v):
    return (np.isscalar(v) and not isinstance(v, (str, bytes, bool, np.bool_)) and not np.iscomplexobj(v)
            and np.isfinite(v))


def _integer(v):
    return isinstance(v, (int, np.integer)) and not isinstance(v, (bool, np.bool_))


def _flag(v):
    return isinstance(v, (bool, np.bool_))


def _check(rules):
    """Raise ValueError with the message of the first rule whose test fails."""
    for ok, message in rules:
        if not ok():
            raise ValueError(message)


def _models(thickness, vp, vs, rho):
    """[B, L] float64 model arrays, validated by the reference's rules and messages."""
    named = (("thickness", thickness), ("vp", vp), ("vs", vs), ("rho", rho))
    arrays = []
    for name, value in named:
        _check([(lambda: not np.iscomplexobj(value), f"{name} must be real")])
        a = np.ascontiguousarray(value, dtype=np.float64)
        a = a[None, :] if a.ndim == 1 else a
        _check([(lambda: a.ndim == 2 and a.size > 0 and bool(np.isfinite(a).all()),
                 f"{name} must be a finite nonempty [B,L] or [L] array")])
        arrays.append(a)
    h, a, b, r = arrays
    solid = b > 0
    _check([
        (lambda: all(x.shape == h.shape for x in arrays), "model array shapes must agree"),
        (lambda: h.shape[1] <= 256, "at most 256 layers are supported"),
        (lambda: not (np.any(h < 0) or np.any(a <= 0) or np.any(r <= 0) or np.any(b < 0)),
         "thickness/vs must be non-negative and vp/rho positive"),
        (lambda: not (np.any(b[:, -1] <= 0) or np.any(b[:, 1:] == 0)),
         "only a single first water layer may have vs=0; the half-space must be solid"),
        (lambda: not np.any(a[solid] / b[solid] <= np.sqrt(4. / 3.)),
         "solid vp/vs must exceed sqrt(4/3) for positive bulk modulus"),
    ])
    return arrays


def _periods(periods):
    _check([(lambda: not np.iscomplexobj(periods), "periods must be real")])
    per = np.ascontiguousarray(periods, dtype=np.float64)
    _check([(lambda: per.ndim == 1 and per.size > 0 and bool(np.isfinite(per).all()) and not np.any(per <= 0),
             "periods must be a finite positive 1-D array")])
    return per


def _fp32_range(arrays, per):
    f = np.finfo(np.float32)
    big, tiny = float(f.max), .5 * float(f.smallest_subnormal)
    for name, x in zip(("thickness", "vp", "vs", "rho"), arrays):
        _check([(lambda: not (x.max() > big or np.any((x > 0.) & (x <= tiny))),
                 f"{name} contains values outside the float32 range")])
    _check([(lambda: not (per.max() > big or per.min() < 2. * np.pi / big),
             "periods or their angular frequencies are outside the float32 range")])


# --- period plan -------------------------------------------------------------------------
@dataclass
class _PeriodPlan:
    """Which periods the device solves (`grid`), whether by anchor
    continuation, and the index that restores the caller's order."""
    grid: np.ndarray
    restore: np.ndarray | None
    continuation: bool
    reason: str

    @classmethod
    def build(cls, per, solver, backend):
        rising = per.size < 2 or bool(np.all(np.diff(per) > 0))
        falling = per.size >= 2 and bool(np.all(np.diff(per) < 0))
        wants = solver.frequency_strategy == "anchor_continuation"
        # Continuation needs its periods sorted; any other repeated or
        # unsorted (but not strictly falling) grid is solved once per unique
        # period. A strictly falling grid keeps the caller's order.
        needs_sorted = wants and solver.accuracy != "high" and per.size >= 8
        unique, inverse = per, None
        if not rising and (needs_sorted or not falling):
            unique, inverse = np.unique(per, return_inverse=True)
        continuation, reason = False, "independent frequency search requested"
        if wants:
            near = unique.size >= 3 and bool(np.any(np.log(unique[2:]) - np.log(unique[:-2]) <= .25))
            for blocked, why in ((solver.accuracy == "high", "high accuracy retains independent search"),
                                 (unique.size < 8, "fewer than 8 unique periods; using independent search"),
                                 (not near, "period grid has no interior target within the anchor log-gap limit; "
                                            "using independent search")):
                if blocked:
                    reason = why
                    break
            else:
                continuation, reason = True, f"{backend.upper()} anchor continuation on sorted unique periods"
        if continuation or unique.size < per.size:
            return cls(unique, inverse, continuation, reason)
        return cls(per, None, continuation, reason)

    def back(self, x):
        return x if self.restore is None or x is None else x.take(self.restore, axis=-1)


# --- results -------------------------------------------------------------------------------
class DispersionBatch:
    """Phase (and group) velocities [B, nmodes, T] with the reference's masks,
    status codes and per-trace diagnostics (field names as in
    python/modisp/__init__.py:71-139)."""

    def __init__(self, velocity, periods, nmodes, wave, flags, levels, counts, bounds, backend, reason, **meta):
        self.phase_velocity = velocity
        self.precision = velocity.dtype.name
        self.period = np.array(periods, copy=True)
        self.mode = np.arange(nmodes)
        self.wave = wave
        self.trace_flags = flags
        self.search_level = levels
        self.root_count = counts
        self.search_bounds = np.array(bounds, copy=True)
        self.backend_used = backend
        self.backend_reason = reason
        defaults = dict(accuracy=None, close_root_rtol=.005, search_strategy=None, sampling_rtol=None, secular=None,
                        refinement=None, love_strategy=None, group_velocity=None, group_velocity_flags=None,
                        prediction_flags=None, group_rtol=None, frequency_strategy="independent",
                        frequency_strategy_used="independent",
                        frequency_strategy_reason="independent frequency search requested", anchor_stride=4,
                        continuity_flags=None, continuity_unresolved=False)
        unknown = set(meta) - set(defaults)
        if unknown:
            raise TypeError(f"unknown result fields {sorted(unknown)}")
        defaults.update(meta)
        unresolved = defaults.pop("continuity_unresolved")
        for k, v in defaults.items():
            setattr(self, k, v)
        self.close_root_rtol = float(self.close_root_rtol)
        self.anchor_stride = int(self.anchor_stride)
        (self.status, self.mode_mask, self.close_root_mask, self.close_root_risk,
         self.group_velocity_mask) = result_masks(velocity, flags, counts, self.close_root_rtol, self.group_velocity,
                                                  self.group_velocity_flags)
        self.group_velocity_valid = self.group_velocity_mask
        self.numerical_flags = flags & np.uint32(0xFFFFFFFF ^ (TRACE_RANGE_TRUNCATED | TRACE_AUTO_RELAXED))
        self.approximate_search = (flags & (TRACE_APPROXIMATE_SEARCH | TRACE_CONTINUATION_APPROXIMATE)) != 0
        uncertain = self.approximate_search.copy()
        if unresolved and self.continuity_flags is not None:
            uncertain |= (self.continuity_flags & ContinuityFlag.UNRESOLVED) != 0
        self.mode_index_uncertain = uncertain
        self.potentially_ambiguous = uncertain | self.close_root_risk
        self.mode_convention = ("ascending recovered roots; close-cluster omissions may shift indices"
                                if self.approximate_search.any() else "ascending roots in the reported search interval")

    def check_modes(self, **options):
        """Cross-period mode checks of this result (modisp_amd.modes)."""
        from .modes import check_mode_reliability
        return check_mode_reliability(self, **options)

    @property
    def auto_relaxed(self):
        return (self.trace_flags & TRACE_AUTO_RELAXED) != 0 if self.accuracy == "auto" else None

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


class _Outputs:
    """Result arrays of one solve, filled chunk by chunk (or taken whole from a
    single chunk's engine output)."""

    def __init__(self):
        self.velocity = self.flags = self.levels = self.counts = None
        self.group = self.group_flags = self.prediction = None

    def put(self, sl, B, out, with_group, with_prediction):
        velocity, flags, _, levels, counts, extras = out
        parts = dict(velocity=velocity, flags=flags, levels=levels, counts=counts,
                     group=extras["group_velocity"] if with_group else None,
                     group_flags=extras["group_velocity_flags"] if with_group else None,
                     prediction=extras["prediction_flags"] if with_prediction else None)
        whole = sl.start == 0 and sl.stop == B
        for name, part in parts.items():
            if part is None:
                continue
            if whole:
                setattr(self, name, part)
                continue
            if getattr(self, name) is None:
                setattr(self, name, np.empty((B,) + part.shape[1:], dtype=part.dtype))
            getattr(self, name)[sl] = part


# --- solver --------------------------------------------------------------------------------
class Solver:
    """Reusable phase dispersion solver on the ROCm backend; one instance per
    concurrent caller. Options are the reference's (python/modisp/__init__.py:
    160-271); see `_engine_options` for what each one selects here."""

    def __init__(self, nc=512, cmin=None, cmax=None, c_tol_rel=None, f_accept=1., block=256,
                 search_strategy="shape_adaptive", workspace_bytes=256 * 1024 ** 2, chunk_size=None, *,
                 backend="auto", device=0, threads=None, profile=False, accuracy=None,
                 adaptive_period_threshold=None, close_root_rtol=.005, sampling_rtol=None, secular=None,
                 refinement=None, love_strategy="phase", precision="float64", group_velocity=False,
                 group_rtol=.001, frequency_strategy="independent", anchor_stride=4, continuity_check=True):
        uniform = search_strategy == "uniform"
        _check([
            (lambda: backend in ("auto", "cpu", "cuda", "rocm"), "backend must be auto, cpu, cuda or rocm"),
            (lambda: search_strategy in _STRATEGIES, "invalid search_strategy"),
            (lambda: _flag(continuity_check), "continuity_check must be boolean"),
            (lambda: _flag(group_velocity), "group_velocity must be boolean"),
            (lambda: _real_scalar(group_rtol) and 0 < group_rtol < 1,
             "group_rtol must be finite and between 0 and 1"),
            (lambda: frequency_strategy in ("independent", "anchor_continuation"),
             "frequency_strategy must be independent or anchor_continuation"),
            (lambda: _integer(anchor_stride) and anchor_stride >= 2, "anchor_stride must be an integer of at least 2"),
            (lambda: isinstance(precision, str) and precision in ("float64", "float32"),
             "precision must be float64 or float32"),
            (lambda: not (precision == "float32" and (backend == "cpu" or uniform)),
             "precision='float32' requires an adaptive CUDA solver"),
            (lambda: secular in (None, "dunkin", "fast"), "secular must be None, 'dunkin' or 'fast'"),
            (lambda: refinement in (None, "legacy", "brent", "chandrupatla"),
             "refinement must be legacy, brent or chandrupatla"),
            (lambda: love_strategy in ("phase", "scan"), "love_strategy must be phase or scan"),
        ])
        if uniform and (group_velocity or frequency_strategy != "independent"):
            raise NotImplementedError("group velocity and frequency continuation require an adaptive solver")
        refinement = refinement or ("legacy" if uniform else "brent")
        if uniform and refinement != "legacy":
            raise NotImplementedError("uniform supports only legacy refinement")
        if uniform and secular == "fast":
            raise NotImplementedError("secular='fast' is not available on the legacy uniform CUDA path")
        if accuracy is None or (uniform and accuracy == "auto"):
            accuracy = "uniform" if uniform else DEFAULT_ACCURACY
        elif uniform:
            raise ValueError("accuracy applies only to adaptive strategies; uniform is the legacy fixed path")
        _check([(lambda: accuracy in ("auto", "high", "balanced", "fast", "uniform"),
                 "accuracy must be auto, high, balanced or fast")])
        preset = _PRESETS[accuracy]
        if adaptive_period_threshold is not None:
            _check([(lambda: accuracy not in ("high", "uniform"),
                     "adaptive_period_threshold applies only to auto/balanced/fast"),
                    (lambda: _real_scalar(adaptive_period_threshold) and adaptive_period_threshold > 0,
                     "adaptive_period_threshold must be finite and positive")])
        _check([(lambda: _real_scalar(close_root_rtol) and 0 < close_root_rtol < 1,
                 "close_root_rtol must be finite and between 0 and 1"),
                (lambda: sampling_rtol is None or search_strategy == "shape_adaptive",
                 "sampling_rtol applies only to shape_adaptive")])
        self._auto_sampling = accuracy == "auto" and search_strategy == "shape_adaptive" and sampling_rtol is None
        sampling_rtol = preset.sampling if sampling_rtol is None else sampling_rtol
        c_tol_rel = preset.c_tol if c_tol_rel is None else c_tol_rel
        _check([
            (lambda: _real_scalar(sampling_rtol) and 0 < sampling_rtol <= 1,
             "sampling_rtol must be finite and between 0 and 1"),
            (lambda: cmin is None or (np.isfinite(cmin) and cmin > 0), "cmin must be finite and positive"),
            (lambda: cmax is None or (np.isfinite(cmax) and cmax > 0), "cmax must be finite and positive"),
            (lambda: cmin is None or cmax is None or cmax > cmin, "cmax must exceed cmin"),
            (lambda: np.isfinite(c_tol_rel) and c_tol_rel > 0, "c_tol_rel must be positive"),
            (lambda: np.isfinite(f_accept) and f_accept >= 0, "f_accept must be non-negative"),
            (lambda: nc >= 3, "nc must be at least 3"),
            (lambda: _integer(workspace_bytes) and workspace_bytes >= 1, "workspace_bytes must be a positive integer"),
            (lambda: chunk_size is None or (_integer(chunk_size) and chunk_size >= 1), "chunk_size must be positive"),
            (lambda: _integer(device) and device >= 0, "device must be a non-negative integer"),
            (lambda: threads is None or (_integer(threads) and 1 <= threads <= 256), "threads must be 1..256"),
            (lambda: _flag(profile), "profile must be boolean"),
        ])
        self.nc, self.cmin, self.cmax, self.block = int(nc), cmin, cmax, int(block)
        self.c_tol_rel, self.f_accept = float(c_tol_rel), float(f_accept)
        self.search_strategy, self.workspace_bytes, self.chunk_size = search_strategy, int(workspace_bytes), chunk_size
        self.backend, self.device, self.threads, self.profile = backend, int(device), threads, bool(profile)
        self.accuracy, self.adaptive_period_threshold = accuracy, adaptive_period_threshold
        self.close_root_rtol, self.sampling_rtol = float(close_root_rtol), float(sampling_rtol)
        self.secular, self.refinement, self.love_strategy = secular, refinement, love_strategy
        self.precision, self.group_velocity, self.group_rtol = precision, bool(group_velocity), float(group_rtol)
        self.frequency_strategy, self.anchor_stride = frequency_strategy, int(anchor_stride)
        self.continuity_check = bool(continuity_check)
        self._engine = None
        self._engine_key = None
        self._calls = 0
        self._last_stats = None
        self._key = None
        self.last_result = None
        self._lock = threading.RLock()

    # -- engine ----------------------------------------------------------------------------
    def _engine_options(self, continuation):
        """QeSolver arguments: the preset's code and strategy (search level
        codes and approximate flags), the refinement and Love methods, the
        requested precision for the scan and the returned velocities, and, for
        search_strategy="uniform", the legacy nc-node uniform grid."""
        uniform = self.search_strategy == "uniform"
        threshold = self.adaptive_period_threshold or (1.0 if self.accuracy in ("balanced", "auto") else
                                                       .63 if self.accuracy == "fast" else 0.)
        return dict(accuracy=_PRESETS[self.accuracy].code, period_threshold=float(threshold),
                    close_root_rtol=self.close_root_rtol, shape=self.search_strategy == "shape_adaptive",
                    sampling_rtol=-1. if self._auto_sampling else self.sampling_rtol,
                    secular=0 if self.secular == "dunkin" else 1, refinement=_REFINEMENT_CODE[self.refinement],
                    love_phase=int(self.love_strategy == "phase"), single_precision=self.precision == "float32",
                    group_velocity=self.group_velocity, group_rtol=self.group_rtol, continuation=continuation,
                    anchor_stride=self.anchor_stride, grid="uniform" if uniform else "adaptive", nc=self.nc,
                    scan="float32" if uniform else None)

    def _capacity(self, B, L, T):
        per_model = T * _TRACE_BYTES + L * _ROW_BYTES
        fit = (self.workspace_bytes - _FIXED_BYTES) // per_model
        if fit < 1:
            raise MemoryError("workspace_bytes is too small for one model and the backend's fixed workspace; "
                              "increase the budget or reduce threads")
        return int(min(B, fit, self.chunk_size or B))

    def _acquire(self, capacity, L, T, nmodes, continuation):
        key = (L, T, int(nmodes), continuation, tuple(sorted(self._engine_options(continuation).items())))
        if self._engine is not None and self._engine_key[0] == key and self._engine_key[1] >= capacity:
            return self._engine_key[1], False
        self.close()
        self._engine = QeSolver(capacity, L, T, int(nmodes), **self._engine_options(continuation))
        self._engine_key = (key, capacity)
        return capacity, True

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

    # -- one solve ------------------------------------------------------------------------
    def _solve(self, thickness, vp, vs, rho, periods, nmodes, wave, _bounds=None, _row_periods=None):
        """One solve. `_bounds` ([B, 2]) and `_row_periods` (one period per
        model, `periods` then holding the first) serve the continuity review."""
        clock = time.perf_counter()
        _check([(lambda: wave in ("rayleigh", "love"), "wave must be rayleigh or love"),
                (lambda: _integer(nmodes) and 1 <= nmodes <= 256, "nmodes must be 1..256")])
        models = _models(thickness, vp, vs, rho)
        per = _periods(periods)
        if self.precision == "float32":
            _fp32_range(models, per)
        B, L = models[0].shape
        uniform = self.search_strategy == "uniform"
        if uniform and (wave != "rayleigh" or np.any(models[2] == 0) or L < 2):
            raise NotImplementedError("uniform is the legacy CUDA solid-Rayleigh option; "
                                      "use qe_adaptive for this combination")
        backend, reason = self._select()
        plan = _PeriodPlan.build(per, self, backend)
        if _row_periods is not None:
            plan = _PeriodPlan(per[:1], None, False, "independent frequency search requested")
        wave_code = int(wave == "love")
        low = .2 if uniform and self.cmin is None else self.cmin
        bnd = (_auto_bounds(*models, wave_code, low, self.cmax) if _bounds is None
               else np.ascontiguousarray(_bounds, dtype=np.float64))
        _check([(lambda: bool(np.isfinite(bnd).all()) and not np.any(bnd[:, 0] <= 0),
                 "automatic search range could not be constructed")])
        T = plan.grid.size
        capacity, allocated = self._acquire(self._capacity(B, L, T), L, T, nmodes, plan.continuation)
        rows = None if _row_periods is None else np.ascontiguousarray(_row_periods, dtype=np.float64).reshape(B, 1)
        out, stats, chunks = _Outputs(), {k: 0 for k in STAT_NAMES}, 0
        for start in range(0, B, capacity):
            sl = slice(start, min(B, start + capacity))
            res = self._engine.run(*(m[sl] for m in models), plan.grid if rows is None else rows[sl], bnd[sl],
                                   self.c_tol_rel, self.f_accept, wave_code, self.cmin is None, self.profile)
            out.put(sl, B, res, self.group_velocity, plan.continuation)
            for k, v in res[2].items():
                stats[k] = stats.get(k, 0) + int(v)
            chunks += 1
        prediction = out.prediction
        if self.frequency_strategy == "anchor_continuation" and prediction is None:
            prediction = np.full((B, T), int(PredictionFlag.BYPASSED), dtype=np.uint32)
        if plan.continuation:
            for name, bit in _CONTINUATION_COUNTERS:
                stats["continuation_fallback_" + name] = int(np.count_nonzero(prediction & int(bit)))
            stats["continuation_accepted"] = int(np.count_nonzero(prediction & int(PredictionFlag.ACCEPTED)))
            stats["continuation_fallbacks"] = int(np.count_nonzero(prediction & int(PredictionFlag.FALLBACK)))
        continuity = self._continuity(models, plan, wave, bnd, out, prediction, stats)
        velocity, flags, levels, counts = (plan.back(x) for x in (out.velocity, out.flags, out.levels, out.counts))
        group, gflags, prediction, continuity = (plan.back(x) for x in (out.group, out.group_flags, prediction,
                                                                        continuity))
        if uniform:
            flags |= np.uint32(TRACE_RANGE_TRUNCATED)
            levels[:] = -2
        if self.cmin is not None or self.cmax is not None:
            flags |= np.uint32(TRACE_RANGE_TRUNCATED)
        counts[(flags & TRACE_SEARCH_UNCONVERGED) != 0] = -1
        self._record_stats(stats, backend, wave, flags, plan, per.size, chunks, capacity, allocated, clock)
        result = DispersionBatch(
            velocity, per, nmodes, wave, flags, levels, counts, bnd, backend, reason, accuracy=self.accuracy,
            close_root_rtol=self.close_root_rtol, search_strategy=self.search_strategy,
            sampling_rtol=self.sampling_rtol if self.search_strategy == "shape_adaptive" else None,
            secular="dunkin" if uniform else (self.secular or "fast"), refinement=self.refinement,
            love_strategy=self.love_strategy if wave == "love" else None, group_velocity=group,
            group_velocity_flags=gflags, prediction_flags=prediction,
            group_rtol=self.group_rtol if self.group_velocity else None, frequency_strategy=self.frequency_strategy,
            frequency_strategy_used="anchor_continuation" if plan.continuation else "independent",
            frequency_strategy_reason=plan.reason, anchor_stride=self.anchor_stride, continuity_flags=continuity,
            continuity_unresolved=stats.get("continuity_unresolved", 0) > 0)
        self._key = (B, L, T, self.nc, int(nmodes))
        self._calls += 1
        self.last_result = result
        return result

    def _continuity(self, models, plan, wave, bounds, out, prediction, stats):
        """Screen adjacent periods of approximate presets and review suspects in FP64."""
        if not _PRESETS[self.accuracy].approximate:
            return None
        if not (self.continuity_check or plan.continuation or self.accuracy == "auto"):
            return None
        flags, suspect = continuity_screen(plan.grid, out.velocity, prediction, self.continuity_check)
        if self.accuracy == "auto":
            failed = (out.flags & (TRACE_UNRESOLVED_BRACKET | TRACE_SEARCH_UNCONVERGED | TRACE_ROOT_CONFLICT)) != 0
            if failed.any():
                flags[failed] |= int(ContinuityFlag.SEARCH)
                suspect = True
        for k in ("suspected", "local_searches", "full_searches", "repaired", "unresolved", "evaluations"):
            stats["continuity_" + k] = 0
        if suspect:
            _review(self, models, plan.grid, wave, bounds, out.velocity, out.flags, out.levels, out.counts, out.group,
                    out.group_flags, flags, stats)
        return flags

    def _record_stats(self, stats, backend, wave, flags, plan, requested, chunks, capacity, allocated, clock):
        uniform = self.search_strategy == "uniform"
        stats["evaluations"] = sum(int(stats.get(k, 0)) for k in ("scan_nodes", "fp64_rechecks", "refine_evaluations",
                                                                  "probe_evaluations", "probe_evaluations_fp64"))
        attempts = stats.get("continuation_accepted", 0) + stats.get("continuation_fallbacks", 0)
        stats.update(
            backend=backend, precision=self.precision, accuracy=self.accuracy, requested_strategy=self.search_strategy,
            strategy="uniform" if uniform else "rocm_exhaustive_scan", search_grid="uniform" if uniform else "adaptive",
            sampling_rtol=self.sampling_rtol, secular="dunkin" if uniform else (self.secular or "fast"),
            secular_kernel="sh" if wave == "love" else "dunkin_fast", refinement=self.refinement,
            love_strategy=self.love_strategy if wave == "love" else None, close_root_rtol=self.close_root_rtol,
            refinement_tolerance=self.c_tol_rel, chunks=chunks, capacity_models=int(capacity),
            engine_allocations=int(allocated), group_velocity=self.group_velocity, group_rtol=self.group_rtol,
            frequency_strategy=self.frequency_strategy,
            frequency_strategy_used="anchor_continuation" if plan.continuation else "independent",
            frequency_strategy_reason=plan.reason, anchor_stride=self.anchor_stride,
            periods_requested=int(requested), periods_computed=int(plan.grid.size),
            approximate_search_enabled=_PRESETS[self.accuracy].approximate,
            approximate_traces=int(np.count_nonzero(flags & TRACE_APPROXIMATE_SEARCH)),
            close_root_risk_traces=int(np.count_nonzero(flags & TRACE_CLOSE_ROOT_RISK)),
            continuation_local_success_rate=stats.get("continuation_accepted", 0) / attempts if attempts else None,
            continuity_check=self.continuity_check and _PRESETS[self.accuracy].approximate,
            workspace_budget_bytes=self.workspace_bytes)
        if self.accuracy == "auto":
            stats["auto_relaxed_traces"] = int(np.count_nonzero(flags & TRACE_AUTO_RELAXED))
            stats["sampling_rtol_max"] = (None if self.search_strategy != "shape_adaptive" else
                                          .05 if self._auto_sampling else self.sampling_rtol)
        if self.profile:
            stats["timings_s"] = {"solver_total": time.perf_counter() - clock}
        self._last_stats = stats

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
        love = last is not None and last.wave == "love"
        report = dict(backend="rocm", precision=precision, accuracy=self.accuracy, secular=self.secular or "fast",
                      secular_kernel="sh" if love else "dunkin_fast", close_root_rtol=self.close_root_rtol,
                      refinement_tolerance=self.c_tol_rel, refinement=self.refinement,
                      love_strategy=self.love_strategy, group_velocity=self.group_velocity, group_rtol=self.group_rtol,
                      frequency_strategy=self.frequency_strategy, anchor_stride=self.anchor_stride)
        if last is not None:
            report.update(frequency_strategy_used=last.frequency_strategy_used,
                          frequency_strategy_reason=last.frequency_strategy_reason)
        if self._key is not None:
            cfg = self._engine.config if self._engine is not None else None
            report.update(qe_kernel_report(self._key[1], int(self.accuracy != "high"),
                                           self.search_strategy == "shape_adaptive", precision == "float32",
                                           config=cfg))
            if love and self.love_strategy == "phase":
                report["scan_kernel_report"] = dict(report)
                report.update(love_phase_kernel_report(self._key[1], precision == "float32"))
                report["scan_kernel"] = report.get("search_kernel")
                report["search_kernel"] = "love_phase_with_scan_fallback"
        return report

    def stats(self):
        if self._last_stats is None:
            raise RuntimeError("call the solver first")
        return dict(self._last_stats)


# --- continuity review -----------------------------------------------------------------------
_REASONS = int(ContinuityFlag.ENVELOPE | ContinuityFlag.CLOSE_FP32 | ContinuityFlag.PREDICTION | ContinuityFlag.SEARCH)
_VERIFIER_FAILURE = TRACE_SEARCH_UNCONVERGED | TRACE_UNRESOLVED_BRACKET | TRACE_F_REJECTED


def _review(solver, models, periods, wave, bounds, velocity, trace, levels, counts, group, group_flags, flags, stats):
    """FP64 review of the traces the continuity screen flagged (the reference
    semantics are python/modisp/_continuity.py:17-205): each suspect is
    re-solved alone at accuracy="high" with two guard roots beyond the
    requested modes, first in a window cut just above its observed roots when
    the trace showed both guards, then in the full window. A verified result
    that differs from the reported one replaces it (REPAIRED), one that cannot
    be verified leaves the trace UNRESOLVED, and every suspect is REVIEWED.
    The arrays are updated in place."""
    rows, cols = np.nonzero(flags & _REASONS)
    n = rows.size
    stats["continuity_suspected"] = n
    if not n:
        return
    modes = velocity.shape[1]
    guards = min(256, modes + 2)
    limits = bounds[rows].copy()
    # Local windows: keep the lower bound (it fixes the modes' ranks) and cut
    # the upper bound above the highest observed root by 2.5 of the largest
    # observed gap (or 2% of the root), if both guards were found.
    seen = velocity[rows, :, cols].astype(np.float64)
    top = np.nanmax(np.where(np.isfinite(seen), seen, 0.), axis=1)
    if modes > 1:
        gaps = np.diff(seen, axis=1)
        margin = np.nanmax(np.where(np.isfinite(gaps), gaps, 0.), axis=1)
    else:
        margin = .1 * top
    cut = top + np.maximum(2.5 * margin, .02 * top)
    local = ((counts[rows, cols] >= modes + 2) & (guards == modes + 2) & (cut > limits[:, 0])
             & (cut < .98 * limits[:, 1]))
    settled = np.zeros(n, dtype=bool)
    verifier_options = dict(backend="rocm", device=solver.device, accuracy="high", cmin=solver.cmin,
                            cmax=solver.cmax, c_tol_rel=min(solver.c_tol_rel, 1e-6), f_accept=solver.f_accept,
                            secular=solver.secular, refinement=solver.refinement,
                            group_velocity=solver.group_velocity, group_rtol=solver.group_rtol,
                            close_root_rtol=solver.close_root_rtol, continuity_check=False,
                            workspace_bytes=solver.workspace_bytes, chunk_size=solver.chunk_size)
    with Solver(**verifier_options) as verifier:
        for stage, members in (("local", local), ("full", None)):
            idx = np.flatnonzero(members if members is not None else ~settled)
            if not idx.size:
                continue
            window = limits[idx].copy()
            if stage == "local":
                window[:, 1] = cut[idx]
            stats[f"continuity_{stage}_searches"] += idx.size
            try:
                res = verifier._solve(*(m[rows[idx]] for m in models), periods[cols[idx]][:1], guards, wave,
                                      _bounds=window, _row_periods=periods[cols[idx]])
            except MemoryError:
                continue
            stats["continuity_evaluations"] += verifier.stats().get("evaluations", 0)
            ok = _verified(res, stage == "local", modes, guards, velocity.dtype, verifier.c_tol_rel)
            _adopt(res, idx[ok], np.flatnonzero(ok), rows, cols, models, velocity, trace, levels, counts, group,
                   group_flags, flags, solver.c_tol_rel, stats)
            settled[idx[ok]] = True
    flags[rows[~settled], cols[~settled]] |= int(ContinuityFlag.UNRESOLVED)
    stats["continuity_unresolved"] += int(np.count_nonzero(~settled))
    flags[rows, cols] |= int(ContinuityFlag.REVIEWED)


def _verified(res, local, modes, guards, dtype, tol):
    """Which verifier rows are conclusive: a clean FP64 search whose known
    roots (guards included) are finite and strictly ascending in the result
    precision; a local search must also have found both guards below its cut."""
    count = res.root_count[:, 0].astype(np.int64)
    ok = (count >= 0) & (res.search_level[:, 0] >= 0) & ((res.trace_flags[:, 0] & _VERIFIER_FAILURE) == 0)
    if local:
        ok &= count >= modes + 2
    known = np.minimum(count, guards)
    roots = res.phase_velocity[:, :, 0].astype(dtype)
    inside = np.arange(roots.shape[1])[None, :] < known[:, None]
    finite = np.where(inside, np.isfinite(roots), True).all(axis=1)
    pairs = inside[:, 1:]
    rising = np.where(pairs, roots[:, 1:] > roots[:, :-1], True).all(axis=1)
    ok &= finite & rising
    if local:
        last = np.take_along_axis(res.phase_velocity[:, :, 0], np.maximum(known - 1, 0)[:, None], 1)[:, 0]
        ok &= (known > 0) & (last < res.search_bounds[:, 1] * (1 - 4 * tol))
    return ok


def _adopt(res, members, res_rows, rows, cols, models, velocity, trace, levels, counts, group, group_flags, flags, tol,
           stats):
    """Write the verified results that differ from the reported ones."""
    if not members.size:
        return
    modes = velocity.shape[1]
    r, c = rows[members], cols[members]
    new = res.phase_velocity[res_rows, :modes, 0].astype(velocity.dtype)
    old = velocity[r, :, c]
    count = np.minimum(res.root_count[res_rows, 0], modes + 2)
    same = np.isclose(old, new, rtol=tol, atol=0, equal_nan=True).all(axis=1)
    identical = ((old == new) | (np.isnan(old) & np.isnan(new))).all(axis=1)
    reason = flags[r, c]
    change = (((reason & int(ContinuityFlag.SEARCH)) != 0) | (counts[r, c] != count) | ~same
              | (((reason & int(ContinuityFlag.CLOSE_FP32)) != 0) & ~identical))
    if not change.any():
        return
    k = np.flatnonzero(change)
    r, c, rr = r[k], c[k], res_rows[k]
    velocity[r, :, c] = new[k]
    counts[r, c] = count[k]
    word = res.trace_flags[rr, 0].copy()
    # A cutoff flag raised only by a guard root does not describe the reported modes.
    cutoff = .99 * models[2][r, -1]
    every = res.phase_velocity[rr, :, 0]
    guard_hit = np.any(np.nan_to_num(every[:, modes:], nan=-np.inf) >= cutoff[:, None], axis=1)
    mode_hit = np.any(np.nan_to_num(every[:, :modes], nan=-np.inf) >= cutoff[:, None], axis=1)
    word = np.where(((word & TRACE_NEAR_CUTOFF) != 0) & guard_hit & ~mode_hit, word & ~np.uint32(TRACE_NEAR_CUTOFF),
                    word)
    if velocity.dtype == np.float32:
        word |= np.uint32(TRACE_FP64_REPAIR)
    trace[r, c] = word
    levels[r, c] = res.search_level[rr, 0]
    if group is not None:
        group[r, :, c] = res.group_velocity[rr, :modes, 0]
        group_flags[r, :, c] = res.group_velocity_flags[rr, :modes, 0]
    flags[r, c] |= int(ContinuityFlag.REPAIRED)
    stats["continuity_repaired"] += int(k.size)


# --- convenience functions -------------------------------------------------------------------
def dispersion_batch(thickness, vp, vs, rho, periods, nmodes=5, wave="rayleigh", solver=None, names=None,
                     search_strategy=None, accuracy=None, **options):
    """One call on a batch [B, L]; with `solver`, no other option may differ from it."""
    if solver is not None:
        conflicting = (options or (search_strategy is not None and search_strategy != solver.search_strategy)
                       or (accuracy is not None and accuracy != solver.accuracy
                           and not (accuracy == "auto" and solver.search_strategy == "uniform")))
        if conflicting:
            raise ValueError("options conflict with the supplied Solver")
        return solver(thickness, vp, vs, rho, periods, nmodes, wave, names)
    with Solver(search_strategy=search_strategy or "shape_adaptive", accuracy=accuracy, **options) as s:
        return s(thickness, vp, vs, rho, periods, nmodes, wave, names)


def dispersion(model, periods, nmodes=5, wave="rayleigh", **options):
    """One model: an [L, 4] array of columns h, vp, vs, rho, or a dict of the four arrays."""
    batch = model
    if not isinstance(batch, dict):
        arr = np.asarray(batch)
        if arr.ndim != 2 or arr.shape[1] != 4:
            raise ValueError("model must be LayeredModel, an array with columns h/vp/vs/rho, or a batch dict")
        batch = {k: arr[:, j] for j, k in enumerate(("thickness", "vp", "vs", "rho"))}
    return dispersion_batch(**batch, periods=periods, nmodes=nmodes, wave=wave, **options)
