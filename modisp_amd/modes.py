"""Cross-period checks of recovered mode ranks (heuristic diagnostics).

The interface (``check_mode_reliability``, ``ModeReliabilityFlag``,
``ModeReliability`` and its fields) is the reference's
(python/modisp/reliability.py); it judges a finished solve without changing
it. Periods are visited in ascending order; each period is compared with the
previous distinct one (and, for the trend test, the one before that):

* NONFINITE / NUMERICAL_FAILURE / SEARCH_UNCERTAIN from the values and the
  trace flags; CLOSE_ROOT and RANK_AMBIGUOUS (non-ascending) within a trace;
* UNASSESSED for the first period, PERIOD_GAP where the log step exceeds the
  limit, DUPLICATE_MISMATCH between repeated periods;
* ROOT_COUNT_CHANGE where the count or the set of finite modes changes;
* STEP_JUMP for a relative change above ``max_relative_step``;
* RANK_AMBIGUOUS where nearest-value matching between the two periods does
  not map each mode to itself (both directions);
* PREDICTION_MISMATCH where a mode leaves the linear-in-log-period trend of
  the two previous periods by more than ``prediction_rtol``.

Here every test is a vector operation over [models, modes]; rank matching
uses a [B, K, K] distance array with non-finite entries excluded.
"""
from __future__ import annotations

from dataclasses import dataclass
from enum import IntFlag

import numpy as np

__all__ = ["ModeReliabilityFlag", "ModeReliability", "check_mode_reliability"]


class ModeReliabilityFlag(IntFlag):
    UNASSESSED = 1
    NUMERICAL_FAILURE = 2
    ROOT_COUNT_CHANGE = 4
    CLOSE_ROOT = 8
    RANK_AMBIGUOUS = 16
    STEP_JUMP = 32
    PREDICTION_MISMATCH = 64
    DUPLICATE_MISMATCH = 128
    NONFINITE = 256
    PERIOD_GAP = 512
    SEARCH_UNCERTAIN = 1024


# Trace flags that mark a numerical failure / an approximate search.
_FAILURE_BITS = 1 | 4 | 8 | 32 | 64 | 128 | 512 | 1024
_UNCERTAIN_BITS = 2048 | 4096


@dataclass(frozen=True)
class ModeReliability:
    flags: np.ndarray              # [B, K, T] in the input's period order
    assessed: np.ndarray           # [B, K, T]: a finite, sufficiently sampled incoming edge
    previous_index: np.ndarray     # [T]: input index of the period compared with (-1 for the first)
    sorted_order: np.ndarray       # [T]: stable ascending order of the periods

    @property
    def suspect(self):
        return self.flags != 0

    @property
    def passes_checks(self):
        """Passed the observed checks; no claim of modal identity or completeness."""
        return self.assessed & ~self.suspect

    def summary(self):
        return dict(cells=int(self.flags.size), assessed=int(self.assessed.sum()),
                    passes_checks=int(self.passes_checks.sum()),
                    flags={f.name: int(np.count_nonzero(self.flags & int(f))) for f in ModeReliabilityFlag})


def _log_step(a, b):
    """log(b / a) for a < b, exact for nearly equal periods."""
    return float(np.log1p((b - a) / a)) if b - a < a else float(np.log(b) - np.log(a))


def _positive(name, x):
    if not np.isscalar(x) or np.iscomplexobj(x) or not np.isfinite(x) or x <= 0:
        raise ValueError(f"{name} must be finite and positive")


def _inputs(result, periods, trace_flags, root_count, close_root_rtol):
    if hasattr(result, "phase_velocity"):
        if periods is not None or trace_flags is not None or root_count is not None:
            raise ValueError("periods/trace_flags/root_count are supplied by the result")
        periods, trace_flags, root_count = result.period, result.trace_flags, result.root_count
        close_root_rtol = result.close_root_rtol if close_root_rtol is None else close_root_rtol
        result = result.phase_velocity
    if np.iscomplexobj(result) or np.iscomplexobj(periods):
        raise ValueError("velocities and periods must be real")
    v = np.asarray(result, dtype=float)
    p = np.asarray(periods, dtype=float)
    if (v.ndim != 3 or min(v.shape) < 1 or p.ndim != 1 or p.size != v.shape[2] or not np.isfinite(p).all()
            or np.any(p <= 0)):
        raise ValueError("expected velocity [B,K,T] and finite positive periods [T]")
    if np.isinf(v).any() or np.any(v[np.isfinite(v)] <= 0):
        raise ValueError("velocities must be positive finite values or NaN")
    B, K, T = v.shape
    if trace_flags is None:
        tf = np.zeros((B, T), dtype=np.uint32)
    else:
        tf = np.asarray(trace_flags)
        if tf.shape != (B, T) or tf.dtype.kind not in "iu" or np.any(tf < 0):
            raise ValueError("trace_flags must be nonnegative integer [B,T]")
    if root_count is None:
        counts = np.isfinite(v).sum(axis=1)
    else:
        counts = np.asarray(root_count)
        if counts.shape != (B, T) or counts.dtype.kind not in "iu":
            raise ValueError("root_count must be integer [B,T]")
    return v, p, tf.astype(np.int64), counts, .005 if close_root_rtol is None else close_root_rtol


def _rank_mismatch(a, b, fa, fb):
    """[B, K] modes at the later period whose nearest earlier mode is not the
    same rank, or whose nearest earlier mode's nearest later mode is not it."""
    K = a.shape[1]
    cost = np.abs(a[:, :, None] - b[:, None, :])                 # [B, earlier, later]
    cost[~(fa[:, :, None] & fb[:, None, :])] = np.inf
    nearest = np.argmin(cost, axis=1)                             # [B, later] -> earlier rank
    back = np.argmin(cost, axis=2)                                # [B, earlier] -> later rank
    ranks = np.arange(K)[None, :]
    round_trip = np.take_along_axis(back, nearest, axis=1)
    return fb & fa.any(axis=1)[:, None] & ((nearest != ranks) | (round_trip != ranks))


def check_mode_reliability(result, periods=None, *, trace_flags=None, root_count=None, close_root_rtol=None,
                           max_relative_step=.1, prediction_rtol=.05, max_log_period_gap=.25):
    """Flags [B, K, T] of recovered ranks across sorted periods, for a
    DispersionBatch or a velocity array [B, K, T] with its periods."""
    v, p, tf, counts, close_rtol = _inputs(result, periods, trace_flags, root_count, close_root_rtol)
    for name, x in (("close_root_rtol", close_rtol), ("max_relative_step", max_relative_step),
                    ("prediction_rtol", prediction_rtol), ("max_log_period_gap", max_log_period_gap)):
        _positive(name, x)
    F = ModeReliabilityFlag
    B, K, T = v.shape
    finite = np.isfinite(v)
    flags = np.zeros(v.shape, dtype=np.uint16)
    assessed = np.zeros(v.shape, dtype=bool)
    previous = np.full(T, -1, dtype=np.int64)

    def mark(mask, bit, where=(slice(None),) * 3):
        flags[where] |= np.where(mask, np.uint16(int(bit)), np.uint16(0))

    failed = ((tf & _FAILURE_BITS) != 0) | (counts < 0)                     # [B, T]
    mark(~finite, F.NONFINITE)
    mark(np.broadcast_to(failed[:, None, :], v.shape), F.NUMERICAL_FAILURE)
    mark(np.broadcast_to(((tf & _UNCERTAIN_BITS) != 0)[:, None, :], v.shape), F.SEARCH_UNCERTAIN)
    if K > 1:
        gap = np.diff(v, axis=1)
        near = np.isfinite(gap) & (gap <= close_rtol * v[:, 1:, :])
        disorder = np.isfinite(gap) & (gap <= 0)
        lower, upper = (slice(None), slice(0, -1), slice(None)), (slice(None), slice(1, None), slice(None))
        for mask, bit in ((near, F.CLOSE_ROOT), (disorder, F.RANK_AMBIGUOUS)):
            mark(mask, bit, lower)   # both members of a pair
            mark(mask, bit, upper)
    order = np.argsort(p, kind="stable")
    chain = []   # input indices of the distinct periods visited so far, ascending
    for idx in order:
        if not chain:
            flags[:, :, idx] |= np.uint16(int(F.UNASSESSED))
            chain.append(int(idx))
            continue
        prev = chain[-1]
        previous[idx] = prev
        a, b = v[:, :, prev], v[:, :, idx]
        fa, fb = finite[:, :, prev], finite[:, :, idx]
        both = fa & fb
        clean = both & ~failed[:, prev, None] & ~failed[:, idx, None]
        changed = (counts[:, prev] != counts[:, idx]) | np.any(fa != fb, axis=1)
        flags[:, :, idx] |= np.where(changed[:, None], np.uint16(int(F.ROOT_COUNT_CHANGE)), np.uint16(0))
        if p[idx] == p[prev]:                     # a repeated period: the two results must agree
            differ = (fa != fb) | (both & (np.abs(a - b) > 1e-10 * np.fmax(a, b)))
            bit = np.where(differ, np.uint16(int(F.DUPLICATE_MISMATCH)), np.uint16(0))
            flags[:, :, idx] |= bit
            flags[:, :, prev] |= bit
            assessed[:, :, idx] = clean
            continue
        step_log = _log_step(p[prev], p[idx])
        if step_log > max_log_period_gap:
            flags[:, :, idx] |= np.uint16(int(F.PERIOD_GAP))
        else:
            assessed[:, :, idx] = clean
        rel = np.zeros_like(a)
        np.divide(np.abs(b - a), np.fmax(a, b), out=rel, where=both)
        flags[:, :, idx] |= np.where(both & (rel > max_relative_step), np.uint16(int(F.STEP_JUMP)), np.uint16(0))
        flags[:, :, idx] |= np.where(_rank_mismatch(a, b, fa, fb), np.uint16(int(F.RANK_AMBIGUOUS)), np.uint16(0))
        if len(chain) >= 2:
            older = chain[-2]
            step_old = _log_step(p[older], p[prev])
            if step_old <= max_log_period_gap and step_log <= max_log_period_gap:
                with np.errstate(invalid="ignore"):
                    trend = a + (a - v[:, :, older]) * (step_log / step_old)
                    off = both & finite[:, :, older] & (np.abs(b - trend) > prediction_rtol * b)
                flags[:, :, idx] |= np.where(off, np.uint16(int(F.PREDICTION_MISMATCH)), np.uint16(0))
        chain.append(int(idx))
    return ModeReliability(flags, assessed, previous, order)
