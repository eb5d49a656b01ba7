"""Comparison of this backend's results with the reference's golden vectors.

The vectors (``tests/golden/reference/*.npz``, written by
``tests/golden/make_reference_vectors.py`` from the reference CPU backend)
hold, per seeded workload subset, the reference's outputs under its default
``shape_adaptive`` strategy (prefix ``shape_``) and under ``qe_adaptive``
(prefix ``qe_``). A trace matches when:

* root counts are equal (traces the reference leaves unresolved, count -1,
  are not compared);
* the validity masks of the reported modes are equal and phase velocities
  agree to 1e-9 relative;
* the result-defined trace flags agree: search unconverged (4), near cutoff
  (16), f rejected (128) and Love phase fallback (16384), and close-root risk
  (4096) where the reference accepted the trace with its primary search (its
  high adaptive fallback ladder never sets 4096 for Rayleigh traces);
* the search level agrees where the reference accepted the trace with its
  primary search (8 shape, 3 adaptive, 11 Love phase, -3/-4 empty); its
  fallback ladders (levels 5 and 3 under ``shape``) have no counterpart;
* group velocities: the same flags, and values within the reference
  finite-difference polish's own bound, 0.25 group_rtol max(c, |U|)
  (cpp/common/group_velocity.hpp:87-92), of this backend's analytic U. Where
  the reference's finite-difference polish fails (POLISH_FAILED 2,
  DERIVATIVE_UNSTABLE 4 or ILL_CONDITIONED 8) and this backend's analytic
  derivative returns a valid U with no flag, the cell is counted
  (`analytic_only`) rather than a deviation: there is no reference value, and
  the analytic U is held to the oracle's analytic derivative by the other GPU
  tests.

Every trace that does not match is a deviation and must be named, with the
fields that differ, in ``tests/golden/reference/deviations.json``.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

VECTORS = Path(__file__).resolve().parent / "golden" / "reference"
DEVIATIONS = VECTORS / "deviations.json"
MODEL = ("thickness", "vp", "vs", "rho")
RTOL = 1e-9
GROUP_RTOL = 1e-3
FLAG_BITS = 4 | 16 | 128 | 4096 | 16384
FD_FAILURES = (2, 4, 8)
PRIMARY_LEVELS = {"shape": (8, 11, -3, -4), "qe": (3, 11, -3, -4)}


def rectangular_files():
    return sorted(p for p in VECTORS.glob("*.npz") if p.stem != "uneven_rayleigh")


def compare_trace(got, ref, i, t, nmodes, strategy):
    """Fields of trace (model i, period t) that differ; `got`/`ref` hold arrays
    [B, modes, T] / [B, T] under the names of the reference's result tuple."""
    fields = []
    rc = int(ref["root_count"][i, t])
    if rc >= 0 and int(got["root_count"][i, t]) != rc:
        fields.append("count")
    gv, rv = got["phase_velocity"][i, :nmodes, t], ref["phase_velocity"][i, :nmodes, t]
    if rc >= 0:
        if not np.array_equal(np.isfinite(gv), np.isfinite(rv)):
            fields.append("mask")
        ok = np.isfinite(gv) & np.isfinite(rv)
        if ok.any() and np.max(np.abs(gv[ok] - rv[ok]) / rv[ok]) > RTOL:
            fields.append("velocity")
    rl = int(ref["search_level"][i, t])
    bits = FLAG_BITS if rl in PRIMARY_LEVELS[strategy] else FLAG_BITS & ~4096
    if (int(got["trace_flags"][i, t]) ^ int(ref["trace_flags"][i, t])) & bits:
        fields.append("flags")
    if rl in PRIMARY_LEVELS[strategy] and int(got["search_level"][i, t]) != rl:
        fields.append("level")
    if ref.get("group_velocity") is not None and got.get("group_velocity") is not None:
        gf, rf = got["group_velocity_flags"][i, :nmodes, t], ref["group_velocity_flags"][i, :nmodes, t]
        analytic_only = (gf == 0) & np.isin(rf, FD_FAILURES)
        if np.any((gf != rf) & ~analytic_only):
            fields.append("group_flags")
        gg, rg = got["group_velocity"][i, :nmodes, t], ref["group_velocity"][i, :nmodes, t]
        ok = np.isfinite(gg) & np.isfinite(rg) & np.isfinite(rv)
        if ok.any():
            bound = .25 * GROUP_RTOL * np.maximum(rv[ok], np.abs(rg[ok]))
            if np.any(np.abs(gg[ok] - rg[ok]) > bound + RTOL * np.abs(rg[ok])):
                fields.append("group")
    return fields


def reference_arrays(z, strategy):
    keys = ("phase_velocity", "trace_flags", "search_level", "root_count", "group_velocity",
            "group_velocity_flags")
    return {k: z[f"{strategy}_{k}"] for k in keys}


def differences(got, z, strategy, nmodes):
    ref = reference_arrays(z, strategy)
    B, T = ref["root_count"].shape
    out = {}
    for i in range(B):
        for t in range(T):
            f = compare_trace(got, ref, i, t, nmodes, strategy)
            if f:
                out[f"{i},{t}"] = f
    return out


def load_deviations():
    if not DEVIATIONS.exists():
        return {}
    return json.loads(DEVIATIONS.read_text())["deviations"]


def analytic_only_cells(got, z, strategy):
    """Cells whose group velocity only this backend's analytic derivative gives."""
    rf = z[f"{strategy}_group_velocity_flags"]
    gf = got["group_velocity_flags"][:, :rf.shape[1], :rf.shape[2]]
    return int(np.count_nonzero((gf == 0) & np.isin(rf, FD_FAILURES)))
