"""Name and verify every trace on which the device differs from the reference vectors.

Input: the JSON written by ``tests/test_gpu_reference_vectors.py`` with
MODISP_DEVIATION_DUMP set (each differing trace: the fields that differ, the
device's reported roots and group velocities). Output:
``tests/golden/reference/deviations.json``, which that GPU test then holds the
device to exactly. Every deviation must fall in one of three categories,
each verified here with the oracle's FP64 evaluator (pinned to the reference
by ``tests/test_oracle_pins.py`` and ``tests/test_reference_vectors.py``);
anything else stops the script:

* ``reference_missed_root``: the reference's search misses roots that the
  device returns: one of a collapsing pair narrower than its sampling, or a
  root at the top of the window (a mode at its cutoff, c -> Vs half-space).
  ``truth`` holds the first K sign changes of F in the window (a dense
  1,000,001-node scan joined with the device's roots, each refined to 1e-14);
  the device's reported roots are truth's first modes and the reference's
  roots are a subset of truth.
* ``reference_fd_group_error``: the reference's finite-difference group
  velocity differs from the group velocity of the dispersion curve itself
  (``oracle.curve_group_velocity``: Richardson-extrapolated central
  differences of phase roots in frequency) by more than its own bound
  0.25 group_rtol max(c, |U|); the device's analytic U equals the curve's.
* ``reference_love_phase_fallback``: the reference's Love phase path failed
  its F sign check on a phase bracket narrowed to 16 ulp and fell back to the
  shape search, which returned the same roots; the device's phase count
  accepted them (level 11, no 16384 flag).

    python tests/golden/make_deviations.py <dump.json>
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import oracle  # noqa: E402
import reference_vectors as rv  # noqa: E402

DENSE = 1_000_001


def trace_inputs(stem, i, t):
    z = np.load(rv.VECTORS / f"{stem}.npz")
    if "layers" in z.files:
        L = int(z["layers"][i])
        model = [np.ascontiguousarray(z[k][i, :L]) for k in rv.MODEL]
        period, nmodes = float(z["periods"][i, t]), int(z["nmodes"][i])
    else:
        model = [np.ascontiguousarray(z[k][i]) for k in rv.MODEL]
        period, nmodes = float(z["periods"][t]), int(z["nmodes"])
    return z, model, period, z["bounds"][i], nmodes, int(z["wave"])


def sign_change(model, period, r, wave, rel=1e-10):
    fl, fr = oracle.secular(*model, period, np.array([r * (1 - rel), r * (1 + rel)]), wave=wave)
    return bool(np.signbit(fl) != np.signbit(fr))


def first_roots(model, period, window, K, wave, seeds):
    """The first K sign changes of F in the window, refined to 1e-14."""
    lo, hi = float(window[0]), float(window[1])
    c = np.concatenate([np.linspace(lo, hi, DENSE), [s for s in seeds if lo < s < hi]])
    c = np.unique(c)
    f = oracle.secular(*model, period, c, wave=wave)
    s = np.signbit(f)
    idx = np.nonzero(s[1:] != s[:-1])[0][:K]
    out = []
    for j in idx:
        r, _ = oracle.refine(*model, period, c[j], c[j + 1], 1e-14, wave=wave)
        out.append(r)
    return np.array(out)


def categorize(label, key, entry):
    stem, strategy = label.split("|")
    i, t = map(int, key.split(","))
    z, model, period, window, nmodes, wave = trace_inputs(stem, i, t)
    K = nmodes + 2
    ref_roots = z[f"{strategy}_phase_velocity"][i, :nmodes, t]
    got = np.array(entry["roots"][:nmodes], dtype=float)
    fields = set(entry["fields"])
    out = dict(fields=sorted(fields), roots=entry["roots"])
    if fields & {"count", "mask", "velocity"}:
        seeds = [r for r in got if np.isfinite(r)]
        truth = first_roots(model, period, window, K, wave, seeds)
        assert all(sign_change(model, period, r, wave) for r in truth), (label, key, "truth not sign changes")
        m = min(nmodes, truth.size)
        assert np.allclose(got[:m], truth[:m], rtol=1e-9, atol=0), (label, key, got, truth)
        ref_ok = ref_roots[np.isfinite(ref_roots)]
        assert all(np.any(np.abs(truth - r) <= 1e-9 * r) for r in ref_ok), (label, key, "reference root not in truth")
        missed = [float(r) for r in truth if not np.any(np.abs(ref_ok - r) <= 1e-9 * r)]
        out.update(category="reference_missed_root", truth=[float(x) for x in truth], missed=missed,
                   reference_count=int(z[f"{strategy}_root_count"][i, t]))
        return out
    if "group" in fields:
        rg = z[f"{strategy}_group_velocity"][i, :nmodes, t]
        gg = np.array(entry["group"][:nmodes], dtype=float)
        bound = .25 * rv.GROUP_RTOL * np.maximum(ref_roots, np.abs(rg))
        modes = [int(j) for j in np.nonzero(np.isfinite(rg) & np.isfinite(gg) & (np.abs(gg - rg) > bound))[0]]
        curve = []
        for j in modes:
            u = oracle.curve_group_velocity(*model, period, np.asarray(window)[None], j, wave)
            assert abs(gg[j] - u) <= 1e-9 * abs(u), (label, key, j, gg[j], u)
            assert abs(rg[j] - u) > bound[j], (label, key, j, rg[j], u)
            curve.append(float(u))
        out.update(category="reference_fd_group_error", group_modes=modes, curve_group=curve,
                   reference_group=[float(rg[j]) for j in modes])
        assert fields <= {"group", "group_flags"}, (label, key, fields)
        return out
    if fields <= {"flags", "level"} and wave == 1 and int(z[f"{strategy}_trace_flags"][i, t]) & 16384:
        ok = np.isfinite(ref_roots)
        assert np.array_equal(ok, np.isfinite(got)) and np.allclose(got[ok], ref_roots[ok], rtol=1e-9, atol=0)
        out.update(category="reference_love_phase_fallback")
        return out
    raise SystemExit(f"unexplained deviation {label} {key}: {sorted(fields)}")


def main(dump):
    raw = json.loads(Path(dump).read_text())
    named = {}
    for label in sorted(raw):
        for key in sorted(raw[label], key=lambda k: tuple(map(int, k.split(",")))):
            named.setdefault(label, {})[key] = categorize(label, key, raw[label][key])
            print(label, key, named[label][key]["category"], flush=True)
    doc = dict(
        note="Traces on which the device differs from the reference vectors, by category; written and verified by "
             "tests/golden/make_deviations.py, re-verified by tests/test_reference_vectors.py, held exactly by "
             "tests/test_gpu_reference_vectors.py.",
        deviations=named)
    rv.DEVIATIONS.write_text(json.dumps(doc, indent=1, sort_keys=True) + "\n")


if __name__ == "__main__":
    main(sys.argv[1])
