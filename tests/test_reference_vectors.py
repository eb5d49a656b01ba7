"""The reference's golden vectors, the oracle against them, and the named deviations (CPU).

* The vector files are the ones the manifest lists (sha256).
* The oracle's restatement of the reference's ``qe_adaptive`` search, FD group
  velocity and Love phase path reproduces every ``qe_*`` vector bit for bit:
  phase velocities, trace flags, search levels, root counts, group velocities
  and their flags, on subsets of all five workloads.
* Every named deviation (``deviations.json``) is re-verified with the oracle's
  FP64 evaluator: roots the reference missed (one of a collapsing pair, or a
  root at the top of the window) are genuine sign changes of F and the
  device's roots are the first sign changes; a reference
  finite-difference group velocity that misses the dispersion curve's own
  derivative by more than its bound; a reference Love phase fallback that
  returned the same roots.
"""
import hashlib
import json

import numpy as np
import pytest

import oracle
import reference_vectors as rv

MANIFEST = json.loads((rv.VECTORS / "manifest.json").read_text())
FILES = [c["file"] for c in MANIFEST["cases"]]


@pytest.mark.parametrize("case", MANIFEST["cases"], ids=lambda c: c["file"])
def test_vector_files_match_the_manifest(case):
    data = (rv.VECTORS / case["file"]).read_bytes()
    assert hashlib.sha256(data).hexdigest() == case["sha256"]


def assert_bitwise(got, z, prefix, sl=(slice(None),) * 3, label=""):
    for key in ("phase_velocity", "group_velocity"):
        a, b = got[key], z[f"{prefix}_{key}"][sl]
        assert np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)]), \
            f"{label} {key}"
    for key in ("trace_flags", "search_level", "root_count"):
        assert np.array_equal(got[key], z[f"{prefix}_{key}"][sl[0], sl[2]]), f"{label} {key}"
    assert np.array_equal(got["group_velocity_flags"], z[f"{prefix}_group_velocity_flags"][sl]), f"{label} group flags"


@pytest.mark.parametrize("path", rv.rectangular_files(), ids=lambda p: p.stem)
def test_oracle_reproduces_the_reference_qe_vectors(path):
    z = np.load(path)
    got = oracle.solve(*(z[k] for k in rv.MODEL), z["periods"], z["bounds"], int(z["nmodes"]), int(z["wave"]),
                       tol=1e-12, refinement=1, love_phase=1, group=1, group_rtol=rv.GROUP_RTOL)
    assert_bitwise(got, z, "qe", label=path.stem)


def test_oracle_reproduces_the_reference_uneven_qe_vectors():
    z = np.load(rv.VECTORS / "uneven_rayleigh.npz")
    for i in range(z["layers"].size):
        L, T, M = int(z["layers"][i]), int(z["nperiods"][i]), int(z["nmodes"][i])
        got = oracle.solve(*(z[k][i, :L] for k in rv.MODEL), z["periods"][i, :T], z["bounds"][i:i + 1], M, 0,
                           tol=1e-12, refinement=1, love_phase=1, group=1, group_rtol=rv.GROUP_RTOL)
        assert_bitwise(got, z, "qe", (slice(i, i + 1), slice(0, M), slice(0, T)), label=f"uneven model {i}")


def test_the_reference_strategies_disagree_on_collapsing_pairs():
    """The two reference strategies themselves return different mode sets on
    collapsing-pair traces, which is why the device is compared with both."""
    for stem, key in (("sweep_b12_s104_rayleigh", (6, 2)), ("sweep_b32_s106_rayleigh", (14, 5)),
                      ("sweep_rayleigh", (6, 6))):
        z = np.load(rv.VECTORS / f"{stem}.npz")
        a = z["shape_phase_velocity"][key[0], :, key[1]]
        b = z["qe_phase_velocity"][key[0], :, key[1]]
        assert not np.allclose(a, b, rtol=1e-6, equal_nan=True), stem


def _inputs(stem, i, t):
    z = np.load(rv.VECTORS / f"{stem}.npz")
    if "layers" in z.files:
        L = int(z["layers"][i])
        return z, [z[k][i, :L] for k in rv.MODEL], float(z["periods"][i, t]), z["bounds"][i], int(z["nmodes"][i]), 0
    return z, [z[k][i] for k in rv.MODEL], float(z["periods"][t]), z["bounds"][i], int(z["nmodes"]), int(z["wave"])


DEVIATIONS = [(label, key, dev) for label, d in sorted(rv.load_deviations().items()) for key, dev in sorted(d.items())]


@pytest.mark.parametrize("label,key,dev", DEVIATIONS, ids=[f"{lb}:{k}" for lb, k, _ in DEVIATIONS])
def test_named_deviation_is_verified(label, key, dev):
    stem, strategy = label.split("|")
    i, t = map(int, key.split(","))
    z, model, period, window, nmodes, wave = _inputs(stem, i, t)
    ref = z[f"{strategy}_phase_velocity"][i, :nmodes, t]
    got = np.array(dev["roots"][:nmodes], dtype=float)
    if dev["category"] == "reference_missed_root":
        truth = np.array(dev["truth"])
        assert truth.size <= nmodes + 2 and np.all(np.diff(truth) > 0)
        assert np.all((truth > window[0]) & (truth < window[1]))
        for r in truth:   # every listed root is an FP64 sign change of F
            fl, fr = oracle.secular(*model, period, np.array([r * (1 - 1e-10), r * (1 + 1e-10)]), wave=wave)
            assert np.signbit(fl) != np.signbit(fr), r
        m = min(nmodes, truth.size)
        np.testing.assert_allclose(got[:m], truth[:m], rtol=1e-9)
        assert np.all(np.isnan(got[m:]))
        ok = ref[np.isfinite(ref)]
        assert all(np.any(np.abs(truth - r) <= 1e-9 * r) for r in ok)   # the reference's roots are in truth
        missed = np.array(dev["missed"])
        assert missed.size >= 1 and all(np.any(np.abs(truth - r) <= 1e-12 * r) for r in missed)
        assert not any(np.any(np.abs(ok - r) <= 1e-9 * r) for r in missed)
        for r in missed:   # one of a pair closer than close_root_rtol, or at the window top
            gaps = np.abs(truth - r)[np.abs(truth - r) > 0]
            assert (gaps.size and gaps.min() <= .005 * r) or window[1] - r <= 1e-9 * r, r
    elif dev["category"] == "reference_fd_group_error":
        rg = z[f"{strategy}_group_velocity"][i, :nmodes, t]
        for j, u, ru in zip(dev["group_modes"], dev["curve_group"], dev["reference_group"]):
            assert rg[j] == ru
            curve = oracle.curve_group_velocity(*model, period, np.asarray(window)[None], j, wave)
            assert abs(curve - u) <= 1e-10 * abs(u)
            assert abs(ru - u) > .25 * rv.GROUP_RTOL * max(ref[j], abs(ru))
            # the oracle's analytic derivative, which the device follows, is the curve's
            assert abs(oracle.group_analytic(*model, period, ref[j], wave) - u) <= 1e-9 * abs(u)
    elif dev["category"] == "reference_love_phase_fallback":
        assert int(z[f"{strategy}_trace_flags"][i, t]) & 16384 and wave == 1
        ok = np.isfinite(ref)
        np.testing.assert_allclose(got[ok], ref[ok], rtol=1e-9)
        assert np.array_equal(np.isfinite(got), ok)
    else:
        pytest.fail(f"unknown deviation category {dev['category']}")
