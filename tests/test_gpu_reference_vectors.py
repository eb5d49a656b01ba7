"""GPU results against the reference's own outputs (golden vectors).

Each seeded subset of the five workloads was solved by the reference CPU
backend under its default strategy (``shape_adaptive``) and under
``qe_adaptive`` (``tests/golden/make_reference_vectors.py``). The device
solves the same inputs through the C ABI with both scan precisions and the
matching ``shape`` setting; every trace must match the reference (see
``tests/reference_vectors.py`` for the fields) except the deviations named in
``tests/golden/reference/deviations.json``, where the device must return the
recorded roots (verified there as FP64 sign changes of the oracle's
evaluator by ``tests/test_reference_vectors.py``).

Set MODISP_DEVIATION_DUMP=<path> to write every difference seen to a JSON
file instead of failing on the first one.
"""
import json
import os

import numpy as np
import pytest

import reference_vectors as rv

pytestmark = pytest.mark.gpu

DUMP = os.environ.get("MODISP_DEVIATION_DUMP")
_seen = {}


@pytest.fixture(scope="module")
def md():
    import modisp_amd
    if not modisp_amd.available(0):
        pytest.fail("no gfx950 device visible to the gpu tests")
    modisp_amd.set_device(0)
    return modisp_amd


@pytest.fixture(scope="module", autouse=True)
def dump_at_end():
    yield
    if DUMP:
        with open(DUMP, "w") as fh:
            json.dump(_seen, fh, indent=1, sort_keys=True)


def check(label, diff, roots, groups):
    """Compare the differing traces with the named deviations of `label`: the
    same traces and fields, the recorded roots, and for a reference
    finite-difference group error the recorded curve-difference velocities."""
    expected = rv.load_deviations().get(label, {})
    if DUMP:
        _seen[label] = {k: dict(fields=v, roots=[float(x) for x in roots[k]], group=[float(x) for x in groups[k]])
                        for k, v in diff.items()}
        return
    assert sorted(diff) == sorted(expected), f"{label}: differing traces {sorted(diff)} vs named {sorted(expected)}"
    for key, fields in diff.items():
        dev = expected[key]
        assert sorted(fields) == sorted(dev["fields"]), f"{label} trace {key}: {fields}"
        want = np.array(dev["roots"], dtype=float)
        got = np.array(roots[key], dtype=float)
        assert np.array_equal(np.isfinite(got), np.isfinite(want)), f"{label} trace {key}: {got} vs {want}"
        ok = np.isfinite(want)
        np.testing.assert_allclose(got[ok], want[ok], rtol=rv.RTOL, err_msg=f"{label} trace {key}")
        for mode, u in zip(dev.get("group_modes", []), dev.get("curve_group", [])):
            assert abs(groups[key][mode] - u) <= rv.RTOL * abs(u), f"{label} trace {key} mode {mode} group"


@pytest.mark.parametrize("scan", ["float32", "float64"])
@pytest.mark.parametrize("strategy", ["shape", "qe"])
@pytest.mark.parametrize("path", rv.rectangular_files(), ids=lambda p: p.stem)
def test_rectangular_subset(md, path, strategy, scan):
    z = np.load(path)
    model = [np.ascontiguousarray(z[k]) for k in rv.MODEL]
    B, L = model[0].shape
    per, nmodes, wave = z["periods"], int(z["nmodes"]), int(z["wave"])
    print(f"{path.stem} {strategy} {scan}: B={B} L={L} T={per.size} modes={nmodes} wave={wave}", flush=True)
    with md.QeSolver(B, L, per.size, nmodes, accuracy=0, shape=strategy == "shape", group_velocity=True,
                     scan=scan) as s:
        vel, flags, st, lev, cnt, ex = s.run(*model, per, z["bounds"], 1e-12, 1., wave, False)
    got = dict(phase_velocity=vel, trace_flags=flags, search_level=lev, root_count=cnt,
               group_velocity=ex["group_velocity"], group_velocity_flags=ex["group_velocity_flags"])
    diff = rv.differences(got, z, strategy, nmodes)
    roots = {k: vel[int(k.split(",")[0]), :, int(k.split(",")[1])] for k in diff}
    groups = {k: ex["group_velocity"][int(k.split(",")[0]), :, int(k.split(",")[1])] for k in diff}
    check(f"{path.stem}|{strategy}", diff, roots, groups)


@pytest.mark.parametrize("scan", ["float32", "float64"])
@pytest.mark.parametrize("strategy", ["shape", "qe"])
def test_uneven_subset(md, strategy, scan):
    z = np.load(rv.VECTORS / "uneven_rayleigh.npz")
    layers, nper, modes = z["layers"], z["nperiods"], z["nmodes"]
    B = layers.size
    rows, row_off, per, per_off = [], [0], [], [0]
    for i in range(B):
        rows.append(np.stack([z[k][i, :layers[i]] for k in rv.MODEL]))
        row_off.append(row_off[-1] + int(layers[i]))
        per.append(z["periods"][i, :nper[i]])
        per_off.append(per_off[-1] + int(nper[i]))
    packed = np.concatenate(rows, axis=1)
    print(f"uneven {strategy} {scan}: B={B} traces={per_off[-1]}", flush=True)
    with md.QeSolver(B, int(layers.max()), 1, 1, accuracy=0, shape=strategy == "shape", group_velocity=True,
                     scan=scan) as s:
        out = s.run_ragged(*packed, row_off, np.concatenate(per), per_off, modes, z["bounds"], 1e-12, 0)
    Kmax, Tmax = int(modes.max()), int(nper.max())
    got = dict(phase_velocity=np.full((B, Kmax, Tmax), np.nan), group_velocity=np.full((B, Kmax, Tmax), np.nan),
               group_velocity_flags=np.zeros((B, Kmax, Tmax), np.uint32), trace_flags=np.zeros((B, Tmax), np.uint32),
               search_level=np.zeros((B, Tmax), np.int32), root_count=np.zeros((B, Tmax), np.int32))
    cell = 0
    for i in range(B):
        M, T = int(modes[i]), int(nper[i])
        sl = slice(per_off[i], per_off[i + 1])
        got["phase_velocity"][i, :M, :T] = out["velocity"][cell:cell + M * T].reshape(M, T)
        got["group_velocity"][i, :M, :T] = out["group_velocity"][cell:cell + M * T].reshape(M, T)
        got["group_velocity_flags"][i, :M, :T] = out["group_velocity_flags"][cell:cell + M * T].reshape(M, T)
        got["trace_flags"][i, :T] = out["flags"][sl]
        got["search_level"][i, :T] = out["levels"][sl]
        got["root_count"][i, :T] = out["counts"][sl]
        cell += M * T
    diff = {}
    ref = rv.reference_arrays(z, strategy)
    for i in range(B):
        for t in range(int(nper[i])):
            f = rv.compare_trace(got, ref, i, t, int(modes[i]), strategy)
            if f:
                diff[f"{i},{t}"] = f
    roots = {k: got["phase_velocity"][int(k.split(",")[0]), :, int(k.split(",")[1])] for k in diff}
    groups = {k: got["group_velocity"][int(k.split(",")[0]), :, int(k.split(",")[1])] for k in diff}
    check(f"uneven_rayleigh|{strategy}", diff, roots, groups)
