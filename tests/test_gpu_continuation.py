"""Anchor continuation across periods (the uneven-batch method).

Every trace the continuation accepts must carry exactly the roots of the
independent search (counts equal, velocities within 1e-9 relative, group
velocities equal to the analytic value at the root); anchors and fallbacks are
solved by the same pipeline as the independent run. The oracle's independent
search checks a sample of models. Prediction flags follow the reference's
meanings (cpp/common/continuation.hpp:6-10): every period is exactly one of
anchor, accepted or fallback, and accepted periods report the approximate
search bits and level 12.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
RTOL = 1e-9


@pytest.fixture(scope="module")
def md():
    import modisp_amd
    if not modisp_amd.available(0):
        pytest.fail("no gfx950 device visible to the gpu tests")
    modisp_amd.set_device(0)
    return modisp_amd


def _check_against_independent(ind, cont, pred, label):
    anchor, accepted, fallback = (pred & 1) != 0, (pred & 2) != 0, (pred & 4) != 0
    assert np.all(anchor.astype(int) + accepted.astype(int) + fallback.astype(int) == 1), label
    assert accepted.sum() >= 0.3 * (~anchor).sum(), f"{label}: only {accepted.sum()} of {(~anchor).sum()} accepted"
    np.testing.assert_array_equal(cont["counts"], ind["counts"], err_msg=f"{label} counts")
    v, w = cont["velocity"], ind["velocity"]
    np.testing.assert_array_equal(np.isfinite(v), np.isfinite(w), err_msg=f"{label} mask")
    ok = np.isfinite(w)
    assert (np.abs(v[ok] - w[ok]) / w[ok]).max() <= RTOL, label
    return accepted


@pytest.mark.parametrize("wave", [0, 1], ids=["rayleigh", "love"])
def test_rectangular_continuation_equals_independent_search(md, wave):
    from modisp_amd import workloads
    wl = workloads.precision_batch(B=64, seed=41)
    per = np.geomspace(0.5, 60., 48)
    window = workloads.scan_window(wl.vs)
    B, L = wl.thickness.shape
    runs = {}
    for cont in (False, True):
        with md.QeSolver(B, L, per.size, 5, group_velocity=True, continuation=cont) as s:
            vel, flags, st, levels, counts, ex = s.run(wl.thickness, wl.vp, wl.vs, wl.rho, per, window, 1e-12, 1.,
                                                       wave, False)
        runs[cont] = dict(velocity=vel, flags=flags, levels=levels, counts=counts, group=ex["group_velocity"],
                          pred=ex["prediction_flags"])
    ind, con = runs[False], runs[True]
    accepted = _check_against_independent(ind, con, con["pred"], f"wave {wave}")
    assert np.all(con["levels"][accepted] == 12)
    assert np.all((con["flags"][accepted] & (2048 | 65536)) == (2048 | 65536))
    g, h = con["group"], ind["group"]
    both = np.isfinite(g) & np.isfinite(h)
    assert (np.abs(g[both] - h[both]) / np.abs(h[both])).max() <= RTOL
    # oracle's independent search on a sample of models
    for i in range(0, B, 16):
        ref = oracle.solve(*(getattr(wl, k)[i:i + 1] for k in ("thickness", "vp", "vs", "rho")), per,
                           window[i:i + 1], 5, wave, tol=1e-12)
        resolved = ref["root_count"][0] >= 0
        np.testing.assert_array_equal(con["counts"][i][resolved], ref["root_count"][0][resolved])
        rv = ref["phase_velocity"][0]
        ok = np.isfinite(rv) & resolved[None, :]
        assert (np.abs(con["velocity"][i][ok] - rv[ok]) / rv[ok]).max() <= RTOL


def test_ragged_continuation_on_the_uneven_batch(md):
    from modisp_amd import workloads
    wl = workloads.uneven_batch(B=48, seed=106)
    rows, row_off, per, per_off, modes, bounds = [], [0], [], [0], [], []
    for i in range(wl.B):
        n = wl.layers[i]
        rows.append(np.stack([a[i, :n] for a in (wl.thickness, wl.vp, wl.vs, wl.rho)]))
        row_off.append(row_off[-1] + n)
        per.append(wl.periods[i])
        per_off.append(per_off[-1] + wl.periods[i].size)
        modes.append(int(wl.nmodes[i]))
        bounds.append([0.8 * wl.vs[i, :n].min(), np.nextafter(wl.vs[i, n - 1], 0.)])
    packed = np.concatenate(rows, axis=1)
    args = (*packed, row_off, np.concatenate(per), per_off, modes, np.array(bounds), 1e-12, 0)
    out = {}
    for cont in (False, True):
        with md.QeSolver(wl.B, int(max(wl.layers)), 1, 1, continuation=cont) as s:
            out[cont] = s.run_ragged(*args)
    pred = out[True]["prediction_flags"]
    assert np.all(out[False]["prediction_flags"] == 8)
    _check_against_independent(out[False], out[True], pred, "uneven")
