"""hipGraph replay of the device pipeline and the launch-shape autotuner.

A captured graph is replayed several times, with eager torch work and
synchronisations between the replays (outputs are cleared by a kernel each
time); every replay must equal the eager run bit for bit. The autotuner must
pick one of its candidates, and every launch shape must give the same results.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import modisp_amd
    if not modisp_amd.available(0):
        pytest.fail("no gfx950 device visible to the gpu tests")
    modisp_amd.set_device(0)
    return torch, modisp_amd


def _device_case(torch, wl, window):
    dev = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    B, L = wl.thickness.shape
    T = wl.periods.size
    bufs = dict(thickness=t(wl.thickness), vp=t(wl.vp), vs=t(wl.vs), rho=t(wl.rho), periods=t(wl.periods),
                bounds=t(window), velocity=torch.empty(B * 5 * T, dtype=torch.float64, device=dev),
                flags=torch.empty(B * T, dtype=torch.int32, device=dev),
                levels=torch.empty(B * T, dtype=torch.int32, device=dev),
                counts=torch.empty(B * T, dtype=torch.int32, device=dev),
                group=torch.empty(B * 5 * T, dtype=torch.float64, device=dev),
                group_flags=torch.empty(B * 5 * T, dtype=torch.int32, device=dev))
    return bufs, {k: v.data_ptr() for k, v in bufs.items()}


OUTPUTS = ("velocity", "flags", "levels", "counts", "group", "group_flags")


def _snapshot(bufs):
    return {k: bufs[k].cpu().numpy().copy() for k in OUTPUTS}


def _equal(a, b):
    for k in OUTPUTS:
        np.testing.assert_array_equal(a[k].view(np.uint8), b[k].view(np.uint8), err_msg=k)


@pytest.mark.parametrize("wave", [0, 1], ids=["rayleigh", "love"])
def test_graph_replays_equal_the_eager_run(env, wave):
    torch, md = env
    from modisp_amd import workloads
    wl = workloads.precision_batch(B=96, seed=31)
    window = workloads.scan_window(wl.vs)
    B, L = wl.thickness.shape
    bufs, ptr = _device_case(torch, wl, window)
    with md.QeSolver(B, L, wl.periods.size, 5, group_velocity=True) as s:
        s.run_device_batch(B, ptr, 1e-12, wave)
        s.synchronize()
        eager = _snapshot(bufs)
    with md.QeSolver(B, L, wl.periods.size, 5, group_velocity=True, graph=True) as g:
        for replay in range(4):
            for k in OUTPUTS:
                bufs[k].zero_()            # eager torch work between replays
            torch.cuda.synchronize()
            g.run_device_batch(B, ptr, 1e-12, wave)   # call 0 runs eagerly and captures; 1..3 replay
            g.synchronize()
            _equal(_snapshot(bufs), eager)
            assert g.last_timing()[0] > 0.


def test_autotuner_picks_a_candidate_and_shapes_agree(env):
    torch, md = env
    from modisp_amd import workloads
    wl = workloads.precision_batch(B=256, seed=32)
    window = workloads.scan_window(wl.vs)
    B, L = wl.thickness.shape
    bufs, ptr = _device_case(torch, wl, window)
    with md.QeSolver(B, L, wl.periods.size, 5, group_velocity=True) as s:
        s.run_device_batch(B, ptr, 1e-12, 0)
        s.synchronize()
        ref = _snapshot(bufs)
        w, b = s.autotune_device_batch(B, ptr, 1e-12, 0)
        assert w in (1, 2, 4, 8) and b in (8, 16, 24)
        s.run_device_batch(B, ptr, 1e-12, 0)
        s.synchronize()
        _equal(_snapshot(bufs), ref)
    for w in (1, 2, 8):
        with md.QeSolver(B, L, wl.periods.size, 5, group_velocity=True, scan_waves=w, consumer_blocks=4) as s:
            s.run_device_batch(B, ptr, 1e-12, 0)
            s.synchronize()
            _equal(_snapshot(bufs), ref)


# --- ragged batches resident on the device -------------------------------------------------------
RAGGED_OUT = ("velocity", "flags", "levels", "counts", "group", "group_flags")


def _ragged_case(torch, md, packed):
    dev = torch.device("cuda", 0)
    geom, cells = md.ragged_geometry(packed["row_offsets"], packed["period_offsets"], packed["modes"])
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    bufs = {k: t(packed[k]) for k in ("thickness", "vp", "vs", "rho", "row_offsets", "periods", "period_offsets",
                                      "modes", "bounds")}
    bufs["cell_offsets"] = t(cells)
    nc, nt = geom["ncells"], geom["ntraces"]
    bufs.update(velocity=torch.empty(nc, dtype=torch.float64, device=dev),
                flags=torch.empty(nt, dtype=torch.int32, device=dev),
                levels=torch.empty(nt, dtype=torch.int32, device=dev),
                counts=torch.empty(nt, dtype=torch.int32, device=dev),
                group=torch.empty(nc, dtype=torch.float64, device=dev),
                group_flags=torch.empty(nc, dtype=torch.int32, device=dev))
    return geom, bufs, {k: v.data_ptr() for k, v in bufs.items()}


def _host_ragged(md, packed, geom):
    with md.QeSolver(geom["B"], geom["max_rows"], 1, 1, group_velocity=True) as s:
        out = s.run_ragged(*(packed[k] for k in ("thickness", "vp", "vs", "rho", "row_offsets", "periods",
                                                 "period_offsets", "modes", "bounds")), 1e-12, 0)
    return dict(velocity=out["velocity"], flags=out["flags"].view(np.int32), levels=out["levels"],
                counts=out["counts"], group=out["group_velocity"], group_flags=out["group_velocity_flags"].view(np.int32))


def _ragged_equal(bufs, ref):
    for k in RAGGED_OUT:
        got = bufs[k].cpu().numpy()
        np.testing.assert_array_equal(got.view(np.uint8), np.ascontiguousarray(ref[k]).view(np.uint8), err_msg=k)


def test_ragged_device_graph_replays_equal_the_host_ragged_run(env):
    """md_rocm_run_ragged_device (device preparation of the uneven recipe:
    3..40 finite layers, low-velocity zones, per-model periods and modes)
    equals md_rocm_run_ragged (host preparation) bit for bit, eagerly and on
    every replay of its captured graph."""
    torch, md = env
    from modisp_amd import workloads
    wl = workloads.uneven_batch(B=96, seed=107)
    packed = workloads.pack_ragged(wl, period_step=4)
    geom, bufs, ptr = _ragged_case(torch, md, packed)
    ref = _host_ragged(md, packed, geom)
    with md.QeSolver(geom["B"], geom["max_rows"], 1, 1, group_velocity=True) as s:
        s.run_ragged_device_batch(geom, ptr, 1e-12, 0)
        s.synchronize()
        _ragged_equal(bufs, ref)
        assert s.last_stats()["invalid_models"] == 0
    with md.QeSolver(geom["B"], geom["max_rows"], 1, 1, group_velocity=True, graph=True) as g:
        for replay in range(4):
            for k in RAGGED_OUT:
                bufs[k].zero_()            # eager torch work between replays
            torch.cuda.synchronize()
            g.run_ragged_device_batch(geom, ptr, 1e-12, 0)   # call 0 runs eagerly and captures; 1..3 replay
            g.synchronize()
            _ragged_equal(bufs, ref)
            assert g.last_timing()[0] > 0.


def test_ragged_device_autotuner_and_inconsistent_offsets(env):
    """The autotuner picks a candidate and keeps the results; a model whose
    cell range does not hold modes x periods cells is solved as an empty model
    (count 0, level -4, its cells untouched) and counted, and every other
    model's results are unchanged."""
    torch, md = env
    from modisp_amd import workloads
    wl = workloads.uneven_batch(B=64, seed=108)
    packed = workloads.pack_ragged(wl, period_step=5)
    geom, bufs, ptr = _ragged_case(torch, md, packed)
    ref = _host_ragged(md, packed, geom)
    with md.QeSolver(geom["B"], geom["max_rows"], 1, 1, group_velocity=True) as s:
        w, b = s.autotune_ragged_device_batch(geom, ptr, 1e-12, 0)
        assert w in (1, 2, 4, 8) and b in (8, 16, 24)
        s.run_ragged_device_batch(geom, ptr, 1e-12, 0)
        s.synchronize()
        _ragged_equal(bufs, ref)
        bad = 7
        mode_bad = packed["modes"].copy()
        mode_bad[bad] += 1                       # cell range no longer modes x periods
        bufs["modes"].copy_(torch.from_numpy(mode_bad))
        bufs["velocity"].fill_(12345.)
        torch.cuda.synchronize()
        s.run_ragged_device_batch(geom, ptr, 1e-12, 0)
        s.synchronize()
        assert s.last_stats()["invalid_models"] == 1
    _, cells = md.ragged_geometry(packed["row_offsets"], packed["period_offsets"], packed["modes"])
    po = packed["period_offsets"]
    tr, cl = slice(po[bad], po[bad + 1]), slice(cells[bad], cells[bad + 1])
    assert (bufs["counts"].cpu().numpy()[tr] == 0).all() and (bufs["levels"].cpu().numpy()[tr] == -4).all()
    assert (bufs["velocity"].cpu().numpy()[cl] == 12345.).all()
    keep_t = np.ones(geom["ntraces"], bool)
    keep_t[tr] = False
    keep_c = np.ones(geom["ncells"], bool)
    keep_c[cl] = False
    for k, keep in (("velocity", keep_c), ("counts", keep_t), ("levels", keep_t), ("flags", keep_t)):
        np.testing.assert_array_equal(bufs[k].cpu().numpy()[keep].view(np.uint8),
                                      np.ascontiguousarray(ref[k][keep]).view(np.uint8), err_msg=k)
