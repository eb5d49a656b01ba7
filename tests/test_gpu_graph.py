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
        assert w in (1, 2, 4, 8) and b in (4, 8, 16)
        s.run_device_batch(B, ptr, 1e-12, 0)
        s.synchronize()
        _equal(_snapshot(bufs), ref)
    for w in (1, 2, 8):
        with md.QeSolver(B, L, wl.periods.size, 5, group_velocity=True, scan_waves=w, consumer_blocks=4) as s:
            s.run_device_batch(B, ptr, 1e-12, 0)
            s.synchronize()
            _equal(_snapshot(bufs), ref)
