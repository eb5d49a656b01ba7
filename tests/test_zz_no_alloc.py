"""md_rocm_run_device_static, the entry point that allocates nothing, against md_rocm_run_device.

Two calls in a row through the C ABI (the second with a smaller batch, so the
fixed per-GPU workspace holds the first call's contents when the second
starts) must equal md_rocm_run_device bit for bit, also when the static path
splits the batch into several trace chunks. A batch whose prepared inputs do
not fit the workspace is refused before any launch.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OUTPUTS = ("velocity", "flags", "levels", "counts", "group", "group_flags")


@pytest.fixture(scope="module")
def env():
    import torch
    import modisp_amd
    if not modisp_amd.available(0):
        pytest.fail("no gfx950 device visible to the gpu tests")
    modisp_amd.set_device(0)
    return torch, modisp_amd


def _device_case(torch, h, vp, vs, rho, periods, window, modes):
    dev = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    B, T = h.shape[0], periods.size
    bufs = dict(thickness=t(h), vp=t(vp), vs=t(vs), rho=t(rho), periods=t(periods), bounds=t(window),
                velocity=torch.empty(B * modes * T, dtype=torch.float64, device=dev),
                flags=torch.empty(B * T, dtype=torch.int32, device=dev),
                levels=torch.empty(B * T, dtype=torch.int32, device=dev),
                counts=torch.empty(B * T, dtype=torch.int32, device=dev),
                group=torch.empty(B * modes * T, dtype=torch.float64, device=dev),
                group_flags=torch.empty(B * modes * T, dtype=torch.int32, device=dev))
    return bufs, {k: v.data_ptr() for k, v in bufs.items()}


def _scramble(torch, bufs):
    for k in OUTPUTS:
        bufs[k].fill_(-7)
    torch.cuda.synchronize()


def _snapshot(bufs):
    return {k: bufs[k].cpu().numpy().copy() for k in OUTPUTS}


def _equal(a, b, label):
    for k in OUTPUTS:
        np.testing.assert_array_equal(a[k].view(np.uint8), b[k].view(np.uint8), err_msg=f"{label} {k}")


@pytest.mark.parametrize("wave", [0, 1], ids=["rayleigh", "love"])
@pytest.mark.parametrize("chunk", [0, 2500], ids=["one_chunk", "chunks_of_2500"])
def test_static_entry_equals_run_device_on_two_calls(env, wave, chunk):
    torch, md = env
    from modisp_amd import workloads
    wl = workloads.precision_batch(B=96, seed=41)
    window = workloads.scan_window(wl.vs)
    B, L = wl.thickness.shape
    T = wl.periods.size
    small = 40
    big_bufs, big = _device_case(torch, wl.thickness, wl.vp, wl.vs, wl.rho, wl.periods, window, 5)
    sm_bufs, sm = _device_case(torch, wl.thickness[:small], wl.vp[:small], wl.vs[:small], wl.rho[:small],
                               wl.periods, window[:small], 5)
    print(f"wave {wave} chunk {chunk}: B={B} then {small}, T={T}", flush=True)
    with md.QeSolver(B, L, T, 5, group_velocity=True, scan="float32") as s:
        s.run_device_batch(B, big, 1e-12, wave)
        s.synchronize()
        want_big = _snapshot(big_bufs)
        s.run_device_batch(small, sm, 1e-12, wave)
        s.synchronize()
        want_small = _snapshot(sm_bufs)
    with md.QeSolver(B, L, T, 5, group_velocity=True, scan="float32", chunk_traces=chunk) as s:
        _scramble(torch, big_bufs)
        s.run_device_static_batch(B, big, 1e-12, wave)
        s.synchronize()
        _equal(_snapshot(big_bufs), want_big, "first call")
        _scramble(torch, sm_bufs)
        s.run_device_static_batch(small, sm, 1e-12, wave)
        s.synchronize()
        _equal(_snapshot(sm_bufs), want_small, "second, smaller call")


def test_static_entry_refuses_a_batch_beyond_its_workspace(env):
    torch, md = env
    from modisp_amd import workloads
    B, modes = 30000, 5   # 4.7 KB of resident inputs per model: 141 MB > 128 MiB
    rng = np.random.default_rng(5)
    h, vp, vs, rho = workloads.monotone_models(rng, B, 13)
    periods = np.geomspace(.5, 60., 100)
    bufs, ptr = _device_case(torch, h, vp, vs, rho, periods, workloads.scan_window(vs), modes)
    _scramble(torch, bufs)
    with md.QeSolver(B, 13, periods.size, modes, group_velocity=True, scan="float32") as s:
        with pytest.raises(ValueError, match="static workspace"):
            s.run_device_static_batch(B, ptr, 1e-12, 0)
        s.synchronize()
    assert np.all(bufs["counts"].cpu().numpy() == -7)   # nothing ran
