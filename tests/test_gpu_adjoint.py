"""Device adjoint dc/d(h, vp, vs, rho) behind torch.autograd.

Bars: the weighted gradient sum equals the oracle's FP64 implicit phase
gradient at the same roots (1e-8 relative to the model's largest entry), and
forward+adjoint gradients of a misfit loss agree with central finite
differences of the device forward solve (relative parameter step 1e-6) to
1e-6 relative.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

MODEL = ("thickness", "vp", "vs", "rho")


@pytest.fixture(scope="module")
def torch_md():
    import torch
    import modisp_amd
    if not modisp_amd.available(0):
        pytest.fail("no gfx950 device visible to the gpu tests")
    modisp_amd.set_device(0)
    return torch, modisp_amd


def _water_models(B, rng):
    from modisp_amd import workloads
    h, vp, vs, rho = workloads.monotone_models(rng, B, 8)
    vs[:, 0], vp[:, 0], rho[:, 0], h[:, 0] = 0., 1.5, 1.03, rng.uniform(.5, 3., B)
    return h, vp, vs, rho


def _cases():
    from modisp_amd import workloads
    wl = workloads.gradient_batch(B=6, seed=3)
    yield "gradient-rayleigh", wl.thickness, wl.vp, wl.vs, wl.rho, wl.periods[::10], 3, 0
    yield "gradient-love", wl.thickness, wl.vp, wl.vs, wl.rho, wl.periods[::10], 3, 1
    h, vp, vs, rho = _water_models(4, np.random.default_rng(11))
    yield "water-rayleigh", h, vp, vs, rho, np.array([1., 4., 12.]), 3, 0
    # a root 1e-9 below a layer's shear speed (r -> 0 in that layer)
    for wave, name in ((0, "shear-crossing-rayleigh"), (1, "shear-crossing-love")):
        h, vp, vs, rho, period, _ = oracle.shear_speed_crossing(wave)
        yield name, h[None], vp[None], vs[None], rho[None], np.array([period]), 1, wave


@pytest.mark.parametrize("case", list(_cases()), ids=lambda c: c[0])
def test_adjoint_matches_oracle_implicit_gradient(torch_md, case):
    torch, md = torch_md
    from modisp_amd.autograd import phase_velocity
    name, h, vp, vs, rho, periods, modes, wave = case
    dev = torch.device("cuda", 0)
    T = [torch.tensor(x, device=dev, requires_grad=True) for x in (h, vp, vs, rho)]
    b = md.bounds(h, vp, vs, rho, wave)
    window = np.stack([np.maximum(b[:, 0], 0.8 * np.where(vs > 0, vs, np.inf).min(axis=1)), b[:, 1]], axis=1)
    c = phase_velocity(*T, periods, window, nmodes=modes, wave=wave)
    rng = np.random.default_rng(5)
    w = torch.tensor(rng.normal(size=tuple(c.shape)), device=dev)
    loss = torch.nansum(c * w)
    loss.backward()
    got = np.stack([t.grad.cpu().numpy() for t in T], axis=-1)      # [B, L, 4]
    roots = c.detach().cpu().numpy()
    wn = w.cpu().numpy()
    assert np.isfinite(roots).sum() >= roots.shape[0] * len(periods)
    for i in range(h.shape[0]):
        want = np.zeros((h.shape[1], 4))
        for m in range(modes):
            for f, p in enumerate(periods):
                if np.isfinite(roots[i, m, f]):
                    want += wn[i, m, f] * oracle.phase_gradient(h[i], vp[i], vs[i], rho[i], p, roots[i, m, f], wave)
        if wave == 0 and vs[i, 0] == 0.:
            want[0, 2] = 0.  # the water layer has no shear speed to differentiate
        scale = np.abs(want).max()
        np.testing.assert_allclose(got[i], want, rtol=0, atol=1e-8 * scale, err_msg=f"{name} model {i}")


@pytest.mark.parametrize("wave", [0, 1], ids=["rayleigh", "love"])
def test_forward_adjoint_matches_central_differences(torch_md, wave):
    torch, md = torch_md
    from modisp_amd import QeSolver, workloads
    from modisp_amd.autograd import phase_velocity
    wl = workloads.gradient_batch(B=3, seed=9)
    periods = wl.periods[::12]
    dev = torch.device("cuda", 0)
    base = [np.array(getattr(wl, k)) for k in MODEL]
    window = workloads.scan_window(wl.vs)
    B, L = wl.thickness.shape
    solver = QeSolver(B, L, periods.size, 3)

    def loss_of(arrays, grad=False):
        T = [torch.tensor(x, device=dev, requires_grad=grad) for x in arrays]
        c = phase_velocity(*T, periods, window, nmodes=3, wave=wave, solver=solver, tol=1e-14)
        ok = torch.isfinite(c) & torch.isfinite(target_t)
        loss = ((torch.where(ok, c, 0.) - torch.where(ok, target_t, 0.)) ** 2).sum()
        if grad:
            loss.backward()
            return loss.item(), [t.grad.cpu().numpy() for t in T]
        return loss.item()

    with torch.no_grad():   # targets: the same models with Vs perturbed (the workload's recipe)
        Tt = [torch.tensor(x, device=dev) for x in (base[0], base[1], wl.extra["target_vs"], base[3])]
        target_t = phase_velocity(*Tt, periods, window, nmodes=3, wave=wave, tol=1e-14)
    _, grads = loss_of(base, grad=True)
    rng = np.random.default_rng(21 + wave)
    checked = 0
    while checked < 16:
        p, i, j = rng.integers(0, 4), rng.integers(0, B), rng.integers(0, L)
        if j == L - 1 and p in (0, 2):
            continue   # half-space thickness is not a parameter; its Vs moves the fixed window's top
        if wave == 1 and p == 1:
            continue   # Love roots do not depend on vp
        x0 = base[p][i, j]
        step = 1e-6 * abs(x0)
        up = [a.copy() for a in base]
        dn = [a.copy() for a in base]
        up[p][i, j] = x0 + step
        dn[p][i, j] = x0 - step
        fd = (loss_of(up) - loss_of(dn)) / (2 * step)
        g = grads[p][i, j]
        scale = max(abs(fd), 1e-3 * np.abs(grads[p][i]).max(), 1e-12)
        assert abs(g - fd) <= 1e-6 * scale + 1e-9, (p, i, j, g, fd)
        checked += 1
