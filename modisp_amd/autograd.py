"""torch.autograd binding of the device root solver.

``phase_velocity(thickness, vp, vs, rho, periods, ...)`` returns the phase
velocities [B, modes, T] of the first `modes` roots (NaN where a mode does not
exist) as a float64 CUDA tensor that is differentiable with respect to the
four model tensors. The backward pass calls the adjoint kernel
(``md_rocm_adjoint``), which recomputes each root's propagator in FP64 and
applies the implicit-function rule dc/dp = -(dF/dp)/(dF/dc); it sums
grad_output * dc/dp over the model's roots on the device. Both passes run on
torch's current stream, so no host synchronisation is inserted.

The reference has no gradients; its roots are what ``forward`` computes
(python/modisp/__init__.py:277-507 for the result layout).
"""
from __future__ import annotations

import torch

from . import QeSolver, workloads

_WAVES = {"rayleigh": 0, "love": 1, 0: 0, 1: 1}


def _check(t, name, dims):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{name} must be a CUDA tensor")
    if t.dim() != dims:
        raise ValueError(f"{name} must have {dims} dimensions")


class PhaseVelocity(torch.autograd.Function):
    """c = roots(thickness, vp, vs, rho); see phase_velocity()."""

    @staticmethod
    def forward(ctx, thickness, vp, vs, rho, periods, window, solver, wave, tol):
        B = thickness.shape[0]
        h, a, b, r = (x.detach().to(torch.float64).contiguous() for x in (thickness, vp, vs, rho))
        per = periods.detach().to(torch.float64).contiguous()
        win = window.detach().to(torch.float64).contiguous()
        dev = h.device
        vel = torch.empty((B, solver.nmodes, solver.nf), dtype=torch.float64, device=dev)
        flags = torch.empty((B, solver.nf), dtype=torch.int32, device=dev)
        levels = torch.empty_like(flags)
        counts = torch.empty_like(flags)
        solver.set_stream(torch.cuda.current_stream(dev).cuda_stream)
        solver.run_device_batch(B, dict(thickness=h.data_ptr(), vp=a.data_ptr(), vs=b.data_ptr(), rho=r.data_ptr(),
                                        periods=per.data_ptr(), bounds=win.data_ptr(), velocity=vel.data_ptr(),
                                        flags=flags.data_ptr(), levels=levels.data_ptr(), counts=counts.data_ptr()),
                                tol, wave, row_periods=per.dim() == 2)
        ctx.save_for_backward(h, a, b, r, per, vel)
        ctx.solver, ctx.wave = solver, wave
        ctx.mark_non_differentiable(counts)
        return vel, counts

    @staticmethod
    def backward(ctx, grad_vel, grad_counts):
        h, a, b, r, per, vel = ctx.saved_tensors
        solver = ctx.solver
        w = grad_vel.detach().to(torch.float64).contiguous()
        grads = [torch.empty_like(h) for _ in range(4)]
        solver.set_stream(torch.cuda.current_stream(h.device).cuda_stream)
        solver.adjoint_device_batch(h.shape[0], dict(
            thickness=h.data_ptr(), vp=a.data_ptr(), vs=b.data_ptr(), rho=r.data_ptr(), periods=per.data_ptr(),
            roots=vel.data_ptr(), weight=w.data_ptr(), grad_thickness=grads[0].data_ptr(),
            grad_vp=grads[1].data_ptr(), grad_vs=grads[2].data_ptr(), grad_rho=grads[3].data_ptr()),
            ctx.wave, row_periods=per.dim() == 2)
        return (*grads, None, None, None, None, None)


def phase_velocity(thickness, vp, vs, rho, periods, window=None, nmodes=3, wave="rayleigh", *, solver=None,
                   tol=1e-12, scan="float32"):
    """Differentiable multimode phase velocities [B, nmodes, T] (NaN: no root).

    thickness, vp, vs, rho: [B, L] CUDA tensors (the last row is the
    half-space); periods: [T] or [B, T]; window: [B, 2] search bounds (default
    [0.8 min Vs, Vs_halfspace)). A QeSolver can be passed to reuse its device
    buffers across calls (its geometry must match)."""
    for t, n in ((thickness, "thickness"), (vp, "vp"), (vs, "vs"), (rho, "rho")):
        _check(t, n, 2)
    B, L = thickness.shape
    per = torch.as_tensor(periods, dtype=torch.float64, device=thickness.device)
    if window is None:
        win = torch.from_numpy(workloads.scan_window(vs.detach().cpu().numpy())).to(thickness.device)
    else:
        win = torch.as_tensor(window, dtype=torch.float64, device=thickness.device)
    if solver is None:
        solver = QeSolver(B, L, per.shape[-1], nmodes, scan=scan)
    vel, _ = PhaseVelocity.apply(thickness, vp, vs, rho, per, win, solver, _WAVES[wave], tol)
    return vel
