"""Seeded synthetic layered-model workloads (km, km/s, g/cm^3, s).

These generate the five benchmark/parity workloads of this backend. Models
are [B, L] float64 arrays whose last row is the half-space (thickness 0);
the layout matches the reference's batch convention
(python/modisp/models.py:23-47, 224-270): zero-thickness padding rows are
exactly neutral, so ragged batches may be padded.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

__all__ = ["Workload", "precision_batch", "plain_batch", "uneven_batch", "pack_ragged", "gradient_batch", "sweep_batch",
           "monotone_models", "WORKLOADS"]


@dataclass
class Workload:
    name: str
    thickness: np.ndarray
    vp: np.ndarray
    vs: np.ndarray
    rho: np.ndarray
    periods: np.ndarray            # [T] shared, or list of per-model arrays
    waves: tuple = ("rayleigh",)
    nmodes: int | np.ndarray = 3   # scalar, or per-model counts
    layers: np.ndarray | None = None   # per-model row counts for ragged batches
    extra: dict = field(default_factory=dict)

    @property
    def B(self):
        return self.thickness.shape[0]

    def subset(self, index):
        index = np.asarray(index)
        per = self.periods if isinstance(self.periods, np.ndarray) else [self.periods[i] for i in index]
        nm = self.nmodes if np.isscalar(self.nmodes) else np.asarray(self.nmodes)[index]
        lay = None if self.layers is None else self.layers[index]
        extra = {k: (v[index] if isinstance(v, np.ndarray) and v.shape[:1] == (self.B,) else v)
                 for k, v in self.extra.items()}
        return Workload(self.name, self.thickness[index], self.vp[index], self.vs[index], self.rho[index], per,
                        self.waves, nm, lay, extra)


def _density(vp):
    return 0.32 * vp + 0.77


def monotone_models(rng, B, L, h_range=(0.5, 5.), jitter=0.1):
    """Monotone-increasing Vs from U(1.2,2.0) at the top to U(3.8,4.6) in the
    half-space, linear between with N(0, jitter) noise, Vp = 1.732 Vs,
    rho = 0.32 Vp + 0.77; thickness U(h_range), half-space thickness 0."""
    top = rng.uniform(1.2, 2.0, B)
    bottom = rng.uniform(3.8, 4.6, B)
    x = np.linspace(0., 1., L)[None, :]
    vs = top[:, None] + (bottom - top)[:, None] * x
    noise = rng.normal(0., jitter, (B, L))
    noise[:, 0] = 0.
    noise[:, -1] = 0.
    vs = np.maximum.accumulate(vs + noise, axis=1)
    vs[:, -1] = np.maximum(vs[:, -1], vs[:, -2])
    vp = 1.732 * vs
    h = rng.uniform(h_range[0], h_range[1], (B, L))
    h[:, -1] = 0.
    return h, vp, vs, _density(vp)


def plain_batch(B=1024, seed=0):
    """Plain FP64 CPU-runnable case: 1024 models, 6 layers + half-space, 60
    log-spaced periods 1-50 s, Rayleigh, 3 modes, phase velocity only."""
    rng = np.random.default_rng(seed)
    h, vp, vs, rho = monotone_models(rng, B, 7)
    return Workload("plain_fp64", h, vp, vs, rho, np.geomspace(1., 50., 60), ("rayleigh",), 3)


def precision_batch(B=16384, seed=1):
    """Precision form: 16384 models, 12 layers + half-space, 100 log-spaced
    periods 0.5-60 s, Rayleigh and Love, 5 modes each, FP64 group velocity;
    512-sample FP32 scan over [0.8 min(Vs), Vs_halfspace]."""
    rng = np.random.default_rng(seed)
    h, vp, vs, rho = monotone_models(rng, B, 13)
    return Workload("precision_fp32_scan", h, vp, vs, rho, np.geomspace(0.5, 60., 100), ("rayleigh", "love"), 5,
                    extra=dict(nc=512))


def uneven_batch(B=65536, seed=2):
    """Uneven batching: finite-layer counts U{3..40}, 20% with one low-velocity
    layer (Vs x U(0.7,0.9)), thickness log-uniform 0.05-10 km, per-model period
    counts U{20..200} log-spaced 0.2-100 s, Rayleigh, per-model modes U{1..10}.
    Returned padded to the largest row count; `layers` holds the true counts."""
    rng = np.random.default_rng(seed)
    nfinite = rng.integers(3, 41, B)
    L = int(nfinite.max()) + 1
    top = rng.uniform(1.2, 2.0, B)
    bottom = rng.uniform(3.8, 4.6, B)
    h = np.zeros((B, L))
    vs = np.zeros((B, L))
    for i in range(B):
        n = nfinite[i]
        x = np.linspace(0., 1., n + 1)
        v = top[i] + (bottom[i] - top[i]) * x
        v[1:-1] += rng.normal(0., 0.1, n - 1)
        v = np.maximum.accumulate(v)
        h[i, :n] = np.exp(rng.uniform(np.log(0.05), np.log(10.), n))
        vs[i, :n] = v[:-1]
        vs[i, n:] = v[-1]          # padding rows repeat the half-space, thickness 0
    lvz = rng.random(B) < 0.2
    for i in np.flatnonzero(lvz):
        j = rng.integers(0, nfinite[i])
        vs[i, j] *= rng.uniform(0.7, 0.9)
    vp = 1.732 * vs
    nper = rng.integers(20, 201, B)
    periods = [np.geomspace(0.2, 100., n) for n in nper]
    modes = rng.integers(1, 11, B)
    return Workload("uneven_batch", h, vp, vs, _density(vp), periods, ("rayleigh",), modes, nfinite + 1,
                    extra=dict(lvz=lvz))


def pack_ragged(wl, period_step=1):
    """Packs a ragged workload (per-model `layers`, period lists and mode
    counts) into the arrays of the ragged entry points: thickness, vp, vs,
    rho [rows], row_offsets and period_offsets [B+1], periods, modes [B] and
    the window bounds [B, 2] = [0.8 min Vs, nextafter(Vs_halfspace, 0)] of
    each model's own rows. Every `period_step`-th period is kept."""
    B = wl.B
    lay = np.asarray(wl.layers if wl.layers is not None else np.full(B, wl.thickness.shape[1]), dtype=np.int64)
    keep = np.arange(wl.thickness.shape[1])[None, :] < lay[:, None]
    out = {k: getattr(wl, k)[keep] for k in ("thickness", "vp", "vs", "rho")}
    per = [np.asarray(p)[::period_step] for p in (wl.periods if isinstance(wl.periods, list) else [wl.periods] * B)]
    out["row_offsets"] = np.concatenate([[0], np.cumsum(lay)]).astype(np.int64)
    out["period_offsets"] = np.concatenate([[0], np.cumsum([p.size for p in per])]).astype(np.int64)
    out["periods"] = np.concatenate(per)
    out["modes"] = (np.full(B, wl.nmodes) if np.isscalar(wl.nmodes) else np.asarray(wl.nmodes)).astype(np.int32)
    vs_min = np.where(keep, wl.vs, np.inf).min(axis=1)
    half = wl.vs[np.arange(B), lay - 1]
    out["bounds"] = np.stack([0.8 * vs_min, np.nextafter(half, 0.)], axis=1)
    return out


def gradient_batch(B=4096, seed=3):
    """Differentiation: 4096 models of 20 layers + half-space (monotone recipe),
    50 log-spaced periods 2-40 s, Rayleigh fundamental + 2 overtones; target
    models have Vs perturbed by N(0, 0.05)."""
    rng = np.random.default_rng(seed)
    h, vp, vs, rho = monotone_models(rng, B, 21)
    target_vs = vs + rng.normal(0., 0.05, vs.shape)
    target_vs = np.maximum(target_vs, 0.2)
    return Workload("adjoint", h, vp, vs, rho, np.geomspace(2., 40., 50), ("rayleigh",), 3,
                    extra=dict(target_vs=target_vs))


def sweep_batch(B=262144, seed=4):
    """Inversion-scale sweep: 30 layers + half-space, Vs a Gaussian random walk
    (sigma 0.15) clipped to [0.5, 4.8], Vp/Vs U(1.65, 2.0), rho = 0.32 Vp + 0.77,
    200 log-spaced periods 0.5-100 s, Rayleigh and Love, 4 modes, group on."""
    rng = np.random.default_rng(seed)
    L = 31
    start = rng.uniform(1.0, 3.0, B)
    steps = rng.normal(0., 0.15, (B, L))
    steps[:, 0] = 0.
    vs = np.clip(start[:, None] + np.cumsum(steps, axis=1), 0.5, 4.8)
    vp = vs * rng.uniform(1.65, 2.0, (B, L))
    h = rng.uniform(0.5, 5., (B, L))
    h[:, -1] = 0.
    return Workload("sweep", h, vp, vs, _density(vp), np.geomspace(0.5, 100., 200), ("rayleigh", "love"), 4)


WORKLOADS = {"plain_fp64": plain_batch, "precision_fp32_scan": precision_batch, "uneven_batch": uneven_batch,
             "adjoint": gradient_batch, "sweep": sweep_batch}


def scan_window(vs):
    """Per-model scan window [0.8 min(Vs), nextafter(Vs_halfspace, 0)]."""
    vs = np.asarray(vs, dtype=np.float64)
    solid = np.where(vs > 0, vs, np.inf)
    lo = 0.8 * solid.min(axis=1)
    hi = np.nextafter(vs[:, -1], 0.)
    return np.stack([lo, hi], axis=1)
