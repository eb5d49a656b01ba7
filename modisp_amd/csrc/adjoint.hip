// gfx950 adjoint of the phase-velocity roots: dc/d(h, vp, vs, rho) of every
// layer by the implicit-function rule at a converged root,
//
//     dc/dp = -(dF/dp) / (dF/dc) = (dF/dp) c / (k dF/dk),   k = w / c,
//
// (no reference counterpart; the implicit formula is the one the reference
// uses for its group velocity, cpp/common/group_velocity.hpp:2-4). F is the
// normalised secular function of the propagator: every positive factor it
// carries -- the per-layer max-abs rescale and the e^-(p+q) attenuation of
// the compound matrix -- multiplies F, which vanishes at the root, so it
// cancels from the ratio and the factors can be held constant.
//
// Per root (one lane), in FP64:
//   1. forward sweep, half-space to surface, in dual numbers of the
//      wavenumber k: the states e_{j+1} entering each layer and the scales s_j go
//      to a per-lane tape in device memory, and dF/dk comes out at the top;
//   2. reverse sweep, surface to half-space: the adjoint row vector
//      l_j = dF/de_j is carried down with the transposed layer matrix,
//      l_{j+1} = l_j C_j / s_j, and at each layer
//      dF/dp_j = l_j (dC_j/dp_j) e_{j+1} / s_j comes from the layer's compound
//      matrix evaluated with dual numbers seeded in two of its parameters at a
//      time ((h, vp), then (vs, rho));
//   3. the half-space vector and the water top contribute their own terms.
// The propagator is recomputed here from the model, not kept from the root
// search. Each block owns whole models, so the weighted sum over a model's
// roots, sum_roots w dc/dp, accumulates in LDS and is written once.
#include <hip/hip_runtime.h>
#include <cstdint>
#include "pipeline.hpp"
#include "secular.hpp"

namespace mdr {

constexpr int kAdjWaves = 1;  // one wavefront per block: a block owns whole models

__device__ inline PrepRow adj_row(const AdjointArgs &a, int m, int j) {
  const size_t i = static_cast<size_t>(m) * a.L + j;
  return make_prep_row(a.h[i], a.vp[i], a.vs[i], a.rho[i]);
}

// Layer constants of row r in dual numbers, seeded in up to two of the row's
// parameters (0 h, 1 vp, 2 vs, 3 rho; -1 none) as the a and b directions.
__device__ inline Layer<Dual2> seeded_layer(const PrepRow &r, double omega, double iw2, int wave, int da, int db) {
  Layer<Dual2> l;
  const double ka = omega * r.ivp, kb = omega * r.ivs;
  auto d = [&](int which, double dh, double dvp, double dvs, double drho) {
    return which == 0 ? dh : which == 1 ? dvp : which == 2 ? dvs : which == 3 ? drho : 0.;
  };
  auto seed = [&](double v, double dh, double dvp, double dvs, double drho) {
    return Dual2(v, d(da, dh, dvp, dvs, drho), d(db, dh, dvp, dvs, drho));
  };
  l.h = seed(r.h, 1., 0., 0., 0.);
  l.kb = seed(kb, 0., 0., -kb * r.ivs, 0.);
  if (wave == 0) {
    const double gk = r.g2 * iw2;
    l.ka = seed(ka, 0., -ka * r.ivp, 0., 0.);
    l.gk = seed(gk, 0., 0., 2. * gk * r.ivs, 0.);
    l.rho = seed(r.rho, 0., 0., 0., 1.);
    l.irho = seed(r.irho, 0., 0., 0., -r.irho * r.irho);
    l.irho2 = seed(r.irho2, 0., 0., 0., -2. * r.irho2 * r.irho);
  } else {
    const double mu = 0.5 * r.rho * r.g2;
    l.ka = Dual2(0.);
    l.gk = Dual2(0.);
    l.rho = seed(mu, 0., 0., 2. * mu * r.ivs, mu * r.irho);
    l.irho = seed(r.imu, 0., 0., -2. * r.imu * r.ivs, -r.imu * r.irho);
    l.irho2 = Dual2(0.);
  }
  l.pad = Dual2(0.);
  return l;
}

// Sum over one root of weight * dc/dp into the block's LDS accumulators
// g[0..3][j] (h, vp, vs, rho).
__device__ void root_adjoint(const AdjointArgs &a, int m, int top, double omega, double c, double weight,
                             double *tape, int stride, double *g) {
  const int L = a.L;
  const double iw2 = 1. / (omega * omega);
  const Dual2 k(omega / c, 1., 0.);  // dual direction a: the wavenumber
  const double kv = omega / c;
  auto add = [&](int p, int j, double v) { atomicAdd(&g[p * kMaxRows + j], v); };
  auto tp = [&](int j, int i) -> double & { return tape[(static_cast<size_t>(j) * 6 + i) * stride]; };
  if (a.wave == 0) {
    // 1. forward sweep in k-duals, taping e_{j+1} and s_j.
    const Dual2 k2 = k * k;
    Dual2 e[5];
    {
      const PrepRow r = adj_row(a, m, L - 1);
      rayleigh_halfspace(seeded_layer(r, omega, iw2, 0, -1, -1), k, k2, e);
    }
    for (int j = L - 2; j >= top; --j) {
      const PrepRow r = adj_row(a, m, j);
      for (int i = 0; i < 5; ++i) tp(j, i) = e[i].v;
      tp(j, 5) = rayleigh_layer(seeded_layer(r, omega, iw2, 0, -1, -1), k, k2, e).v;
    }
    Dual2 F = e[0];
    double lam[5] = {1., 0., 0., 0., 0.};
    double etop[5];
    for (int i = 0; i < 5; ++i) etop[i] = e[i].v;
    if (top) {
      const PrepRow r = adj_row(a, m, 0);
      F = rayleigh_water_top(seeded_layer(r, omega, iw2, 0, -1, -1), k, e);
      const Wave<double> P = wave_factor(kv, omega * r.ivp, r.h);
      lam[0] = P.c;
      lam[1] = -r.rho * P.w;
      lam[2] = lam[3] = lam[4] = 0.;
    }
    const double Fk = F.a;
    if (!(Fk != 0.) || !(fabs(Fk) < INFINITY)) return;
    const double scale = weight * c / (kv * Fk);
    const double kd = kv, k2d = kv * kv;
    if (top) {  // the water layer's own parameters (its vs is zero: no vs term)
      const PrepRow r = adj_row(a, m, 0);
      const Dual2 kk(kd);
      const Dual2 ee[5] = {Dual2(etop[0]), Dual2(etop[1]), Dual2(etop[2]), Dual2(etop[3]), Dual2(etop[4])};
      const Dual2 f1 = rayleigh_water_top(seeded_layer(r, omega, iw2, 0, 0, 1), kk, ee);
      const Dual2 f2 = rayleigh_water_top(seeded_layer(r, omega, iw2, 0, 3, -1), kk, ee);
      add(0, 0, scale * f1.a);
      add(1, 0, scale * f1.b);
      add(3, 0, scale * f2.a);
    }
    // 2. reverse sweep.
    for (int j = top; j <= L - 2; ++j) {
      const PrepRow r = adj_row(a, m, j);
      double ein[5];
      for (int i = 0; i < 5; ++i) ein[i] = tp(j, i);
      const double is = 1. / tp(j, 5);
      const Dual2 kk(kd), kk2(k2d);
      Dual2 n[5];
      const Compound<Dual2> q1 = rayleigh_compound(seeded_layer(r, omega, iw2, 0, 0, 1), kk, kk2);
      compound_apply(q1, ein, n);
      double dh = 0., dvp = 0.;
      for (int i = 0; i < 5; ++i) { dh += lam[i] * n[i].a; dvp += lam[i] * n[i].b; }
      const Compound<Dual2> q2 = rayleigh_compound(seeded_layer(r, omega, iw2, 0, 2, 3), kk, kk2);
      compound_apply(q2, ein, n);
      double dvs = 0., drho = 0.;
      for (int i = 0; i < 5; ++i) { dvs += lam[i] * n[i].a; drho += lam[i] * n[i].b; }
      add(0, j, scale * dh * is);
      add(1, j, scale * dvp * is);
      add(2, j, scale * dvs * is);
      add(3, j, scale * drho * is);
      const Compound<double> q{q1.A.v, q1.B.v, q1.C.v, q1.D.v, q1.E.v, q1.F.v, q1.G.v, q1.H.v,
                               q1.I.v, q1.J.v, q1.K.v, q1.Lq.v, q1.M.v, q1.N.v, q1.O.v, q1.t.v};
      compound_apply_transpose(q, lam);
      for (int i = 0; i < 5; ++i) lam[i] *= is;
    }
    // 3. the half-space vector (rescaled like the forward sweep's).
    {
      const PrepRow r = adj_row(a, m, L - 1);
      const Dual2 kk(kd), kk2(k2d);
      Dual2 eh[5];
      rayleigh_halfspace(seeded_layer(r, omega, iw2, 0, 1, 2), kk, kk2, eh);
      double dvp = 0., dvs = 0.;
      for (int i = 0; i < 5; ++i) { dvp += lam[i] * eh[i].a; dvs += lam[i] * eh[i].b; }
      rayleigh_halfspace(seeded_layer(r, omega, iw2, 0, 3, -1), kk, kk2, eh);
      double drho = 0.;
      for (int i = 0; i < 5; ++i) drho += lam[i] * eh[i].a;
      add(1, L - 1, scale * dvp);
      add(2, L - 1, scale * dvs);
      add(3, L - 1, scale * drho);
    }
  } else {
    // Love: SH 2-vector; a water layer is skipped (top).
    Dual2 e[2];
    {
      const PrepRow r = adj_row(a, m, L - 1);
      const Layer<Dual2> hs = seeded_layer(r, omega, iw2, 1, -1, -1);
      e[0] = hs.rho * m_sqrt((k + hs.kb) * m_fabs(k - hs.kb));
      e[1] = Dual2(1.);
      rescale<2>(e);
    }
    for (int j = L - 2; j >= top; --j) {
      const PrepRow r = adj_row(a, m, j);
      tp(j, 0) = e[0].v;
      tp(j, 1) = e[1].v;
      tp(j, 5) = love_layer(seeded_layer(r, omega, iw2, 1, -1, -1), k, e).v;
    }
    const double Fk = e[0].a;
    if (!(Fk != 0.) || !(fabs(Fk) < INFINITY)) return;
    const double scale = weight * c / (kv * Fk);
    double lam[2] = {1., 0.};
    const Dual2 kk(kv);
    for (int j = top; j <= L - 2; ++j) {
      const PrepRow r = adj_row(a, m, j);
      const double ein[2] = {tp(j, 0), tp(j, 1)};
      const double is = 1. / tp(j, 5);
      Dual2 n[2];
      const ShMatrix<Dual2> q1 = love_matrix(seeded_layer(r, omega, iw2, 1, 0, 2), kk);
      sh_apply(q1, ein, n);
      const double dh = lam[0] * n[0].a + lam[1] * n[1].a, dvs = lam[0] * n[0].b + lam[1] * n[1].b;
      const ShMatrix<Dual2> q2 = love_matrix(seeded_layer(r, omega, iw2, 1, 3, -1), kk);
      sh_apply(q2, ein, n);
      const double drho = lam[0] * n[0].a + lam[1] * n[1].a;
      add(0, j, scale * dh * is);
      add(2, j, scale * dvs * is);
      add(3, j, scale * drho * is);
      const ShMatrix<double> q{q1.c.v, q1.mx.v, q1.wi.v};
      sh_apply_transpose(q, lam);
      lam[0] *= is;
      lam[1] *= is;
    }
    {
      const PrepRow r = adj_row(a, m, L - 1);
      const Layer<Dual2> hs = seeded_layer(r, omega, iw2, 1, 2, 3);
      Dual2 eh[2] = {hs.rho * m_sqrt((kk + hs.kb) * m_fabs(kk - hs.kb)), Dual2(1.)};
      rescale<2>(eh);
      add(2, L - 1, scale * (lam[0] * eh[0].a + lam[1] * eh[1].a));
      add(3, L - 1, scale * (lam[0] * eh[0].b + lam[1] * eh[1].b));
    }
  }
}

__global__ __launch_bounds__(64 * kAdjWaves) void adjoint_kernel(AdjointArgs a) {
  __shared__ double g[4 * kMaxRows];
  const int lane = threadIdx.x;
  const int R = a.modes * a.T;
  double *tape = a.tape + blockIdx.x * 64 + lane;
  const int stride = a.tape_lanes;
  constexpr double two_pi = 6.283185307179586;
  for (int m = blockIdx.x; m < a.B; m += gridDim.x) {
    for (int i = lane; i < 4 * a.L; i += 64) g[(i / a.L) * kMaxRows + i % a.L] = 0.;
    __syncthreads();
    const int top = a.vs[static_cast<size_t>(m) * a.L] == 0. ? 1 : 0;
    for (int r = lane; r < R; r += 64) {
      const int mode = r / a.T, f = r % a.T;
      const size_t cell = (static_cast<size_t>(m) * a.modes + mode) * a.T + f;
      const double c = a.roots[cell], wgt = a.weight[cell];
      if (!(c == c) || !(wgt == wgt) || wgt == 0.) continue;
      const double period = a.row_periods ? a.periods[static_cast<size_t>(m) * a.T + f] : a.periods[f];
      root_adjoint(a, m, top, two_pi / period, c, wgt, tape, stride, g);
    }
    __syncthreads();
    for (int i = lane; i < 4 * a.L; i += 64) {
      const int p = i / a.L, j = i % a.L;
      const size_t o = static_cast<size_t>(m) * a.L + j;
      const double v = g[p * kMaxRows + j];
      (p == 0 ? a.gh : p == 1 ? a.gvp : p == 2 ? a.gvs : a.grho)[o] = v;
    }
    __syncthreads();
  }
}

hipError_t launch_adjoint(const AdjointArgs &a, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(adjoint_kernel, dim3(blocks), dim3(64 * kAdjWaves), 0, s, a);
  return hipGetLastError();
}

}  // namespace mdr
