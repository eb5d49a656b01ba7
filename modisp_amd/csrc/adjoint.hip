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
//   1. forward sweep, half-space to surface, in one-direction dual numbers of
//      the wavenumber k: the states e_{j+1} entering each layer and the
//      scales s_j go to a per-lane tape in device memory, and dF/dk comes out
//      at the top;
//   2. reverse sweep, surface to half-space: the adjoint row vector
//      l_j = dF/de_j is carried down with the transposed layer matrix,
//      l_{j+1} = l_j C_j / s_j, and at each layer
//      dF/dp_j = l_j (dC_j/dp_j) e_{j+1} / s_j is the derivative of the
//      bilinear form l_j C_j e_{j+1}, taken in reverse mode by
//      rayleigh_layer_adjoint (one pass over the layer's wave factors and
//      their closed-form h and wavenumber derivatives, no dual matrix);
//      Love's 2x2 layer matrix is small enough for dual numbers seeded in
//      two parameters at a time;
//   3. the half-space vector and the water top contribute their own terms
//      (dual numbers; once per root).
// The propagator is recomputed here from the model, not kept from the root
// search. A block (one wavefront) takes `group` consecutive models at once
// and spreads their roots over its lanes; the weighted sums over each
// model's roots, sum_roots w dc/dp, accumulate in LDS and are written once.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>
#include "pipeline.hpp"
#include "secular.hpp"

namespace mdr {

__device__ inline PrepRow adj_row(const AdjointArgs &a, int m, int j) {
  const size_t i = static_cast<size_t>(m) * a.L + j;
  return make_prep_row(a.h[i], a.vp[i], a.vs[i], a.rho[i]);
}

// Layer constants of row r with no seeded parameter.
template <int Wave, class T>
__device__ inline Layer<T> plain_layer(const PrepRow &r, double omega, double iw2) {
  Layer<T> l;
  l.h = T(r.h);
  l.kb = T(omega * r.ivs);
  if (Wave == 0) {
    l.ka = T(omega * r.ivp);
    l.gk = T(r.g2 * iw2);
    l.rho = T(r.rho);
    l.irho = T(r.irho);
    l.irho2 = T(r.irho2);
  } else {
    l.ka = T(0.);
    l.gk = T(0.);
    l.rho = T(0.5 * r.rho * r.g2);
    l.irho = T(r.imu);
    l.irho2 = T(0.);
  }
  l.pad = T(0.);
  return l;
}

// Layer constants of row r in dual numbers, seeded in up to two of the row's
// parameters (0 h, 1 vp, 2 vs, 3 rho; -1 none) as the a and b directions.
template <int Wave>
__device__ inline Layer<Dual2> seeded_layer(const PrepRow &r, double omega, double iw2, int da, int db) {
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
  if (Wave == 0) {
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

// Sum over one Rayleigh root of weight * dc/dp into the LDS accumulators
// g[p * L + j] (p: 0 h, 1 vp, 2 vs, 3 rho).
__device__ void rayleigh_root_adjoint(const AdjointArgs &a, int m, int top, double omega, double c, double weight,
                                      double *tape, int stride, double *g) {
  const int L = a.L;
  const double iw2 = 1. / (omega * omega);
  const double kv = omega / c;
  auto add = [&](int p, int j, double v) { atomicAdd(&g[p * L + j], v); };
  auto tp = [&](int j, int i) -> double & { return tape[(static_cast<size_t>(j) * 6 + i) * stride]; };
  // 1. forward sweep in k-duals, taping e_{j+1} and s_j.
  const Dual1 k(kv, 1.), k2 = k * k;
  Dual1 e[5];
  rayleigh_halfspace(plain_layer<0, Dual1>(adj_row(a, m, L - 1), omega, iw2), k, k2, e);
  for (int j = L - 2; j >= top; --j) {
    const PrepRow r = adj_row(a, m, j);
    for (int i = 0; i < 5; ++i) tp(j, i) = e[i].v;
    tp(j, 5) = rayleigh_layer(plain_layer<0, Dual1>(r, omega, iw2), k, k2, e).v;
  }
  Dual1 F = e[0];
  double lam[5] = {1., 0., 0., 0., 0.};
  double etop[5];
  for (int i = 0; i < 5; ++i) etop[i] = e[i].v;
  if (top) {
    const PrepRow r = adj_row(a, m, 0);
    F = rayleigh_water_top(plain_layer<0, Dual1>(r, omega, iw2), k, e);
    const Wave<double> P = wave_factor(kv, omega * r.ivp, r.h);
    lam[0] = P.c;
    lam[1] = -r.rho * P.w;
  }
  const double Fk = F.a;
  if (!(Fk != 0.) || !(fabs(Fk) < INFINITY)) return;
  const double scale = weight * c / (kv * Fk);
  const Dual2 kk(kv), kk2(kv * kv);
  if (top) {  // the water layer's own parameters (its vs is zero: no vs term)
    const PrepRow r = adj_row(a, m, 0);
    const Dual2 ee[5] = {Dual2(etop[0]), Dual2(etop[1]), Dual2(etop[2]), Dual2(etop[3]), Dual2(etop[4])};
    const Dual2 f1 = rayleigh_water_top(seeded_layer<0>(r, omega, iw2, 0, 1), kk, ee);
    const Dual2 f2 = rayleigh_water_top(seeded_layer<0>(r, omega, iw2, 3, -1), kk, ee);
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
    const double ka = omega * r.ivp, kb = omega * r.ivs, gk = r.g2 * iw2;
    const LayerGrad d = rayleigh_layer_adjoint(r.h, ka, kb, gk, r.rho, r.irho, r.irho2, kv, lam, ein);
    for (int i = 0; i < 5; ++i) lam[i] *= is;
    const double f = scale * is;
    add(0, j, f * d.h);
    add(1, j, f * d.ka * (-ka * r.ivp));                              // dka/dvp = -ka / vp
    add(2, j, f * fma(d.kb, -kb * r.ivs, d.gk * (2. * gk * r.ivs)));  // dkb/dvs = -kb / vs, dgk/dvs = 2 gk / vs
    add(3, j, f * d.rho);
  }
  // 3. the half-space vector (rescaled like the forward sweep's).
  {
    const PrepRow r = adj_row(a, m, L - 1);
    Dual2 eh[5];
    rayleigh_halfspace(seeded_layer<0>(r, omega, iw2, 1, 2), kk, kk2, eh);
    double dvp = 0., dvs = 0.;
    for (int i = 0; i < 5; ++i) { dvp += lam[i] * eh[i].a; dvs += lam[i] * eh[i].b; }
    rayleigh_halfspace(seeded_layer<0>(r, omega, iw2, 3, -1), kk, kk2, eh);
    double drho = 0.;
    for (int i = 0; i < 5; ++i) drho += lam[i] * eh[i].a;
    add(1, L - 1, scale * dvp);
    add(2, L - 1, scale * dvs);
    add(3, L - 1, scale * drho);
  }
}

// Love: SH 2-vector; a water layer is skipped (top). No vp terms.
__device__ void love_root_adjoint(const AdjointArgs &a, int m, int top, double omega, double c, double weight,
                                  double *tape, int stride, double *g) {
  const int L = a.L;
  const double iw2 = 1. / (omega * omega);
  const double kv = omega / c;
  auto add = [&](int p, int j, double v) { atomicAdd(&g[p * L + j], v); };
  auto tp = [&](int j, int i) -> double & { return tape[(static_cast<size_t>(j) * 6 + i) * stride]; };
  const Dual1 k(kv, 1.);
  Dual1 e[2];
  {
    const Layer<Dual1> hs = plain_layer<1, Dual1>(adj_row(a, m, L - 1), omega, iw2);
    e[0] = hs.rho * m_sqrt((k + hs.kb) * m_fabs(k - hs.kb));
    e[1] = Dual1(1.);
    rescale<2>(e);
  }
  for (int j = L - 2; j >= top; --j) {
    const PrepRow r = adj_row(a, m, j);
    tp(j, 0) = e[0].v;
    tp(j, 1) = e[1].v;
    tp(j, 5) = love_layer(plain_layer<1, Dual1>(r, omega, iw2), k, e).v;
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
    const ShMatrix<Dual2> q1 = love_matrix(seeded_layer<1>(r, omega, iw2, 0, 2), kk);
    sh_apply(q1, ein, n);
    const double dh = lam[0] * n[0].a + lam[1] * n[1].a, dvs = lam[0] * n[0].b + lam[1] * n[1].b;
    const ShMatrix<Dual2> q2 = love_matrix(seeded_layer<1>(r, omega, iw2, 3, -1), kk);
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
    const Layer<Dual2> hs = seeded_layer<1>(r, omega, iw2, 2, 3);
    Dual2 eh[2] = {hs.rho * m_sqrt((kk + hs.kb) * m_fabs(kk - hs.kb)), Dual2(1.)};
    rescale<2>(eh);
    add(2, L - 1, scale * (lam[0] * eh[0].a + lam[1] * eh[1].a));
    add(3, L - 1, scale * (lam[0] * eh[0].b + lam[1] * eh[1].b));
  }
}

// One wavefront per block. Block b takes models [m0, m0 + group) for
// m0 = b * group, b * group + gridDim.x * group, ...; the group's
// modes * T cells are spread over the lanes in model-major order.
template <int Wave>
__global__ __launch_bounds__(64) void adjoint_kernel(AdjointArgs a) {
  extern __shared__ double g[];  // [group][4][L]
  const int lane = threadIdx.x;
  const int L = a.L, R = a.modes * a.T, per_model = 4 * L;
  double *tape = a.tape + blockIdx.x * 64 + lane;
  const int stride = a.tape_lanes;
  constexpr double two_pi = 6.283185307179586;
  for (int m0 = blockIdx.x * a.group; m0 < a.B; m0 += gridDim.x * a.group) {
    const int nm = min(a.group, a.B - m0);
    for (int i = lane; i < nm * per_model; i += 64) g[i] = 0.;
    __syncthreads();
    for (int q = lane; q < nm * R; q += 64) {
      const int mi = q / R, cell = q - mi * R;
      const int m = m0 + mi, mode = cell / a.T, f = cell - mode * a.T;
      const size_t o = static_cast<size_t>(m) * R + cell;
      const double c = a.roots[o], wgt = a.weight[o];
      if (!(c == c) || !(wgt == wgt) || wgt == 0.) continue;
      const double period = a.row_periods ? a.periods[static_cast<size_t>(m) * a.T + f] : a.periods[f];
      const int top = a.vs[static_cast<size_t>(m) * L] == 0. ? 1 : 0;
      if (Wave == 0)
        rayleigh_root_adjoint(a, m, top, two_pi / period, c, wgt, tape, stride, g + mi * per_model);
      else
        love_root_adjoint(a, m, top, two_pi / period, c, wgt, tape, stride, g + mi * per_model);
    }
    __syncthreads();
    for (int i = lane; i < nm * per_model; i += 64) {
      const int mi = i / per_model, pj = i - mi * per_model, p = pj / L, j = pj - p * L;
      const size_t o = static_cast<size_t>(m0 + mi) * L + j;
      (p == 0 ? a.gh : p == 1 ? a.gvp : p == 2 ? a.gvs : a.grho)[o] = g[i];
    }
    __syncthreads();
  }
}

// FP64 -> FP32 velocities for callers that take float32 results: one
// grid-stride pass on the device, so half the bytes cross to the host.
__global__ __launch_bounds__(256) void narrow_kernel(const double *in, float *out, size_t n) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = static_cast<float>(in[i]);
}

hipError_t launch_narrow(const double *in, float *out, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(narrow_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, in, out, n);
  return hipGetLastError();
}

hipError_t launch_adjoint(const AdjointArgs &a, int blocks, hipStream_t s) {
  const size_t lds = static_cast<size_t>(a.group) * 4 * a.L * sizeof(double);
  if (a.wave == 0)
    hipLaunchKernelGGL(adjoint_kernel<0>, dim3(blocks), dim3(64), lds, s, a);
  else
    hipLaunchKernelGGL(adjoint_kernel<1>, dim3(blocks), dim3(64), lds, s, a);
  return hipGetLastError();
}

}  // namespace mdr
