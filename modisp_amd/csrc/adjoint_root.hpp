// Per-root adjoint sweeps of the phase-velocity roots, dc/d(h, vp, vs, rho)
// of every layer by the implicit-function rule at a converged root,
//
//     dc/dp = -(dF/dp) / (dF/dc) = (dF/dp) c / (k dF/dk),   k = w / c,
//
// shared by the device kernel (modisp_amd/csrc/adjoint.hip, which describes
// the method) and host checks: the model rows, the accumulation of the
// weighted derivatives and the tape are reached through the caller's
// accessors, so the same code runs on a GPU lane (LDS atomics, a
// lane-interleaved tape in device memory) and on the host.
#pragma once
#include "secular.hpp"

namespace mdr {

// Layer constants of row r with no seeded parameter.
template <int Wave, class T, class Row>
MDR_HD Layer<T> plain_layer(const Row &r, double omega, double iw2) {
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
template <int Wave, class Row>
MDR_HD Layer<Dual2> seeded_layer(const Row &r, double omega, double iw2, int da, int db) {
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

// Sum over one Rayleigh root of weight * dc/dp through add(p, j, value)
// (p: 0 h, 1 vp, 2 vs, 3 rho; j the row). row(j) gives row j's prepared
// constants (h, ivp, ivs, g2, rho, irho, irho2, imu); tp(j, i) is the tape
// slot i (0-4 the state entering row j, 5 its scale) of row j.
template <class RowAt, class Add, class Tape>
MDR_HD void rayleigh_root_adjoint(RowAt row, Add add, Tape tp, int L, int top, double omega, double c,
                                  double weight) {
  const double iw2 = 1. / (omega * omega);
  const double kv = omega / c;
  // 1. forward sweep in k-duals, taping e_{j+1} and s_j.
  const Dual1 k(kv, 1.), k2 = k * k;
  Dual1 e[5];
  rayleigh_halfspace(plain_layer<0, Dual1>(row(L - 1), omega, iw2), k, k2, e);
  for (int j = L - 2; j >= top; --j) {
    const auto r = row(j);
    for (int i = 0; i < 5; ++i) tp(j, i) = e[i].v;
    tp(j, 5) = rayleigh_layer(plain_layer<0, Dual1>(r, omega, iw2), k, k2, e).v;
  }
  Dual1 F = e[0];
  double lam[5] = {1., 0., 0., 0., 0.};
  double etop[5];
  for (int i = 0; i < 5; ++i) etop[i] = e[i].v;
  if (top) {
    const auto r = row(0);
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
    const auto r = row(0);
    const Dual2 ee[5] = {Dual2(etop[0]), Dual2(etop[1]), Dual2(etop[2]), Dual2(etop[3]), Dual2(etop[4])};
    const Dual2 f1 = rayleigh_water_top(seeded_layer<0>(r, omega, iw2, 0, 1), kk, ee);
    const Dual2 f2 = rayleigh_water_top(seeded_layer<0>(r, omega, iw2, 3, -1), kk, ee);
    add(0, 0, scale * f1.a);
    add(1, 0, scale * f1.b);
    add(3, 0, scale * f2.a);
  }
  // 2. reverse sweep.
  for (int j = top; j <= L - 2; ++j) {
    const auto r = row(j);
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
    const auto r = row(L - 1);
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
template <class RowAt, class Add, class Tape>
MDR_HD void love_root_adjoint(RowAt row, Add add, Tape tp, int L, int top, double omega, double c, double weight) {
  const double iw2 = 1. / (omega * omega);
  const double kv = omega / c;
  const Dual1 k(kv, 1.);
  Dual1 e[2];
  {
    const Layer<Dual1> hs = plain_layer<1, Dual1>(row(L - 1), omega, iw2);
    e[0] = hs.rho * m_sqrt((k + hs.kb) * m_fabs(k - hs.kb));
    e[1] = Dual1(1.);
    rescale<2>(e);
  }
  for (int j = L - 2; j >= top; --j) {
    const auto r = row(j);
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
    const auto r = row(j);
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
    const auto r = row(L - 1);
    const Layer<Dual2> hs = seeded_layer<1>(r, omega, iw2, 2, 3);
    Dual2 eh[2] = {hs.rho * m_sqrt((kk + hs.kb) * m_fabs(kk - hs.kb)), Dual2(1.)};
    rescale<2>(eh);
    add(2, L - 1, scale * (lam[0] * eh[0].a + lam[1] * eh[1].a));
    add(3, L - 1, scale * (lam[0] * eh[0].b + lam[1] * eh[1].b));
  }
}

}  // namespace mdr
