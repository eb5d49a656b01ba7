// Reverse-mode Rayleigh layer derivative used by the root adjoint
// (rayleigh_layer_adjoint, modisp_amd/csrc/secular.hpp) against
//  * forward-mode dual numbers through the compound matrix itself:
//    S = l^T C e with C from rayleigh_compound<Dual2>, seeded in two layer
//    constants at a time;
//  * central differences of S in long double (for gk and rho, whose
//    derivatives lose digits to cancellation in any double evaluation).
// Samples cover both wave factors propagating, P evanescent and S
// propagating, both evanescent, k exactly at the P wavenumber, and c within
// 1e-7 of the shear speed (the small-z series of the wave factors; thin
// layers at long periods reach it too).
// Prints: worst relative difference reverse vs forward for (h, ka, kb), the
// same for (gk, rho), worst relative difference of the carried row l^T C,
// and the worst relative errors of reverse and of forward mode against the
// long-double differences (gk and rho).
#include <cmath>
#include <cstdio>
#include <random>

// long double instantiations of the secular templates (declared before the
// templates, which look these names up where they are defined)
namespace mdr {
inline long double m_sqrt(long double x) { return sqrtl(x); }
inline void m_sincos(long double x, long double *s, long double *c) { *s = sinl(x); *c = cosl(x); }
inline long double m_exp(long double x) { return expl(x); }
inline long double m_rcp(long double x) { return 1.0L / x; }
inline long double m_fabs(long double x) { return fabsl(x); }
inline long double m_fma(long double a, long double b, long double c) { return a * b + c; }
inline void root_pair(long double r2, long double *r, long double *ir) { *r = sqrtl(r2); *ir = 1.0L / *r; }
}  // namespace mdr
#include "../modisp_amd/csrc/secular.hpp"

using mdr::Dual2;
using LD = long double;

static Dual2 bilinear(const mdr::Compound<Dual2> &q, const mdr::CompoundWeights &w) {
  return q.A * w.A + q.B * w.B + q.C * w.C + q.D * w.D + q.E * w.E + q.F * w.F + q.G * w.G + q.H * w.H +
         q.I * w.I + q.J * w.J + q.K * w.K + q.Lq * w.Lq + q.M * w.M + q.N * w.N + q.O * w.O;
}

// S = l^T C e in long double.
static LD bilinear_ld(double h, double ka, double kb, LD gk, LD rho, double k, const double l[5], const double e[5]) {
  const LD irho = 1.0L / rho, kk = k;
  const mdr::Compound<LD> q =
      mdr::rayleigh_compound(mdr::Layer<LD>{(LD)h, (LD)ka, (LD)kb, gk, rho, irho, irho * irho, 0.0L}, kk, kk * kk);
  LD n[5];
  mdr::compound_apply(q, e, n);
  LD s = 0;
  for (int i = 0; i < 5; ++i) s += (LD)l[i] * n[i];
  return s;
}

int main() {
  std::mt19937_64 g(11);
  auto U = [&](double a, double b) { return std::uniform_real_distribution<double>(a, b)(g); };
  std::normal_distribution<double> N(0., 1.);
  double worst_hk = 0., worst_gr = 0., worst_l = 0., fd_rev = 0., fd_fwd = 0.;
  for (int it = 0; it < 100000; ++it) {
    const double omega = 6.283185307179586 / U(1., 100.);
    const double vp = U(1.5, 8.), vs = vp * U(0.3, 0.7), rho = U(1.5, 3.5);
    const double h = std::exp(U(std::log(0.01), std::log(50.)));
    const double ka = omega / vp, kb = omega / vs, gk = 2. * vs * vs / (omega * omega);
    const int regime = it % 5;
    const double c = regime == 0 ? vp * U(1.01, 1.5) : regime == 1 ? U(vs * 1.001, vp * 0.999)
                   : regime == 4 ? vs * (1. + 1e-7 * U(-1., 1.)) : vs * U(0.5, 0.999);
    const double k = regime == 3 ? ka : omega / c;
    const double irho = 1. / rho, irho2 = irho * irho, k2 = k * k;
    double l[5], e[5];
    for (int i = 0; i < 5; ++i) { l[i] = N(g); e[i] = N(g); }
    const mdr::CompoundWeights w = mdr::compound_weights(l, e, -2. * k2);
    // forward mode: (h, ka), (kb, gk), (rho with 1/rho and 1/rho^2)
    auto layer = [&](int da, int db) {
      auto s = [&](double v, int which, double dv) { return Dual2(v, da == which ? dv : 0., db == which ? dv : 0.); };
      return mdr::Layer<Dual2>{s(h, 0, 1.), s(ka, 1, 1.), s(kb, 2, 1.), s(gk, 3, 1.), s(rho, 4, 1.),
                               s(irho, 4, -irho * irho), s(irho2, 4, -2. * irho2 * irho), Dual2(0.)};
    };
    const Dual2 kd(k), k2d(k2);
    const mdr::Compound<Dual2> q1 = mdr::rayleigh_compound(layer(0, 1), kd, k2d);
    const Dual2 s1 = bilinear(q1, w);
    const Dual2 s2 = bilinear(mdr::rayleigh_compound(layer(2, 3), kd, k2d), w);
    const Dual2 s3 = bilinear(mdr::rayleigh_compound(layer(4, -1), kd, k2d), w);
    const double fwd[5] = {s1.a, s1.b, s2.a, s2.b, s3.a};
    // the carried row from the forward-mode entry values (the values the
    // adjoint's forward sweep tapes)
    double lt[5] = {l[0], l[1], l[2], l[3], l[4]};
    mdr::compound_apply_transpose(mdr::Compound<double>{q1.A.v, q1.B.v, q1.C.v, q1.D.v, q1.E.v, q1.F.v, q1.G.v,
                                                        q1.H.v, q1.I.v, q1.J.v, q1.K.v, q1.Lq.v, q1.M.v, q1.N.v,
                                                        q1.O.v, q1.t.v},
                                  lt);
    double lr[5] = {l[0], l[1], l[2], l[3], l[4]};
    const mdr::LayerGrad d = mdr::rayleigh_layer_adjoint(h, ka, kb, gk, rho, irho, irho2, k, lr, e);
    const double rev[5] = {d.h, d.ka, d.kb, d.gk, d.rho};
    double mag = 0.;
    for (int i = 0; i < 5; ++i) mag += std::fabs(fwd[i]);
    for (int i = 0; i < 5; ++i) {
      const double rel = std::fabs(rev[i] - fwd[i]) / (std::fabs(fwd[i]) + 1e-12 * mag + 1e-300);
      if (i < 3) worst_hk = std::fmax(worst_hk, rel);
      else worst_gr = std::fmax(worst_gr, rel);
    }
    double ln = 0.;
    for (int i = 0; i < 5; ++i) ln = std::fmax(ln, std::fabs(lt[i]));
    for (int i = 0; i < 5; ++i) worst_l = std::fmax(worst_l, std::fabs(lr[i] - lt[i]) / (ln + 1e-300));
    if (it % 10 == 0) {
      const LD dg = 1e-7L * gk, dr = 1e-7L * rho;
      const LD fg = (bilinear_ld(h, ka, kb, gk + dg, rho, k, l, e) - bilinear_ld(h, ka, kb, gk - dg, rho, k, l, e)) /
                    (2 * dg);
      const LD fr = (bilinear_ld(h, ka, kb, gk, rho + dr, k, l, e) - bilinear_ld(h, ka, kb, gk, rho - dr, k, l, e)) /
                    (2 * dr);
      fd_rev = std::fmax(fd_rev, (double)std::fmax(fabsl(d.gk - fg) / fabsl(fg), fabsl(d.rho - fr) / fabsl(fr)));
      fd_fwd = std::fmax(fd_fwd, (double)std::fmax(fabsl(fwd[3] - fg) / fabsl(fg), fabsl(fwd[4] - fr) / fabsl(fr)));
    }
  }
  std::printf("%.3e %.3e %.3e %.3e %.3e\n", worst_hk, worst_gr, worst_l, fd_rev, fd_fwd);
  return 0;
}
