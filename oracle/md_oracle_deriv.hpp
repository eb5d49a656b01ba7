// TEST INFRASTRUCTURE ONLY -- the oracle (see md_oracle.hpp header).
//
// Derivatives of the secular functions:
//  * group_velocity_fd: restatement of the reference's finite-difference
//    implicit group velocity (reference cpp/common/group_velocity.hpp:9-104)
//    and its neighbour window (cpp/cpu/solver.cpp:96-103);
//  * forward-mode (one-direction dual number) derivatives of F with respect
//    to k, omega and every layer parameter, with the positive per-layer
//    attenuation and normalization factors held constant, so they equal the
//    derivatives of F at a root (where F = 0) and stay well conditioned near
//    a body-wave speed (wave_parts). The reference has no analytic
//    derivative; these are this oracle's own, checked in tests against closed
//    forms (half-space U=c, two-layer Love implicit derivative), against
//    Richardson-extrapolated central differences of the released evaluator at
//    roots, and against differences of the dispersion curve itself next to a
//    layer's shear speed. They pin the GPU's analytic group velocity and
//    adjoint dc/dm.
#pragma once
#include "md_oracle.hpp"

namespace mdo {

// ---------------------------------------------------------------------------
// Reference FD group velocity.
// ---------------------------------------------------------------------------
enum : unsigned {
  GROUP_MISSING_ROOT = 1, GROUP_POLISH_FAILED = 2, GROUP_DERIVATIVE_UNSTABLE = 4,
  GROUP_ILL_CONDITIONED = 8, GROUP_CLOSE_ROOT = 16, GROUP_NONFINITE = 32
};
struct GroupResult {
  double velocity = NAN, slope = NAN;
  unsigned flags = 0;
  int evaluations = 0;
};

inline void group_window(const double *roots, int nr, int j, double lower, double upper, double rtol,
                         double &lo, double &hi, unsigned &flag) {
  const double c = roots[j];
  lo = j ? .5 * (roots[j - 1] + c) : lower;
  hi = j + 1 < nr ? .5 * (c + roots[j + 1]) : upper;
  const bool close = (j && c - roots[j - 1] <= rtol * c) || (j + 1 < nr && roots[j + 1] - c <= rtol * roots[j + 1]);
  flag = close ? GROUP_CLOSE_ROOT : 0u;
}

inline double group_shape(const Medium &m, double omega, double c, int &evaluations) {
  ++evaluations;
  return shape_evaluate(m, omega, c).g;
}

inline GroupResult group_velocity_fd(const Medium &source, double omega, double root, double lower, double upper,
                                     double rtol) {
  GroupResult result;
  if (!finite(root) || !(root > 0.)) { result.flags = GROUP_MISSING_ROOT; return result; }
  if (!finite(omega) || !(omega > 0.) || !finite(lower) || !finite(upper)) {
    result.flags = GROUP_NONFINITE;
    return result;
  }
  if (!(lower > 0. && root > lower && root < upper)) { result.flags = GROUP_POLISH_FAILED; return result; }
  Medium m = source;
  m.ka = nullptr; m.kb = nullptr; m.gammk = nullptr; m.refinement = 1;
  double polished = NAN;
  double radius = 2e-6 * root;
  const double limit = std::fmin(.002 * root, .9 * std::fmin(root - lower, upper - root));
  for (int attempt = 0; attempt < 12; ++attempt) {
    const double span = std::fmin(radius, limit), lo = root - span, hi = root + span;
    if (!(span > 0. && lo < root && hi > root)) break;
    const double fl = evaluate(m, omega, lo), fh = evaluate(m, omega, hi);
    result.evaluations += 2;
    if (finite(fl) && finite(fh) && ((fl < 0.) != (fh < 0.) || fl == 0. || fh == 0.)) {
      polished = refine_root(m, omega, lo, hi, fl, fh, 1e-11, result.evaluations);
      break;
    }
    if (span == limit) break;
    radius *= 2.;
  }
  if (!finite(polished) || !(polished > lower && polished < upper)) {
    result.flags = GROUP_POLISH_FAILED;
    return result;
  }
  double hc = std::fmin(1e-3, .2 * std::fmin(std::log1p((polished - lower) / lower),
                                             std::log1p((upper - polished) / polished)));
  double hw = 1e-3;
  if (!(hc > 64. * DBL_EPSILON)) { result.flags = GROUP_DERIVATIVE_UNSTABLE; return result; }
  double cp = group_shape(m, omega, polished * std::exp(hc), result.evaluations);
  double cm = group_shape(m, omega, polished * std::exp(-hc), result.evaluations);
  double wp = group_shape(m, omega * std::exp(hw), polished, result.evaluations);
  double wm = group_shape(m, omega * std::exp(-hw), polished, result.evaluations);
  double previous = NAN;
  unsigned failure = GROUP_DERIVATIVE_UNSTABLE;
  for (int iteration = 0; iteration < 6; ++iteration) {
    const double cp2 = group_shape(m, omega, polished * std::exp(.5 * hc), result.evaluations);
    const double cm2 = group_shape(m, omega, polished * std::exp(-.5 * hc), result.evaluations);
    const double wp2 = group_shape(m, omega * std::exp(.5 * hw), polished, result.evaluations);
    const double wm2 = group_shape(m, omega * std::exp(-.5 * hw), polished, result.evaluations);
    if (!finite(cp) || !finite(cm) || !finite(wp) || !finite(wm) || !finite(cp2) || !finite(cm2) ||
        !finite(wp2) || !finite(wm2)) {
      result.flags = GROUP_NONFINITE;
      return result;
    }
    const double ac = (cp - cm) / (2. * hc), bc = (wp - wm) / (2. * hw);
    const double ah = (cp2 - cm2) / hc, bh = (wp2 - wm2) / hw;
    const double a = (4. * ah - ac) / 3., b = (4. * bh - bc) / 3., den = a + b;
    if (finite(a) && finite(b) && std::fabs(den) > 64. * DBL_EPSILON * (std::fabs(a) + std::fabs(b))) {
      const double u = polished * (a / den);
      const double ea = std::fabs(a - ah) + 16. * DBL_EPSILON * (std::fabs(cp2) + std::fabs(cm2)) / hc;
      const double eb = std::fabs(b - bh) + 16. * DBL_EPSILON * (std::fabs(wp2) + std::fabs(wm2)) / hw;
      const double error = polished * ((std::fabs(b) * ea + std::fabs(a) * eb) / std::fabs(den)) / std::fabs(den);
      const double target = .25 * rtol * std::fmax(polished, std::fabs(u));
      if (finite(u) && finite(error) && finite(previous) && error <= target && std::fabs(u - previous) <= target) {
        result.velocity = u;
        result.slope = a != 0. ? polished * (b / a) : NAN;
        return result;
      }
      previous = u;
      failure = GROUP_DERIVATIVE_UNSTABLE;
    } else {
      failure = GROUP_ILL_CONDITIONED;
      previous = NAN;
    }
    cp = cp2; cm = cm2; wp = wp2; wm = wm2;
    hc *= .5; hw *= .5;
    if (!(hc > 64. * DBL_EPSILON)) break;
  }
  result.flags = failure;
  return result;
}

// ---------------------------------------------------------------------------
// One-direction dual numbers and generic (templated) secular functions.
// ---------------------------------------------------------------------------
struct Dual {
  double v = 0., d = 0.;
  Dual() = default;
  Dual(double value) : v(value), d(0.) {}
  Dual(double value, double deriv) : v(value), d(deriv) {}
};
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
inline bool operator==(Dual a, Dual b) { return a.v == b.v; }
inline double val(Dual a) { return a.v; }
inline double val(double a) { return a; }
inline Dual dsqrt(Dual a) { const double s = std::sqrt(a.v); return {s, a.d / (2. * s)}; }
inline Dual dfabs(Dual a) { return a.v < 0. ? -a : a; }
inline Dual dsin(Dual a) { return {std::sin(a.v), std::cos(a.v) * a.d}; }
inline Dual dcos(Dual a) { return {std::cos(a.v), -std::sin(a.v) * a.d}; }
inline Dual dexp(Dual a) { const double e = std::exp(a.v); return {e, e * a.d}; }
inline Dual scaled(Dual a, double s) { return {a.v / s, a.d / s}; }  // scale held constant
inline double dsqrt(double a) { return std::sqrt(a); }
inline double dfabs(double a) { return std::fabs(a); }
inline double dsin(double a) { return std::sin(a); }
inline double dcos(double a) { return std::cos(a); }
inline double dexp(double a) { return std::exp(a); }
inline double scaled(double a, double s) { return a / s; }

// One wave type's layer factors for derivatives at a root: c = cos p,
// w = sin p / r, x = -r sin p when propagating (k < kx), and e^-p (cosh p,
// sinh p / r, r sinh p) when evanescent, p = r h; pex = p when evanescent
// (else 0) feeds the layer attenuation e^-(pex + sex). Two choices make the
// derivatives exact at a root and well conditioned near r = 0:
//  * the attenuation e^-p is held constant: it multiplies every entry of the
//    layer's matrix, a positive factor of F, which vanishes at the root;
//  * for |z| < 1/4, z = (kx^2 - k^2) h^2, c, w / h and x / h are the power
//    series of cos sqrt z and sin sqrt z / sqrt z in z (cosh and sinh for
//    z < 0) to the z^7 term, so no 1/r appears.
// The values are those of the closed forms to rounding.
inline void wave_parts(Dual k, Dual kx, Dual h, Dual &c, Dual &w, Dual &x, double &pex) {
  static const double inv_fact[16] = {1., 1., 1. / 2., 1. / 6., 1. / 24., 1. / 120., 1. / 720., 1. / 5040.,
                                      1. / 40320., 1. / 362880., 1. / 3628800., 1. / 39916800., 1. / 479001600.,
                                      1. / 6227020800., 1. / 87178291200., 1. / 1307674368000.};
  pex = 0.;
  const Dual r2s = (kx + k) * (kx - k), z = r2s * h * h;
  if (std::fabs(z.v) < .25) {
    const Dual u = -z;
    Dual cs(inv_fact[14]), ss(inv_fact[15]);
    for (int n = 6; n >= 0; --n) {
      cs = cs * u + Dual(inv_fact[2 * n]);
      ss = ss * u + Dual(inv_fact[2 * n + 1]);
    }
    double e = 1.;
    if (k.v > kx.v) { pex = std::sqrt(-z.v); e = std::exp(-pex); }
    c = cs * Dual(e);
    w = h * ss * Dual(e);
    x = -(r2s * h * ss) * Dual(e);
    return;
  }
  if (k.v < kx.v) {
    const Dual r = dsqrt(r2s), p = r * h, sn = dsin(p);
    c = dcos(p);
    w = sn / r;
    x = -r * sn;
    return;
  }
  const Dual r = dsqrt(-r2s), p = r * h;
  pex = p.v;
  const double f = p.v < 16. ? std::exp(-2. * p.v) : 0., cv = (1. + f) * .5, sv = (1. - f) * .5;
  const Dual sh(sv, cv * p.d);  // e^-p sinh p with e^-p held
  c = Dual(cv, sv * p.d);
  w = sh / r;
  x = r * sh;
}

// Released Rayleigh recurrence on a generic scalar (solid or water top). The
// per-layer max-abs normalization uses the values only: the positive scale
// multiplies F and vanishes from any derivative ratio taken at a root.
template <class T>
T rayleigh_generic(const T *d, const T *a, const T *b, const T *rho, int L, T omega, T k) {
  const T k2 = k * k;
  const int last = L - 1;
  const int water = val(b[0]) == 0. ? 1 : 0;
  auto radical = [&](T kx) { return dsqrt((k + kx) * dfabs(k - kx)); };
  T ka0 = omega / a[last], kb0 = omega / b[last];
  T ra = radical(ka0), rb = radical(kb0);
  T t = b[last] / omega;
  T gammk = T(2.) * t * t, gam = gammk * k2, gamm1 = gam - T(1.);
  T rh = rho[last];
  T e[5] = {rh * rh * (gamm1 * gamm1 - gam * gammk * ra * rb), -rh * ra, rh * (gamm1 - gammk * ra * rb), rh * rb,
            k2 - ra * rb};
  for (int j = L - 2; j >= water; --j) {
    const T ka = omega / a[j], kb = omega / b[j];
    t = b[j] / omega;
    gammk = T(2.) * t * t;
    gam = gammk * k2;
    const T h = d[j];
    T cosp, w, x, cosq, y, z;
    double pex, sex;
    wave_parts(k, ka, h, cosp, w, x, pex);
    wave_parts(k, kb, h, cosq, y, z, sex);
    const double exa = pex + sex;
    const T a0(exa == 0. ? 1. : (exa < 60. ? std::exp(-exa) : 0.));  // held, as each factor
    const T cpcq = cosp * cosq, cpy = cosp * y, cpz = cosp * z, cqw = cosq * w, cqx = cosq * x;
    const T xy = x * y, xz = x * z, wy = w * y, wz = w * z;
    const T rhol = rho[j];
    const T g1 = gam - T(1.), tw = gam + g1, gk = gam * gammk, gg1 = gam * g1, g1s = g1 * g1, r2 = rhol * rhol;
    const T apq = a0 - cpcq, tt = T(-2.) * k2;
    T ca[5][5];
    ca[0][0] = cpcq - T(2.) * gg1 * apq - gk * xz - k2 * g1s * wy;
    ca[0][1] = (k2 * cpy - cqx) / rhol;
    ca[0][2] = -(tw * apq + gammk * xz + k2 * g1 * wy) / rhol;
    ca[0][3] = (cpz - k2 * cqw) / rhol;
    ca[0][4] = -(T(2.) * k2 * apq + xz + k2 * k2 * wy) / r2;
    ca[1][0] = (gk * cpz - g1s * cqw) * rhol;
    ca[1][1] = cpcq;
    ca[1][2] = gammk * cpz - g1 * cqw;
    ca[1][3] = -wz;
    ca[1][4] = ca[0][3];
    ca[3][0] = (g1s * cpy - gk * cqx) * rhol;
    ca[3][1] = -xy;
    ca[3][2] = g1 * cpy - gammk * cqx;
    ca[3][3] = cpcq;
    ca[3][4] = ca[0][1];
    ca[4][0] = -(T(2.) * gk * g1s * apq + gk * gk * xz + g1s * g1s * wy) * r2;
    ca[4][1] = ca[3][0];
    ca[4][2] = -(gammk * g1 * tw * apq + gam * gammk * gammk * xz + g1 * g1s * wy) * rhol;
    ca[4][3] = ca[1][0];
    ca[4][4] = ca[0][0];
    ca[2][0] = tt * ca[4][2];
    ca[2][1] = tt * ca[3][2];
    ca[2][2] = a0 + T(2.) * (cpcq - ca[0][0]);
    ca[2][3] = tt * ca[1][2];
    ca[2][4] = tt * ca[0][2];
    T ee[5];
    double scale = 0.;
    for (int i = 0; i < 5; ++i) {
      T s(0.);
      for (int r = 0; r < 5; ++r) s = s + e[r] * ca[r][i];
      ee[i] = s;
      scale = std::fmax(scale, std::fabs(val(s)));
    }
    if (scale < 1e-40) scale = 1.;
    for (int i = 0; i < 5; ++i) e[i] = scaled(ee[i], scale);
  }
  if (water) {
    const T kf = omega / a[0];
    T cp, w, x;
    double pex;
    wave_parts(k, kf, d[0], cp, w, x, pex);
    return cp * e[0] - rho[0] * w * e[1];
  }
  return e[0];
}

template <class T>
T love_generic(const T *h, const T *vs, const T *rho, int L, T omega, T k) {
  const int top = val(vs[0]) == 0. ? 1 : 0, last = L - 1;
  const T kb = omega / vs[last];
  T e1 = rho[last] * dsqrt((k + kb) * dfabs(k - kb));
  T e2 = T(1.) / (vs[last] * vs[last]);
  for (int j = last - 1; j >= top; --j) {
    const T mu = rho[j] * vs[j] * vs[j], kj = omega / vs[j];
    T cq, y, z;
    double sex;
    wave_parts(k, kj, h[j], cq, y, z, sex);
    const T u = e1 * cq + e2 * mu * z, v = e1 * y / mu + e2 * cq;
    double s = std::fmax(std::fabs(val(u)), std::fabs(val(v)));
    if (s < 1e-40) s = 1.;
    e1 = scaled(u, s);
    e2 = scaled(v, s);
  }
  return e1;
}

// Derivatives of F(k, omega; model) at one point. `which` selects the seeded
// direction: 0 -> d/dk, 1 -> d/domega, 2+4*j+{0,1,2,3} -> d/d{h,vp,vs,rho}[j].
inline Dual secular_direction(const double *h, const double *vp, const double *vs, const double *rho, int L,
                              int wave, double omega, double c, int which) {
  std::vector<Dual> dh(L), dvp(L), dvs(L), drho(L);
  for (int j = 0; j < L; ++j) { dh[j] = Dual(h[j]); dvp[j] = Dual(vp[j]); dvs[j] = Dual(vs[j]); drho[j] = Dual(rho[j]); }
  Dual k(omega / c), w(omega);
  if (which == 0) k.d = 1.;
  else if (which == 1) w.d = 1.;
  else {
    const int j = (which - 2) / 4, p = (which - 2) % 4;
    (p == 0 ? dh : p == 1 ? dvp : p == 2 ? dvs : drho)[j].d = 1.;
  }
  return wave == 1 ? love_generic(dh.data(), dvs.data(), drho.data(), L, w, k)
                   : rayleigh_generic(dh.data(), dvp.data(), dvs.data(), drho.data(), L, w, k);
}

// Analytic implicit group velocity U = d(omega)/dk = -F_k / F_omega at a root.
// At a root c from bracket refinement (accurate to the refinement tolerance,
// ~1e-13 relative) the group velocity is evaluated after one Newton step in k
// on F: near a body-wave speed U is steep enough in c that the tolerance alone
// would move it by 1e-8; the step lands on the root to rounding. A step larger
// than 1e-10 relative (c not at a root) is not taken.
inline double group_velocity_analytic(const double *h, const double *vp, const double *vs, const double *rho, int L,
                                      int wave, double omega, double c) {
  Dual fk = secular_direction(h, vp, vs, rho, L, wave, omega, c, 0);
  const double k = omega / c, kn = k - fk.v / fk.d, cn = omega / kn;
  if (std::isfinite(cn) && std::fabs(cn - c) <= 1e-10 * c && cn != c) {
    c = cn;
    fk = secular_direction(h, vp, vs, rho, L, wave, omega, c, 0);
  }
  const Dual fw = secular_direction(h, vp, vs, rho, L, wave, omega, c, 1);
  return -fk.d / fw.d;
}

// The analytic group velocity with the reference's flag contract
// (cpp/common/group_velocity.hpp:9-12, 26-104), in its terms
// A = dF/dln c = -k F_k, B = dF/dln omega = k F_k + omega F_w:
// POLISH_FAILED (2) when the root is not strictly inside (lower, upper);
// ILL_CONDITIONED (8) when |A + B| <= 64 eps (|A| + |B|); DERIVATIVE_UNSTABLE
// (4) when the first-order rounding bound c (|B| dA + |A| dB) / (A + B)^2,
// dA = g |A|, dB = g |B|, g = 32 eps (L + 1), exceeds 0.25 rtol max(c, |U|);
// NONFINITE (32). A flagged root has U = NaN. The bound and thresholds are
// the device kernel's, so that the two agree flag for flag.
inline double group_result_analytic(const double *h, const double *vp, const double *vs, const double *rho, int L,
                                    int wave, double omega, double c, double lower, double upper, double rtol,
                                    unsigned &flags) {
  constexpr double eps = 2.220446049250313e-16;
  flags = 0u;
  if (!(lower > 0. && c > lower && c < upper)) { flags = 2u; return NAN; }
  Dual fk = secular_direction(h, vp, vs, rho, L, wave, omega, c, 0);
  const double k0 = omega / c, kn = k0 - fk.v / fk.d, cn = omega / kn;
  if (std::isfinite(cn) && std::fabs(cn - c) <= 1e-10 * c && cn != c) {
    c = cn;
    fk = secular_direction(h, vp, vs, rho, L, wave, omega, c, 0);
  }
  const Dual fw = secular_direction(h, vp, vs, rho, L, wave, omega, c, 1);
  const double k = omega / c, A = -k * fk.d, B = k * fk.d + omega * fw.d, den = A + B;
  if (!std::isfinite(A) || !std::isfinite(B)) { flags = 32u; return NAN; }
  if (!(std::fabs(den) > 64. * eps * (std::fabs(A) + std::fabs(B)))) { flags = 8u; return NAN; }
  const double u = -fk.d / fw.d;
  if (!std::isfinite(u)) { flags = 32u; return NAN; }
  const double g = 32. * eps * double(L + 1);
  const double err = c * (2. * g * std::fabs(A) * std::fabs(B)) / (den * den);
  if (!(err <= .25 * rtol * std::fmax(c, std::fabs(u)))) { flags = 4u; return NAN; }
  return u;
}

// Implicit phase-velocity gradient dc/dm = -(dF/dm)/(dF/dc) at a root, for
// every layer parameter. out[4*j + {0,1,2,3}] = dc/d{h,vp,vs,rho}[j].
inline void phase_gradient(const double *h, const double *vp, const double *vs, const double *rho, int L, int wave,
                           double omega, double c, double *out) {
  const Dual fk = secular_direction(h, vp, vs, rho, L, wave, omega, c, 0);
  const double fc = fk.d * (-omega / (c * c));  // dF/dc at fixed omega
  for (int j = 0; j < L; ++j)
    for (int p = 0; p < 4; ++p) {
      const Dual fp = secular_direction(h, vp, vs, rho, L, wave, omega, c, 2 + 4 * j + p);
      out[4 * j + p] = -fp.d / fc;
    }
}

}  // namespace mdo
