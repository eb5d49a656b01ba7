// Secular functions of a layered elastic model, evaluated per lane.
//
// Rayleigh: Dunkin delta (compound) matrix propagator, the five-vector of
// second-order minors carried from the half-space to the surface; F = e[0]
// (reference cpp/common/dunkin.hpp:459-566 computes the same function).
// Love: Haskell SH 2x2 propagator; F = e1 (reference physics.hpp:33-53).
//
// GPU formulation:
//  * per-trace layer constants (ka = w/vp, kb = w/vs, gk = 2 (vs/w)^2, rho,
//    1/rho, 1/rho^2, h) are computed once per trace and read from LDS;
//  * the 15 distinct compound entries are contracted directly with the
//    five-vector (fused multiply-adds), never materialising the 5x5 matrix;
//  * an evanescent layer uses one exp(-p) for both cosh-scaling and the
//    attenuation factor a0 = exp(-(p+q)) = exp(-p) exp(-q);
//  * after each layer the five-vector is rescaled by an approximate reciprocal
//    of its largest magnitude, which changes neither the sign nor the zeros
//    of F and costs one hardware reciprocal instead of five divisions.
// The same template serves float (grid scans) and double (brackets,
// refinement); Dual2 (below) reuses it for analytic derivatives, and the
// root adjoint's reverse-mode layer derivative closes the file.
#pragma once
#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MDR_HD __host__ __device__ __forceinline__
#else
#define MDR_HD inline
#endif

namespace mdr {

// Per-trace constants of one layer; 8 slots so a row is 32 B (float) or 64 B (double).
// Rayleigh uses {h, ka, kb, gk, rho, irho, irho2}; Love uses {h, -, kb, -, mu, imu}.
template <class T>
struct Layer {
  T h, ka, kb, gk, rho, irho, irho2, pad;
};

// Math shims: hardware transcendentals for float on the GPU, libm elsewhere.
MDR_HD float m_sqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_sqrtf(x);
#else
  return std::sqrt(x);
#endif
}
MDR_HD double m_sqrt(double x) { return sqrt(x); }
MDR_HD void m_sincos(float x, float *s, float *c) {
#if defined(__HIP_DEVICE_COMPILE__)
  *s = __sinf(x);
  *c = __cosf(x);
#else
  *s = std::sin(x);
  *c = std::cos(x);
#endif
}
// sin and cos in FP64 for the layer phases (|x| = k h, well below 2^20):
// Cody-Waite reduction by pi/2 with a two-term split of pi/2 and FMAs, then
// the Taylor series of sin (to x^17) and cos (to x^16) on [-pi/4, pi/4]
// (truncation below 1e-18) and the quadrant map. Absolute error of a few ulp;
// no large-argument path (the library routine carries one, and its branch).
MDR_HD void sincos_phase(double x, double *s, double *c) {
  const double n = std::rint(x * 0.6366197723675814);       // x / (pi/2)
  double r = std::fma(-n, 1.5707963267948966, x);            // pi/2, leading double
  r = std::fma(-n, 6.123233995736766e-17, r);                // pi/2 - leading double
  const double z = r * r;
  double ps = 2.8114572543455206e-15;                        // 1/17!
  ps = std::fma(ps, z, -7.647163731819816e-13);              // -1/15!
  ps = std::fma(ps, z, 1.6059043836821613e-10);              // 1/13!
  ps = std::fma(ps, z, -2.505210838544172e-08);              // -1/11!
  ps = std::fma(ps, z, 2.7557319223985893e-06);              // 1/9!
  ps = std::fma(ps, z, -1.984126984126984e-04);              // -1/7!
  ps = std::fma(ps, z, 8.333333333333333e-03);               // 1/5!
  ps = std::fma(ps, z, -1.6666666666666666e-01);             // -1/3!
  const double sn = std::fma(r * z, ps, r);
  double pc = 4.779477332387385e-14;                         // 1/16!
  pc = std::fma(pc, z, -1.1470745597729725e-11);             // -1/14!
  pc = std::fma(pc, z, 2.08767569878681e-09);                // 1/12!
  pc = std::fma(pc, z, -2.755731922398589e-07);              // -1/10!
  pc = std::fma(pc, z, 2.48015873015873e-05);                // 1/8!
  pc = std::fma(pc, z, -1.388888888888889e-03);              // -1/6!
  pc = std::fma(pc, z, 4.1666666666666664e-02);              // 1/4!
  pc = std::fma(pc, z, -0.5);                                // -1/2!
  const double cs = std::fma(z, pc, 1.);
  const int q = static_cast<int>(n) & 3;
  const double a = (q & 1) ? cs : sn, b = (q & 1) ? sn : cs;
  *s = (q & 2) ? -a : a;
  *c = ((q + 1) & 2) ? -b : b;
}

MDR_HD void m_sincos(double x, double *s, double *c) {
#if defined(__HIP_DEVICE_COMPILE__)
  sincos_phase(x, s, c);
#else
  *s = std::sin(x);
  *c = std::cos(x);
#endif
}
MDR_HD float m_exp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __expf(x);
#else
  return std::exp(x);
#endif
}
MDR_HD double m_exp(double x) { return exp(x); }
MDR_HD float m_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __frcp_rn(x);
#else
  return 1.f / x;
#endif
}
MDR_HD double m_rcp(double x) { return 1. / x; }
MDR_HD float m_fabs(float x) { return fabsf(x); }
MDR_HD double m_fabs(double x) { return fabs(x); }
MDR_HD float m_fma(float a, float b, float c) { return fmaf(a, b, c); }
MDR_HD double m_fma(double a, double b, double c) { return fma(a, b, c); }

// Approximate reciprocal: the normalization scale only has to be positive.
MDR_HD float fast_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.f / x;
#endif
}
MDR_HD double fast_rcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcp(x);
#else
  return 1. / x;
#endif
}
template <class T> MDR_HD T tiny_scale();
template <> MDR_HD float tiny_scale<float>() { return 1.0e-30f; }
template <> MDR_HD double tiny_scale<double>() { return 1.0e-40; }
#if !defined(__HIPCC__)
template <> MDR_HD long double tiny_scale<long double>() { return 1.0e-40L; }  // host-only checks
#endif

// Max-abs rescale: multiply by ~1/max|e_i| so the state stays O(1). The scale
// is positive, so neither the sign nor the zeros of F change; it is continuous
// in the inputs, so float and double evaluations stay directly comparable.
template <int N, class T>
MDR_HD T rescale(T *e) {
  T m = T(0);
#pragma unroll
  for (int i = 0; i < N; ++i) m = fmax(m, m_fabs(e[i]));
  // A vanishing state is left unscaled (scale 1); one select, not one per entry.
  const bool big = m > tiny_scale<T>();
  const T inv = big ? fast_rcp(m) : T(1);
#pragma unroll
  for (int i = 0; i < N; ++i) e[i] *= inv;
  return big ? m : T(1);
}

// log2 of a (positive) normalization scale, as float. A trapped-mode
// resonance collapses the whole state vector between two close roots; the
// running sum of these logs then has a log|c - c*| cusp that node samples see
// even when the root pair is far narrower than the node spacing.
MDR_HD float scale_log2(float m) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __log2f(m);
#else
  return std::log2(m);
#endif
}
MDR_HD float scale_log2(double m) {
  int ex;
  const double mant = frexp(m, &ex);  // m = mant 2^ex, mant in [0.5, 1)
  return float(ex) + scale_log2(float(mant));
}

// P (or S) factor of one layer: cosine-like c, sine/r-like w, r*sine-like x,
// and the attenuation e^-p (1 when propagating).
template <class T>
struct Wave {
  T c, w, x, att;
};
// r and 1/r from r^2: one hardware reciprocal square root for float; for
// double, that estimate refined by Newton steps (rsq_newton).
MDR_HD void root_pair(float r2, float *r, float *ir) {
#if defined(__HIP_DEVICE_COMPILE__)
  const float q = __builtin_amdgcn_rsqf(r2);
  *ir = q;
  *r = r2 * q;
#else
  *r = std::sqrt(r2);
  *ir = 1.f / *r;
#endif
}
// FP64: 1/sqrt from an estimate by two Newton steps y <- y (3/2 - x y^2 / 2)
// (each doubles the correct bits: the hardware estimate has 24, so the result
// is within 2 ulp), and sqrt = x / sqrt(x).
MDR_HD double rsq_newton(double x, double y) {
  const double h = 0.5 * x;
  y *= std::fma(-h * y, y, 1.5);
  return y * std::fma(-h * y, y, 1.5);
}
MDR_HD void root_pair(double r2, double *r, double *ir) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double y = rsq_newton(r2, __builtin_amdgcn_rsq(r2));
  *ir = y;
  *r = r2 * y;
#else
  *r = sqrt(r2);
  *ir = 1. / *r;
#endif
}
template <class T>
MDR_HD void root_pair(T r2, T *r, T *ir) {
  *r = m_sqrt(r2);
  *ir = m_rcp(*r);
}

template <class T>
MDR_HD Wave<T> wave_factor(T k, T kx, T h) {
  // At k = kx, r^2 = 0 is raised to a tiny positive value: r stays finite,
  // p rounds e^-p to exactly 1, so c = 1, s = 0, x = 0 and the attenuation
  // is 1 -- the r -> 0 limit -- and only w needs its limit h (below).
  T r, ir;
  root_pair(fmax((k + kx) * m_fabs(k - kx), tiny_scale<T>()), &r, &ir);
  const T p = r * h;
  // The branches produce only (cosine-like, sine-like, attenuation); the
  // shared tail forms w and x, so a divergent wave merges three values.
  T c, s, att;
  if (k < kx) {  // propagating: cos p, sin p
    m_sincos(p, &s, &c);
    att = T(1);
  } else {  // evanescent, scaled by e^-p: cosh p, sinh p
    const T e = m_exp(-p);
    const T f = e * e;
    c = (T(1) + f) * T(0.5);
    s = (T(1) - f) * T(0.5);
    att = e;
  }
  Wave<T> o;
  o.c = c;
  o.w = k == kx ? h : s * ir;  // sin p / r or sinh p / r; the r -> 0 limit h at k = kx
  o.x = k < kx ? -r * s : r * s;
  o.att = att;
  return o;
}

// The 15 distinct entries of Dunkin's compound matrix of one layer (scaled by
// the layer's attenuation e^-(p+q)), rows/columns laid out as
// [A E tJ F O; D B tH N F; I G K H J; C M tG B E; Lq C tI D A], t = -2 k^2.
template <class T>
struct Compound {
  T A, B, C, D, E, F, G, H, I, J, K, Lq, M, N, O, t;
};

template <class T>
MDR_HD Compound<T> compound_from_waves(const Wave<T> &P, const Wave<T> &S, T gk, T rho, T irho, T irho2, T k2) {
  const T a0 = P.att * S.att;
  const T cpcq = P.c * S.c, cpy = P.c * S.w, cpz = P.c * S.x, cqw = S.c * P.w, cqx = S.c * P.x;
  const T xy = P.x * S.w, xz = P.x * S.x, wy = P.w * S.w, wz = P.w * S.x;
  const T gam = gk * k2, gamm1 = gam - T(1), twgm1 = gam + gamm1;
  const T gmgmk = gam * gk, gmgm1 = gam * gamm1, gm1sq = gamm1 * gamm1;
  const T a0pq = a0 - cpcq, k4 = k2 * k2;
  Compound<T> q;
  q.A = cpcq - T(2) * gmgm1 * a0pq - gmgmk * xz - k2 * gm1sq * wy;
  q.D = (k2 * cpy - cqx) * irho;
  q.I = -(twgm1 * a0pq + gk * xz + k2 * gamm1 * wy) * irho;
  q.C = (cpz - k2 * cqw) * irho;
  q.Lq = -(T(2) * k2 * a0pq + xz + k4 * wy) * irho2;
  q.E = (gmgmk * cpz - gm1sq * cqw) * rho;
  q.B = cpcq;
  q.G = gk * cpz - gamm1 * cqw;
  q.M = -wz;
  q.N = -xy;
  q.F = (gm1sq * cpy - gmgmk * cqx) * rho;
  q.H = gamm1 * cpy - gk * cqx;
  q.O = -(T(2) * gmgmk * gm1sq * a0pq + gmgmk * gmgmk * xz + gm1sq * gm1sq * wy) * (rho * rho);
  q.J = -(gk * gamm1 * twgm1 * a0pq + gam * gk * gk * xz + gamm1 * gm1sq * wy) * rho;
  q.K = a0 + T(2) * (cpcq - q.A);
  q.t = T(-2) * k2;
  return q;
}
template <class T>
MDR_HD Compound<T> rayleigh_compound(const Layer<T> &ly, T k, T k2) {
  return compound_from_waves(wave_factor(k, ly.ka, ly.h), wave_factor(k, ly.kb, ly.h), ly.gk, ly.rho, ly.irho,
                             ly.irho2, k2);
}

// n = C e (the propagation step, FMA contraction of the 15 entries).
template <class T, class U>
MDR_HD void compound_apply(const Compound<T> &q, const U e[5], T n[5]) {
  const U e0 = e[0], e1 = e[1], e2 = e[2], e3 = e[3], e4 = e[4];
  const T e2t = e2 * q.t;
  n[0] = m_fma(q.A, T(e0), m_fma(q.E, T(e1), m_fma(q.F, T(e3), m_fma(q.O, T(e4), q.J * e2t))));
  n[1] = m_fma(q.D, T(e0), m_fma(q.B, T(e1), m_fma(q.N, T(e3), m_fma(q.F, T(e4), q.H * e2t))));
  n[2] = m_fma(q.I, T(e0), m_fma(q.G, T(e1), m_fma(q.K, T(e2), m_fma(q.H, T(e3), q.J * T(e4)))));
  n[3] = m_fma(q.C, T(e0), m_fma(q.M, T(e1), m_fma(q.B, T(e3), m_fma(q.E, T(e4), q.G * e2t))));
  n[4] = m_fma(q.Lq, T(e0), m_fma(q.C, T(e1), m_fma(q.D, T(e3), m_fma(q.A, T(e4), q.I * e2t))));
}

// l <- l C (row vector times the compound matrix: the adjoint step).
template <class T>
MDR_HD void compound_apply_transpose(const Compound<T> &q, T l[5]) {
  const T l0 = l[0], l1 = l[1], l2 = l[2], l3 = l[3], l4 = l[4];
  l[0] = q.A * l0 + q.D * l1 + q.I * l2 + q.C * l3 + q.Lq * l4;
  l[1] = q.E * l0 + q.B * l1 + q.G * l2 + q.M * l3 + q.C * l4;
  l[2] = q.t * (q.J * l0 + q.H * l1 + q.G * l3 + q.I * l4) + q.K * l2;
  l[3] = q.F * l0 + q.N * l1 + q.H * l2 + q.B * l3 + q.D * l4;
  l[4] = q.O * l0 + q.F * l1 + q.J * l2 + q.E * l3 + q.A * l4;
}

// One Rayleigh layer: e <- C(layer) e, then rescale; returns the scale. The
// same entries as rayleigh_compound, contracted as they are formed so that
// few of them are live at once (the dual-number instantiations would spill).
template <class T>
MDR_HD T rayleigh_layer(const Layer<T> &ly, T k, T k2, T e[5]) {
  const Wave<T> P = wave_factor(k, ly.ka, ly.h);
  const Wave<T> S = wave_factor(k, ly.kb, ly.h);
  const T a0 = P.att * S.att;
  const T cpcq = P.c * S.c, cpy = P.c * S.w, cpz = P.c * S.x, cqw = S.c * P.w, cqx = S.c * P.x;
  const T xy = P.x * S.w, xz = P.x * S.x, wy = P.w * S.w, wz = P.w * S.x;
  const T gk = ly.gk, gam = gk * k2, gamm1 = gam - T(1), twgm1 = gam + gamm1;
  const T gmgmk = gam * gk, gmgm1 = gam * gamm1, gm1sq = gamm1 * gamm1;
  const T rho = ly.rho, irho = ly.irho, irho2 = ly.irho2;
  const T a0pq = a0 - cpcq, k4 = k2 * k2;
  const T A = cpcq - T(2) * gmgm1 * a0pq - gmgmk * xz - k2 * gm1sq * wy;
  const T D = (k2 * cpy - cqx) * irho;
  const T I = -(twgm1 * a0pq + gk * xz + k2 * gamm1 * wy) * irho;
  const T C = (cpz - k2 * cqw) * irho;
  const T Lq = -(T(2) * k2 * a0pq + xz + k4 * wy) * irho2;
  const T E = (gmgmk * cpz - gm1sq * cqw) * rho;
  const T B = cpcq;
  const T G = gk * cpz - gamm1 * cqw;
  const T M = -wz;
  const T N = -xy;
  const T F = (gm1sq * cpy - gmgmk * cqx) * rho;
  const T H = gamm1 * cpy - gk * cqx;
  const T O = -(T(2) * gmgmk * gm1sq * a0pq + gmgmk * gmgmk * xz + gm1sq * gm1sq * wy) * (rho * rho);
  const T J = -(gk * gamm1 * twgm1 * a0pq + gam * gk * gk * xz + gamm1 * gm1sq * wy) * rho;
  const T K = a0 + T(2) * (cpcq - A);
  const T e0 = e[0], e1 = e[1], e2 = e[2], e3 = e[3], e4 = e[4];
  const T e2t = e2 * (T(-2) * k2);
  e[0] = m_fma(A, e0, m_fma(E, e1, m_fma(F, e3, m_fma(O, e4, J * e2t))));
  e[1] = m_fma(D, e0, m_fma(B, e1, m_fma(N, e3, m_fma(F, e4, H * e2t))));
  e[2] = m_fma(I, e0, m_fma(G, e1, m_fma(K, e2, m_fma(H, e3, J * e4))));
  e[3] = m_fma(C, e0, m_fma(M, e1, m_fma(B, e3, m_fma(E, e4, G * e2t))));
  e[4] = m_fma(Lq, e0, m_fma(C, e1, m_fma(D, e3, m_fma(A, e4, I * e2t))));
  return rescale<5>(e);
}

// Half-space five-vector (decaying P and S waves below the stack).
template <class T>
MDR_HD T rayleigh_halfspace(const Layer<T> &hs, T k, T k2, T e[5]) {
  const T ra = m_sqrt((k + hs.ka) * m_fabs(k - hs.ka));
  const T rb = m_sqrt((k + hs.kb) * m_fabs(k - hs.kb));
  const T gk = hs.gk, gam = gk * k2, gamm1 = gam - T(1), rh = hs.rho, rarb = ra * rb;
  e[0] = rh * rh * (gamm1 * gamm1 - gam * gk * rarb);
  e[1] = -rh * ra;
  e[2] = rh * (gamm1 - gk * rarb);
  e[3] = rh * rb;
  e[4] = k2 - rarb;
  return rescale<5>(e);
}

// Pressure-release water top over the solid stack: F = cos(p) e0 - rho_w w e1.
template <class T>
MDR_HD T rayleigh_water_top(const Layer<T> &wl, T k, const T e[5]) {
  const Wave<T> P = wave_factor(k, wl.ka, wl.h);
  const T t0 = P.c * e[0], t1 = wl.rho * P.w * e[1];
  return t0 - t1;
}

// Rayleigh F at wavenumber k for one trace. `lay` holds L rows (row L-1 the
// half-space); `top` is 1 when row 0 is a water layer.
template <class T>
MDR_HD T rayleigh_F(const Layer<T> *lay, int L, int top, T k) {
  const T k2 = k * k;
  T e[5];
  rayleigh_halfspace(lay[L - 1], k, k2, e);
  for (int j = L - 2; j >= top; --j) rayleigh_layer(lay[j], k, k2, e);
  return top ? rayleigh_water_top(lay[0], k, e) : e[0];
}

// As rayleigh_F, also returning the summed log2 of the normalization scales.
template <class T>
MDR_HD T rayleigh_F_scale(const Layer<T> *lay, int L, int top, T k, float *log_scale) {
  const T k2 = k * k;
  T e[5];
  float ls = scale_log2(rayleigh_halfspace(lay[L - 1], k, k2, e));
  for (int j = L - 2; j >= top; --j) ls += scale_log2(rayleigh_layer(lay[j], k, k2, e));
  *log_scale = ls;
  return top ? rayleigh_water_top(lay[0], k, e) : e[0];
}

// Love: one SH layer, state (traction, displacement / mu-scaled). The wave
// factor already carries z = -r sin q (propagating) or r sinh q (evanescent).
template <class T>
MDR_HD T love_layer(const Layer<T> &ly, T k, T e[2]) {
  const Wave<T> S = wave_factor(k, ly.kb, ly.h);
  const T mu = ly.rho, imu = ly.irho;  // Love rows store mu and 1/mu here
  const T u = m_fma(e[0], S.c, e[1] * mu * S.x);
  const T v = m_fma(e[0] * S.w, imu, e[1] * S.c);
  e[0] = u;
  e[1] = v;
  return rescale<2>(e);
}

// The SH layer matrix [[c, mu x], [w / mu, c]] and its two products.
template <class T>
struct ShMatrix {
  T c, mx, wi;
};
template <class T>
MDR_HD ShMatrix<T> love_matrix(const Layer<T> &ly, T k) {
  const Wave<T> S = wave_factor(k, ly.kb, ly.h);
  return ShMatrix<T>{S.c, ly.rho * S.x, S.w * ly.irho};
}
template <class T, class U>
MDR_HD void sh_apply(const ShMatrix<T> &q, const U e[2], T n[2]) {
  n[0] = q.c * T(e[0]) + q.mx * T(e[1]);
  n[1] = q.wi * T(e[0]) + q.c * T(e[1]);
}
template <class T>
MDR_HD void sh_apply_transpose(const ShMatrix<T> &q, T l[2]) {
  const T l0 = l[0], l1 = l[1];
  l[0] = q.c * l0 + q.wi * l1;
  l[1] = q.mx * l0 + q.c * l1;
}

// Love F: half-space state (mu r, 1) -- the reference's (rho r, 1/vs^2) times
// vs^2 > 0 -- propagated up to the first solid row `top`.
template <class T>
MDR_HD T love_F(const Layer<T> *lay, int L, int top, T k) {
  const Layer<T> &hs = lay[L - 1];
  T e[2];
  e[0] = hs.rho * m_sqrt((k + hs.kb) * m_fabs(k - hs.kb));
  e[1] = T(1);
  rescale<2>(e);
  for (int j = L - 2; j >= top; --j) love_layer(lay[j], k, e);
  return e[0];
}

template <class T>
MDR_HD T love_F_scale(const Layer<T> *lay, int L, int top, T k, float *log_scale) {
  const Layer<T> &hs = lay[L - 1];
  T e[2];
  e[0] = hs.rho * m_sqrt((k + hs.kb) * m_fabs(k - hs.kb));
  e[1] = T(1);
  float ls = scale_log2(rescale<2>(e));
  for (int j = L - 2; j >= top; --j) ls += scale_log2(love_layer(lay[j], k, e));
  *log_scale = ls;
  return e[0];
}

}  // namespace mdr

// ---------------------------------------------------------------------------
// Two-direction forward-mode dual numbers: value plus derivatives along two
// seeded inputs (here the wavenumber k and the angular frequency w). The
// secular templates above instantiate with T = Dual2 unchanged; the
// normalization scale is held constant (it multiplies F, which vanishes at a
// root, so it cancels from any derivative ratio taken there).
// ---------------------------------------------------------------------------
namespace mdr {
struct Dual2 {
  double v, a, b;
  MDR_HD Dual2() : v(0.), a(0.), b(0.) {}
  MDR_HD Dual2(double x) : v(x), a(0.), b(0.) {}
  MDR_HD Dual2(double x, double da, double db) : v(x), a(da), b(db) {}
};
MDR_HD Dual2 operator+(Dual2 x, Dual2 y) { return {x.v + y.v, x.a + y.a, x.b + y.b}; }
MDR_HD Dual2 operator-(Dual2 x, Dual2 y) { return {x.v - y.v, x.a - y.a, x.b - y.b}; }
MDR_HD Dual2 operator-(Dual2 x) { return {-x.v, -x.a, -x.b}; }
MDR_HD Dual2 operator*(Dual2 x, Dual2 y) { return {x.v * y.v, x.a * y.v + x.v * y.a, x.b * y.v + x.v * y.b}; }
MDR_HD Dual2 operator*(Dual2 x, double s) { return {x.v * s, x.a * s, x.b * s}; }
MDR_HD Dual2 &operator*=(Dual2 &x, double s) { x = x * s; return x; }
MDR_HD bool operator<(Dual2 x, Dual2 y) { return x.v < y.v; }
MDR_HD bool operator==(Dual2 x, Dual2 y) { return x.v == y.v; }
MDR_HD Dual2 m_sqrt(Dual2 x) {
  const double s = sqrt(x.v), d = 0.5 / s;
  return {s, x.a * d, x.b * d};
}
MDR_HD void m_sincos(Dual2 x, Dual2 *s, Dual2 *c) {
  double sn, cs;
  m_sincos(x.v, &sn, &cs);
  *s = Dual2(sn, cs * x.a, cs * x.b);
  *c = Dual2(cs, -sn * x.a, -sn * x.b);
}
MDR_HD Dual2 m_exp(Dual2 x) {
  const double e = exp(x.v);
  return {e, e * x.a, e * x.b};
}
MDR_HD Dual2 m_rcp(Dual2 x) {
  const double r = 1. / x.v, d = -r * r;
  return {r, x.a * d, x.b * d};
}
MDR_HD Dual2 m_fabs(Dual2 x) { return x.v < 0. ? -x : x; }
MDR_HD Dual2 m_fma(Dual2 x, Dual2 y, Dual2 z) { return x * y + z; }
template <int N>
MDR_HD Dual2 rescale(Dual2 *e) {
  double m = 0.;
#pragma unroll
  for (int i = 0; i < N; ++i) m = fmax(m, fabs(e[i].v));
  if (m > 1.0e-40) {
    const double inv = 1. / m;
#pragma unroll
    for (int i = 0; i < N; ++i) e[i] *= inv;
    return Dual2(m);
  }
  return Dual2(1.);
}
}  // namespace mdr

// ---------------------------------------------------------------------------
// Wave factor for derivatives at a root (dual numbers). The values are those
// of wave_factor; the derivatives differ in two ways that matter only near a
// body-wave speed (r -> 0), both exact at a root:
//  * the attenuation e^-p of an evanescent layer is held constant: it
//    multiplies every entry of the layer's matrix, so it is a positive factor
//    of F, which vanishes at the root; its derivative grows like 1/r and the
//    terms it adds would cancel only to rounding;
//  * for |z| < 1/4, z = (kx^2 - k^2) h^2 (p^2 propagating, -p^2 evanescent),
//    c, w / h and x / h come from the series C(z) = sum (-z)^n / (2n)! (cos
//    sqrt z, cosh sqrt -z) and S(z) = sum (-z)^n / (2n+1)! (sin sqrt z / sqrt
//    z, sinh sqrt -z / sqrt -z) to n = 7 (truncation below 1e-18), so no 1/r
//    appears; the closed forms lose about eps / p^2 of the derivative of w.
// Narrow: the series only for |z| < 1e-4 (to n = 3, truncation below 1e-20),
// and above it e^-p sinh p formed from m = expm1(-p) as -m (2 + m) / 2,
// which keeps full relative precision for small p (1 - e^-2p loses eps / p
// of it): fewer lanes of a wave straddle the two branches.
// ---------------------------------------------------------------------------
namespace mdr {

constexpr double kInvFact[16] = {1., 1., 0.5, 0.16666666666666666, 0.041666666666666664, 0.008333333333333333,
                                 0.001388888888888889, 0.0001984126984126984, 2.48015873015873e-05,
                                 2.7557319223985893e-06, 2.755731922398589e-07, 2.505210838544172e-08,
                                 2.08767569878681e-09, 1.6059043836821613e-10, 1.1470745597729725e-11,
                                 7.647163731819816e-13};  // 1 / n!

template <class D, bool Narrow = false>
MDR_HD Wave<D> wave_factor_held(D k, D kx, D h) {
  constexpr double kZmax = Narrow ? 1e-4 : 0.25;
  constexpr int kTop = Narrow ? 3 : 7;  // series to the z^kTop term
  Wave<D> o;
  const D r2s = (kx + k) * (kx - k);  // signed r^2: > 0 propagating
  const D z = r2s * h * h;
  if (std::fabs(z.v) < kZmax) {
    const D u = -z;
    D cs(kInvFact[2 * kTop]), ss(kInvFact[2 * kTop + 1]);
    for (int n = kTop - 1; n >= 0; --n) {
      cs = cs * u + D(kInvFact[2 * n]);
      ss = ss * u + D(kInvFact[2 * n + 1]);
    }
    const double e = k.v > kx.v ? std::exp(-std::sqrt(-z.v)) : 1.;  // e^-p, held
    o.c = cs * e;
    o.w = h * ss * e;
    o.x = -(r2s * h * ss) * e;
    o.att = D(e);
  } else if (k.v < kx.v) {
    D r, ir;
    root_pair(r2s, &r, &ir);
    const D p = r * h;
    D sn, cn;
    m_sincos(p, &sn, &cn);
    o.c = cn;
    o.w = sn * ir;
    o.x = -(r * sn);
    o.att = D(1.);
  } else {
    D r, ir;
    root_pair(-r2s, &r, &ir);
    const D p = r * h;
    double e, cv, sv;
    if (Narrow) {
      const double m = std::expm1(-p.v);
      e = 1. + m;
      sv = -0.5 * m * (2. + m);
      cv = 1. - sv;
    } else {
      e = std::exp(-p.v);
      const double f = e * e;
      cv = (1. + f) * 0.5;
      sv = (1. - f) * 0.5;
    }
    const D dp = p - D(p.v);  // the derivative part of p
    const D s = D(sv) + dp * cv;
    o.c = D(cv) + dp * sv;
    o.w = s * ir;
    o.x = r * s;
    o.att = D(e);
  }
  return o;
}
MDR_HD Wave<Dual2> wave_factor(Dual2 k, Dual2 kx, Dual2 h) { return wave_factor_held(k, kx, h); }

// One-direction forward-mode dual numbers (value and one derivative): the
// adjoint's forward sweep in k, and the gk-coefficients of its reverse step.
struct Dual1 {
  double v, a;
  MDR_HD Dual1() : v(0.), a(0.) {}
  MDR_HD Dual1(double x) : v(x), a(0.) {}
  MDR_HD Dual1(double x, double d) : v(x), a(d) {}
};
MDR_HD Dual1 operator+(Dual1 x, Dual1 y) { return {x.v + y.v, x.a + y.a}; }
MDR_HD Dual1 operator-(Dual1 x, Dual1 y) { return {x.v - y.v, x.a - y.a}; }
MDR_HD Dual1 operator-(Dual1 x) { return {-x.v, -x.a}; }
MDR_HD Dual1 operator*(Dual1 x, Dual1 y) { return {x.v * y.v, fma(x.a, y.v, x.v * y.a)}; }
MDR_HD Dual1 operator*(Dual1 x, double s) { return {x.v * s, x.a * s}; }
MDR_HD Dual1 &operator*=(Dual1 &x, double s) { x = x * s; return x; }
MDR_HD bool operator<(Dual1 x, Dual1 y) { return x.v < y.v; }
MDR_HD bool operator==(Dual1 x, Dual1 y) { return x.v == y.v; }
MDR_HD Dual1 m_sqrt(Dual1 x) {
  const double s = sqrt(x.v);
  return {s, x.a * (0.5 / s)};
}
// r = sqrt(r2) and 1/r with their k-derivatives from one reciprocal square
// root: dr = dr2 / (2 r), d(1/r) = -dr2 / (2 r^3).
MDR_HD void root_pair(Dual1 r2, Dual1 *r, Dual1 *ir) {
  double s, is;
  root_pair(r2.v, &s, &is);
  const double h = 0.5 * is * r2.a;
  *r = Dual1(s, h);
  *ir = Dual1(is, -h * is * is);
}
MDR_HD void m_sincos(Dual1 x, Dual1 *s, Dual1 *c) {
  double sn, cs;
  m_sincos(x.v, &sn, &cs);
  *s = Dual1(sn, cs * x.a);
  *c = Dual1(cs, -sn * x.a);
}
MDR_HD Dual1 m_exp(Dual1 x) {
  const double e = exp(x.v);
  return {e, e * x.a};
}
MDR_HD Dual1 m_rcp(Dual1 x) {
  const double r = 1. / x.v;
  return {r, -x.a * r * r};
}
MDR_HD Dual1 m_fabs(Dual1 x) { return x.v < 0. ? -x : x; }
MDR_HD Dual1 m_fma(Dual1 x, Dual1 y, Dual1 z) { return x * y + z; }
MDR_HD Wave<Dual1> wave_factor(Dual1 k, Dual1 kx, Dual1 h) { return wave_factor_held(k, kx, h); }
template <int N>
MDR_HD Dual1 rescale(Dual1 *e) {
  double m = 0.;
#pragma unroll
  for (int i = 0; i < N; ++i) m = fmax(m, fabs(e[i].v));
  if (m > 1.0e-40) {
    const double inv = 1. / m;
#pragma unroll
    for (int i = 0; i < N; ++i) e[i] *= inv;
    return Dual1(m);
  }
  return Dual1(1.);
}


}  // namespace mdr

// ---------------------------------------------------------------------------
// Reverse-mode derivative of one Rayleigh layer for the root adjoint: the
// bilinear form S = l^T C e (C the layer's compound matrix at a fixed k) and
// its derivatives with respect to the layer constants h, ka, kb, gk and rho
// (rho also through 1/rho and 1/rho^2). S is linear in ten products of the
// P and S wave factors (a0 = att_P att_S, cpcq = c_P c_S, ...); their
// coefficients beta depend on gk, rho and the entry weights only, so the gk
// derivative is sum_b beta_b' b (beta in duals of gk), the h, ka, kb
// derivatives go through the adjoints of the wave factors, and the rho
// derivative follows from the entries' homogeneity in rho.
// ---------------------------------------------------------------------------
namespace mdr {

// Wave factor with its derivatives along h and along kx, as wave_factor_held
// computes them: the attenuation held constant (datt = 0) and, for |z| < 1/4,
// the series C(z), S(z) with their z-derivatives.
struct WaveGrad {
  Wave<double> v, dh, dkx;
};
MDR_HD WaveGrad wave_factor_grad(double k, double kx, double h) {
  WaveGrad o;
  const double r2s = (kx + k) * (kx - k), z = r2s * h * h;
  if (std::fabs(z) < 0.25) {
    const double u = -z;
    double cs = kInvFact[14], ss = kInvFact[15], dcs = 0., dss = 0.;
    for (int n = 6; n >= 0; --n) {  // Horner with the derivative in u
      dcs = dcs * u + cs;
      dss = dss * u + ss;
      cs = cs * u + kInvFact[2 * n];
      ss = ss * u + kInvFact[2 * n + 1];
    }
    const double cz = -dcs, sz = -dss;  // dC/dz, dS/dz
    const double e = k > kx ? std::exp(-std::sqrt(-z)) : 1.;
    const double zh = 2. * h * r2s, zk = 2. * kx * h * h;  // dz/dh, dz/dkx
    o.v = {cs * e, h * ss * e, -r2s * h * ss * e, e};
    o.dh = {cz * zh * e, (ss + h * sz * zh) * e, -r2s * (ss + h * sz * zh) * e, 0.};
    o.dkx = {cz * zk * e, h * sz * zk * e, -(2. * kx * h * ss + r2s * h * sz * zk) * e, 0.};
    return o;
  }
  double r, ir;
  root_pair((k + kx) * std::fabs(k - kx), &r, &ir);
  const double p = r * h;
  if (k < kx) {  // p = r h, r^2 = kx^2 - k^2: dp/dh = r, dr/dkx = kx / r
    double sn, cn;
    m_sincos(p, &sn, &cn);
    o.v = {cn, sn * ir, -r * sn, 1.};
    o.dh = {-sn * r, cn, -r * r * cn, 0.};
    const double drk = kx * ir, dpk = h * drk;
    o.dkx = {-sn * dpk, (cn * dpk - sn * ir * drk) * ir, -(drk * sn + r * cn * dpk), 0.};
  } else {  // r^2 = k^2 - kx^2: dr/dkx = -kx / r; c = E cosh p, s = E sinh p with E = e^-p held
    const double e = m_exp(-p), f = e * e, c = (1. + f) * 0.5, s = (1. - f) * 0.5;
    o.v = {c, s * ir, r * s, e};
    o.dh = {s * r, c, r * r * c, 0.};
    const double drk = -kx * ir, dpk = h * drk;
    o.dkx = {s * dpk, (c * dpk - s * ir * drk) * ir, drk * s + r * c * dpk, 0.};
  }
  return o;
}

// Weights of the bilinear form l^T C e over the distinct entries of the
// compound matrix laid out as [A E tJ F O; D B tH N F; I G K H J;
// C M tG B E; Lq C tI D A] (t = -2 k^2 folded in): l^T C e = sum_X w_X X.
struct CompoundWeights {
  double A, B, C, D, E, F, G, H, I, J, K, Lq, M, N, O;
};
MDR_HD CompoundWeights compound_weights(const double l[5], const double e[5], double t) {
  const double te2 = t * e[2];
  CompoundWeights w;
  w.A = std::fma(l[0], e[0], l[4] * e[4]);
  w.E = std::fma(l[0], e[1], l[3] * e[4]);
  w.J = std::fma(l[0], te2, l[2] * e[4]);
  w.F = std::fma(l[0], e[3], l[1] * e[4]);
  w.O = l[0] * e[4];
  w.D = std::fma(l[1], e[0], l[4] * e[3]);
  w.B = std::fma(l[1], e[1], l[3] * e[3]);
  w.H = std::fma(l[1], te2, l[2] * e[3]);
  w.N = l[1] * e[3];
  w.I = std::fma(l[2], e[0], l[4] * te2);
  w.G = std::fma(l[2], e[1], l[3] * te2);
  w.K = l[2] * e[2];
  w.C = std::fma(l[3], e[0], l[4] * e[1]);
  w.M = l[3] * e[1];
  w.Lq = l[4] * e[0];
  return w;
}

// dS/d(h, ka, kb, gk, rho) of S = l^T C e for the layer constants given;
// l is replaced by l^T C (the adjoint step, unscaled).
struct LayerGrad {
  double h, ka, kb, gk, rho;
};
MDR_HD LayerGrad rayleigh_layer_adjoint(double h, double ka, double kb, double gk, double rho, double irho,
                                        double irho2, double k, double l[5], const double e[5]) {
  const WaveGrad P = wave_factor_grad(k, ka, h), S = wave_factor_grad(k, kb, h);
  const double k2 = k * k, k4 = k2 * k2;
  const CompoundWeights w = compound_weights(l, e, -2. * k2);
  // the ten products
  const double a0 = P.v.att * S.v.att, cpcq = P.v.c * S.v.c, cpy = P.v.c * S.v.w, cpz = P.v.c * S.v.x;
  const double cqw = S.v.c * P.v.w, cqx = S.v.c * P.v.x, xy = P.v.x * S.v.w, xz = P.v.x * S.v.x;
  const double wy = P.v.w * S.v.w, wz = P.v.w * S.v.x;
  // their coefficients, in duals of gk
  const Dual1 g(gk, 1.), gam = g * k2, gm1 = gam - Dual1(1.), tw = gam + gm1, gg = gam * g, gm = gam * gm1;
  const Dual1 g2s = gm1 * gm1;
  const double rr = rho * rho;
  const double ap = w.A - 2. * w.K;  // K = a0 + 2 (cpcq - A)
  const Dual1 Q = -(gm * (2. * ap) + tw * (irho * w.I) + Dual1(irho2 * (2. * k2 * w.Lq)) +
                    gg * g2s * (rr * 2. * w.O) + g * gm1 * tw * (rho * w.J));
  const Dual1 b_a0 = Q + Dual1(w.K);
  const Dual1 b_cpcq = Dual1(ap + 2. * w.K + w.B) - Q;
  const Dual1 b_cpy = Dual1(irho * k2 * w.D) + g2s * (rho * w.F) + gm1 * w.H;
  const Dual1 b_cqx = -(Dual1(irho * w.D) + gg * (rho * w.F) + g * w.H);
  const Dual1 b_cpz = Dual1(irho * w.C) + gg * (rho * w.E) + g * w.G;
  const Dual1 b_cqw = -(Dual1(irho * k2 * w.C) + g2s * (rho * w.E) + gm1 * w.G);
  const Dual1 b_xz = -(gg * ap + g * (irho * w.I) + Dual1(irho2 * w.Lq) + gg * gg * (rr * w.O) + gg * g * (rho * w.J));
  const Dual1 b_wy = -(g2s * (k2 * ap) + gm1 * (irho * k2 * w.I) + Dual1(irho2 * k4 * w.Lq) + g2s * g2s * (rr * w.O) +
                       gm1 * g2s * (rho * w.J));
  const double b_wz = -w.M, b_xy = -w.N;
  LayerGrad d;
  d.gk = b_a0.a * a0 + b_cpcq.a * cpcq + b_cpy.a * cpy + b_cqx.a * cqx + b_cpz.a * cpz + b_cqw.a * cqw +
         b_xz.a * xz + b_wy.a * wy;
  // adjoints of the wave factors (their attenuation is held: no att terms)
  const double pc = b_cpcq.v * S.v.c + b_cpy.v * S.v.w + b_cpz.v * S.v.x;
  const double pw = b_cqw.v * S.v.c + b_wy.v * S.v.w + b_wz * S.v.x;
  const double px = b_cqx.v * S.v.c + b_xy * S.v.w + b_xz.v * S.v.x;
  const double sc = b_cpcq.v * P.v.c + b_cqw.v * P.v.w + b_cqx.v * P.v.x;
  const double sw = b_cpy.v * P.v.c + b_xy * P.v.x + b_wy.v * P.v.w;
  const double sx = b_cpz.v * P.v.c + b_xz.v * P.v.x + b_wz * P.v.w;
  d.ka = pc * P.dkx.c + pw * P.dkx.w + px * P.dkx.x;
  d.kb = sc * S.dkx.c + sw * S.dkx.w + sx * S.dkx.x;
  d.h = pc * P.dh.c + pw * P.dh.w + px * P.dh.x + sc * S.dh.c + sw * S.dh.w + sx * S.dh.x;
  // The entries are homogeneous in rho (degree -1: D, I, C; -2: Lq; +1: E,
  // F, J; +2: O; 0: the rest), so rho dS/drho = sum_X degree_X w_X X.
  const Compound<double> q = compound_from_waves(P.v, S.v, gk, rho, irho, irho2, k2);
  d.rho = irho * (2. * (w.O * q.O - w.Lq * q.Lq) + w.E * q.E + w.F * q.F + w.J * q.J - w.D * q.D - w.I * q.I -
                  w.C * q.C);
  // l <- l^T C from the entries' values
  compound_apply_transpose(q, l);
  return d;
}

}  // namespace mdr
