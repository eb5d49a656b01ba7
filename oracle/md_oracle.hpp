// TEST INFRASTRUCTURE ONLY -- the oracle. Nothing in modisp_amd/ may include,
// link or call this file; only tests/, __graft_entry__.smoke() and bench.py's
// CPU-baseline leg use it, and only as the checker.
//
// CPU FP64 restatement of the reference FP64 CPU backend (YuanYusung/modisp
// 0.5.0) physics layer. Every function cites the reference file:line whose
// semantics it follows. Floating-point operations are issued in the same order
// as the reference so that, compiled with -ffp-contract=off against glibc, the
// results are bit-identical to the reference's released evaluator. That is what
// lets the frozen reference fixtures (tests/golden/accuracy_high_cpu_linux)
// pin this file bit for bit.
//
// Units: km, km/s, g/cm^3, s, rad/s, rad/km.
#pragma once
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <algorithm>
#include <vector>

namespace mdo {

// Reference cpp/common/dunkin.hpp:569-572: one 2*pi constant, the long double
// literal narrowed to double.
inline double two_pi() { return static_cast<double>(6.283185307179586476925286766559L); }
constexpr double kEps = 2.220446049250313e-16;

// One layered model plus the per-model tables the reference prepares
// (cpp/cpu/solver.cpp:63-68, 465-497) and the per-trace constants
// (solver.cpp:90-94). Arrays are views; ownership stays with the caller.
struct Medium {
  const double *h = nullptr, *vp = nullptr, *vs = nullptr, *rho = nullptr;
  const double *ivp = nullptr, *ivs = nullptr;  // 1/vp^2, 1/vs^2 (0 for water)
  int L = 0;
  int wave = 0;  // 0 Rayleigh (P-SV), 1 Love (SH)
  const double *ka = nullptr, *kb = nullptr, *gammk = nullptr;  // per trace
  int refinement = 1;  // 0 legacy, 1 Brent, 2 Chandrupatla
};

inline bool finite(double x) { return std::isfinite(x); }

// ---------------------------------------------------------------------------
// Rayleigh: normalized Dunkin compound-matrix secular function, released
// evaluator (reference dunkin.hpp:82-221 and 358-448, NORMALIZE=true).
// ---------------------------------------------------------------------------

// Eigenfunction products of one layer (reference dunkin.hpp:82-159).
struct LayerProducts {
  double a0, cpcq, cpy, cpz, cqw, cqx, xy, xz, wy, wz;
};

inline LayerProducts layer_products(double p, double q, double ra, double rb,
                                    double k, double ka, double kb, double h) {
  double cosp = 0., w = 0., x = 0., cosq = 0., y = 0., z = 0., pex = 0., sex = 0.;
  if (k < ka) {  // propagating P
    const double s = std::sin(p);
    cosp = std::cos(p);
    w = s / ra;
    x = -ra * s;
  } else if (k == ka) {  // c == vp limit
    cosp = 1.;
    w = h;
    x = 0.;
  } else {  // evanescent P, scaled by exp(-p)
    pex = p;
    const double fac = (p < 16.) ? std::exp(-2. * p) : 0.;
    cosp = (1. + fac) * 0.5;
    const double s = (1. - fac) * 0.5;
    w = s / ra;
    x = ra * s;
  }
  if (k < kb) {
    const double s = std::sin(q);
    cosq = std::cos(q);
    y = s / rb;
    z = -rb * s;
  } else if (k == kb) {
    cosq = 1.;
    y = h;
    z = 0.;
  } else {
    sex = q;
    const double fac = (q < 16.) ? std::exp(-2. * q) : 0.;
    cosq = (1. + fac) * 0.5;
    const double s = (1. - fac) * 0.5;
    y = s / rb;
    z = rb * s;
  }
  LayerProducts o;
  const double exa = pex + sex;
  o.a0 = (exa == 0.) ? 1. : ((exa < 60.) ? std::exp(-exa) : 0.);
  o.cpcq = cosp * cosq;
  o.cpy = cosp * y;
  o.cpz = cosp * z;
  o.cqw = cosq * w;
  o.cqx = cosq * x;
  o.xy = x * y;
  o.xz = x * z;
  o.wy = w * y;
  o.wz = w * z;
  return o;
}

// The 5x5 compound layer matrix, all 25 entries (reference dunkin.hpp:171-221).
inline void compound_matrix(double k2, double gam, double gammk, double rho,
                            const LayerProducts &v, double ca[5][5]) {
  const double gamm1 = gam - 1.;
  const double twgm1 = gam + gamm1;
  const double gmgmk = gam * gammk;
  const double gmgm1 = gam * gamm1;
  const double gm1sq = gamm1 * gamm1;
  const double rho2 = rho * rho;
  const double a0pq = v.a0 - v.cpcq;
  const double t = -2. * k2;
  ca[0][0] = v.cpcq - 2. * gmgm1 * a0pq - gmgmk * v.xz - k2 * gm1sq * v.wy;
  ca[0][1] = (k2 * v.cpy - v.cqx) / rho;
  ca[0][2] = -(twgm1 * a0pq + gammk * v.xz + k2 * gamm1 * v.wy) / rho;
  ca[0][3] = (v.cpz - k2 * v.cqw) / rho;
  ca[0][4] = -(2. * k2 * a0pq + v.xz + k2 * k2 * v.wy) / rho2;
  ca[1][0] = (gmgmk * v.cpz - gm1sq * v.cqw) * rho;
  ca[1][1] = v.cpcq;
  ca[1][2] = gammk * v.cpz - gamm1 * v.cqw;
  ca[1][3] = -v.wz;
  ca[1][4] = ca[0][3];
  ca[3][0] = (gm1sq * v.cpy - gmgmk * v.cqx) * rho;
  ca[3][1] = -v.xy;
  ca[3][2] = gamm1 * v.cpy - gammk * v.cqx;
  ca[3][3] = ca[1][1];
  ca[3][4] = ca[0][1];
  ca[4][0] = -(2. * gmgmk * gm1sq * a0pq + gmgmk * gmgmk * v.xz + gm1sq * gm1sq * v.wy) * rho2;
  ca[4][1] = ca[3][0];
  ca[4][2] = -(gammk * gamm1 * twgm1 * a0pq + gam * gammk * gammk * v.xz + gamm1 * gm1sq * v.wy) * rho;
  ca[4][3] = ca[1][0];
  ca[4][4] = ca[0][0];
  ca[2][0] = t * ca[4][2];
  ca[2][1] = t * ca[3][2];
  ca[2][2] = v.a0 + 2. * (v.cpcq - ca[0][0]);
  ca[2][3] = t * ca[1][2];
  ca[2][4] = t * ca[0][2];
}

// Returns F = e[0]; optionally the shape feature g (reference dunkin.hpp:443-447).
inline double rayleigh_secular(const Medium &m, double omega, double c, double *shape = nullptr) {
  const double *d = m.h, *a = m.vp, *b = m.vs, *rho = m.rho;
  const double k = omega / c;
  const double k2 = k * k;
  const int last = m.L - 1;
  const int water = (b[0] == 0.) ? 1 : 0;
  const double ka0 = m.ka ? m.ka[last] : omega / a[last];
  const double kb0 = m.kb ? m.kb[last] : omega / b[last];
  double ra = std::sqrt((k + ka0) * std::fabs(k - ka0));
  double rb = std::sqrt((k + kb0) * std::fabs(k - kb0));
  double t = b[last] / omega;
  double gammk = m.gammk ? m.gammk[last] : 2. * t * t;
  double gam = gammk * k2;
  const double gamm1 = gam - 1.;
  const double rh = rho[last];
  double e[5];
  e[0] = rh * rh * (gamm1 * gamm1 - gam * gammk * ra * rb);
  e[1] = -rh * ra;
  e[2] = rh * (gamm1 - gammk * ra * rb);
  e[3] = rh * rb;
  e[4] = k2 - ra * rb;
  for (int j = m.L - 2; j >= water; --j) {
    const double ka = m.ka ? m.ka[j] : omega / a[j];
    const double kb = m.kb ? m.kb[j] : omega / b[j];
    t = b[j] / omega;
    gammk = m.gammk ? m.gammk[j] : 2. * t * t;
    gam = gammk * k2;
    ra = std::sqrt((k + ka) * std::fabs(k - ka));
    rb = std::sqrt((k + kb) * std::fabs(k - kb));
    const double hj = d[j];
    const LayerProducts v = layer_products(ra * hj, rb * hj, ra, rb, k, ka, kb, hj);
    double ca[5][5];
    compound_matrix(k2, gam, gammk, rho[j], v, ca);
    double ee[5];
    for (int i = 0; i < 5; ++i) {
      double s = 0.;
      for (int r = 0; r < 5; ++r) s += e[r] * ca[r][i];
      ee[i] = s;
    }
    // Max-abs normalization (reference dunkin.hpp:335-348, 416-421).
    double scale = 0.;
    for (int i = 0; i < 5; ++i) {
      const double mag = std::fabs(ee[i]);
      if (mag > scale) scale = mag;
    }
    if (scale < 1.0e-40) scale = 1.;
    for (int i = 0; i < 5; ++i) e[i] = ee[i] / scale;
  }
  if (water) {  // pressure-release acoustic top (reference dunkin.hpp:427-442)
    const double kf = omega / a[0];
    const double rf = std::sqrt((k + kf) * std::fabs(k - kf));
    const double p = rf * d[0];
    double cp, w;
    if (k < kf) {
      const double sn = std::sin(p);
      cp = std::cos(p);
      w = sn / rf;
    } else if (k == kf) {
      cp = 1.;
      w = d[0];
    } else {
      const double z = (p < 16.) ? std::exp(-2. * p) : 0.;
      cp = (1. + z) * .5;
      w = (1. - z) * .5 / rf;
    }
    const double t0 = cp * e[0], t1 = rho[0] * w * e[1];
    const double scale = std::fmax(1., std::fmax(std::fabs(t0), std::fabs(t1)));
    if (shape) {
      const double u = t0 / scale, v = t1 / scale;
      *shape = (u - v) / std::sqrt(u * u + v * v + 1e-300);
    }
    return t0 / scale - t1 / scale;
  }
  if (shape) {
    double other = 1e-300;
    for (int i = 1; i < 5; ++i) other += e[i] * e[i];
    *shape = e[0] / std::sqrt(e[0] * e[0] + other);
  }
  return e[0];
}

// ---------------------------------------------------------------------------
// Love: normalized Haskell SH 2x2 recurrence (reference physics.hpp:33-53).
// ---------------------------------------------------------------------------
inline double love_secular(const Medium &m, double omega, double c, double *shape = nullptr) {
  const double *h = m.h, *vs = m.vs, *rho = m.rho;
  const int top = vs[0] == 0. ? 1 : 0, last = m.L - 1;
  const double k = omega / c, kb = m.kb ? m.kb[last] : omega / vs[last];
  double e1 = rho[last] * std::sqrt((k + kb) * std::fabs(k - kb));
  double e2 = 1. / (vs[last] * vs[last]);
  for (int j = last - 1; j >= top; --j) {
    const double mu = rho[j] * vs[j] * vs[j], kj = m.kb ? m.kb[j] : omega / vs[j];
    const double r = std::sqrt((k + kj) * std::fabs(k - kj)), q = r * h[j];
    double cq, y, z;
    if (k < kj) {
      const double sn = std::sin(q);
      cq = std::cos(q);
      y = sn / r;
      z = -r * sn;
    } else if (k == kj) {
      cq = 1.;
      y = h[j];
      z = 0.;
    } else {
      const double f = q < 16. ? std::exp(-2. * q) : 0.;
      cq = (1. + f) * .5;
      y = (1. - f) * .5 / r;
      z = (1. - f) * .5 * r;
    }
    const double u = e1 * cq + e2 * mu * z, v = e1 * y / mu + e2 * cq;
    double s = std::fmax(std::fabs(u), std::fabs(v));
    if (s < 1e-40) s = 1.;
    e1 = u / s;
    e2 = v / s;
  }
  if (shape) *shape = e1 / std::sqrt(e1 * e1 + e2 * e2 + 1e-300);
  return e1;
}

// Evaluator dispatch (reference physics.hpp:60-82, released recurrence).
inline double evaluate(const Medium &m, double omega, double c) {
  return m.wave == 1 ? love_secular(m, omega, c) : rayleigh_secular(m, omega, c);
}
struct ShapeSample { double f, g; };
inline ShapeSample shape_evaluate(const Medium &m, double omega, double c) {
  ShapeSample s;
  s.f = m.wave == 1 ? love_secular(m, omega, c, &s.g) : rayleigh_secular(m, omega, c, &s.g);
  return s;
}

// ---------------------------------------------------------------------------
// Search-range helpers (reference physics.hpp:23-31, 130-208).
// ---------------------------------------------------------------------------
inline bool sh_has_guide(const Medium &m) {
  const int top = m.vs[0] == 0. ? 1 : 0;
  for (int j = top; j < m.L - 1; ++j)
    if (m.h[j] > 0. && m.vs[j] < m.vs[m.L - 1]) return true;
  return false;
}

inline bool ambiguous_sh_cutoff(const Medium &m, double c, double f) {
  return m.wave && f == 0. && std::fabs(c - m.vs[m.L - 1]) <= 64. * kEps * m.vs[m.L - 1];
}

// Mode-density estimate N(c) (reference physics.hpp:130-138).
inline double density_base(const Medium &m, double c) {
  const double ic = 1.0 / (c * c);
  double n = 0.;
  for (int j = 0; j < m.L - 1; ++j) {
    if (m.h[j] == 0.) continue;
    if (m.ivs[j] > ic) n += m.h[j] * std::sqrt(m.ivs[j] - ic);
    if (m.wave == 0 && m.ivp[j] > ic) n += m.h[j] * std::sqrt(m.ivp[j] - ic);
  }
  return 2. * n;
}

// Half-space Rayleigh speed by 56 bisections of the Rayleigh cubic
// x^3 - 8x^2 + (24-16k)x - 16(1-k) (reference physics.hpp:140-144).
inline double rayleigh_reference(double vp, double vs) {
  const double ratio = (vs / vp) * (vs / vp);
  double l = 0., h = 1.;
  for (int j = 0; j < 56; ++j) {
    const double x = (l + h) * .5;
    const double f = ((x - 8.) * x + 24. - 16. * ratio) * x - 16. * (1. - ratio);
    if (f < 0.) l = x; else h = x;
  }
  return vs * std::sqrt((l + h) * .5);
}

// Scholte speed of a fluid over a solid half-space (reference physics.hpp:146-156).
inline double scholte_reference(double vf, double rf, double vp, double vs, double rs) {
  const double ratio = (vs / vp) * (vs / vp), qf = (vs / vf) * (vs / vf);
  double l = 0., h = std::fmax(0., std::fmin(1., vf / vs) * (1. - 1e-12));
  for (int j = 0; j < 64; ++j) {
    const double x = (l + h) * .5, x2 = x * x;
    const double A = std::sqrt(std::fmax(0., 1. - ratio * x2)), B = std::sqrt(std::fmax(0., 1. - x2));
    const double f = 4. * (1. + ratio - ratio * x2) / (1. + A * B) - 4. + x2 +
                     (rf / rs) * x2 * A / std::sqrt(std::fmax(1e-30, 1. - qf * x2));
    if (f < 0.) l = x; else h = x;
  }
  return vs * (l + h) * .5;
}

// Automatic lower search bound: 1/8 of the slowest reference speed
// (reference physics.hpp:158-175).
inline double automatic_lower(const Medium &m) {
  double seed = 1e300;
  const int top = m.vs[0] == 0. ? 1 : 0;
  for (int j = top; j < m.L; ++j) {
    if (m.h[j] == 0. && j < m.L - 1) continue;
    const double vs = m.vs[j], vp = m.vp[j];
    if (!m.wave && !top && vs > 0. && .68 * vs > seed) {
      const double ratio = vs / vp;
      if (ratio > 0. && ratio * ratio <= .75) continue;
    }
    double x = m.wave ? vs : rayleigh_reference(vp, vs);
    if (!m.wave && top) x = std::fmin(x, scholte_reference(m.vp[0], m.rho[0], vp, vs, m.rho[j]));
    seed = std::fmin(seed, x);
  }
  return .125 * seed;
}

// Inverse of N(c) through a monotone table (reference physics.hpp:180-193).
inline double inverse_density(const Medium &m, const double *tx, const double *tn, int nt, double target) {
  int l = 0, h = nt - 1;
  while (h - l > 1) {
    const int q = (l + h) / 2;
    if (tn[q] < target) l = q; else h = q;
  }
  double a = tx[l], b = tx[h], fa = tn[l], fb = tn[h];
  const double eps = 1e-9 * std::fmax(1., std::fabs(target));
  for (int j = 0; j < 48; ++j) {
    double x = a + (b - a) * (target - fa) / (fb - fa);
    if (j % 4 == 3 || !finite(x) || x <= a || x >= b) x = (a + b) * .5;
    const double f = density_base(m, x);
    if (std::fabs(f - target) <= eps || b - a <= 8e-15 * std::fmax(1., std::fabs(x))) return x;
    if (f < target) { a = x; fa = f; } else { b = x; fb = f; }
  }
  return (a + b) * .5;
}

// Dense-grid cell requirement (reference physics.hpp:195-204).
inline double dense_workload(const Medium &m, double period, double lo, double hi) {
  double need = 32. * density_base(m, hi) / period;
  if (!m.wave && m.vs[0] == 0.) {
    const double optical = 2. * m.h[0] / (period * m.vp[0]);
    need = std::fmax(need, 2. * std::fmax(0., hi - lo) * optical * optical / m.vp[0]);
  }
  return need;
}

// Root agreement and ordered witness containment (reference physics.hpp:206-222).
inline bool roots_agree(double current, double previous, double tol, double floor = 0.) {
  return finite(current) && finite(previous) &&
         std::fabs(current - previous) <= std::fmax(4e-12, std::fmax(floor, .25 * tol) * std::fabs(current));
}
inline bool contains_witnesses(const double *roots, int n, const double *witness, int nw, int K, double tol,
                               double floor = 0.) {
  int j = 0;
  for (int i = 0; i < nw; ++i) {
    if (n == K && n > 0 &&
        witness[i] > roots[n - 1] + std::fmax(4e-12, std::fmax(floor, .25 * tol) * std::fabs(roots[n - 1])))
      break;
    while (j < n && !roots_agree(roots[j], witness[i], tol, floor) && roots[j] < witness[i]) ++j;
    if (j >= n || !roots_agree(roots[j], witness[i], tol, floor)) return false;
    ++j;
  }
  return true;
}
inline bool has_close_roots(const double *r, int n, double tol) {  // reference accuracy.hpp:69-72
  for (int i = 1; i < n; ++i)
    if (r[i] - r[i - 1] <= tol * r[i]) return true;
  return false;
}

// ---------------------------------------------------------------------------
// Root refinement (reference refinement.hpp:6-65, physics.hpp:224-237).
// ---------------------------------------------------------------------------
constexpr double kCoordinateFloor = 16. * DBL_EPSILON;  // physics.hpp:128

inline double refinement_target(double a, double b, double tol) {
  return std::fmax(kCoordinateFloor * std::fmax(std::fabs(a), std::fabs(b)),
                   std::fmin((b - a) / 64., .125 * tol * std::fmax(std::fabs(a), std::fabs(b))));
}

inline double refine_legacy(const Medium &m, double omega, double lo, double hi, double fl, double fh,
                            double tol, int &evals) {
  if (lo == hi) return lo;
  const double target = refinement_target(lo, hi, tol);
  for (int j = 0; j < 100 && hi - lo > target; ++j) {
    double x = lo + (hi - lo) * (std::fabs(fl) / (std::fabs(fl) + std::fabs(fh)));
    if (j % 3 == 2 || !finite(x) || x <= lo + .1 * (hi - lo) || x >= hi - .1 * (hi - lo)) x = (lo + hi) * .5;
    const double f = evaluate(m, omega, x);
    ++evals;
    if (!finite(f)) return NAN;
    if (f == 0.) return x;
    if ((f < 0.) == (fl < 0.)) { lo = x; fl = f; } else { hi = x; fh = f; }
  }
  return hi - lo <= target ? (lo + hi) * .5 : NAN;
}

inline bool rfinite(double x) { return x == x && std::fabs(x) != INFINITY; }

inline double refine_root(const Medium &m, double w, double a, double b, double fa, double fb, double tol,
                          int &ev) {
  if (m.refinement == 0) return refine_legacy(m, w, a, b, fa, fb, tol, ev);
  if (a == b) return fa == 0. ? a : NAN;
  if (!rfinite(fa) || !rfinite(fb) || ((fa < 0.) == (fb < 0.) && fa != 0. && fb != 0.)) return NAN;
  if (fa == 0.) return a;
  if (fb == 0.) return b;
  const double stop = refinement_target(a, b, tol);
  if (m.refinement == 2) {  // Chandrupatla (reference refinement.hpp:18-40)
    double x1 = a, x2 = b, x3 = b, f1 = fa, f2 = fb, f3 = fb, t = .5;
    for (int j = 0; j < 160; ++j) {
      if (std::fabs(x2 - x1) <= stop) return .5 * (x1 + x2);
      double x = x1 + t * (x2 - x1);
      if (!(x > std::fmin(x1, x2) && x < std::fmax(x1, x2))) x = .5 * (x1 + x2);
      const double f = evaluate(m, w, x);
      ++ev;
      if (!rfinite(f)) return NAN;
      if (f == 0.) return x;
      x3 = x2; f3 = f2;
      if ((f < 0.) == (f1 < 0.)) { x3 = x1; f3 = f1; } else { x2 = x1; f2 = f1; }
      x1 = x; f1 = f;
      const double dx = std::fabs(x2 - x1);
      if (dx <= stop) return .5 * (x1 + x2);
      const double xi = (x1 - x2) / (x3 - x2), phi = (f1 - f2) / (f3 - f2);
      t = .5;
      if (xi > 0. && xi < 1. && (1. - std::sqrt(1. - xi)) < phi && phi < std::sqrt(xi)) {
        const double alpha = (x3 - x1) / (x2 - x1);
        const double z = f1 / (f1 - f2) * f3 / (f3 - f2) - alpha * f1 / (f3 - f1) * f2 / (f2 - f3);
        if (rfinite(z)) t = z;
      }
      const double limit = std::fmin(.5, .5 * stop / dx);
      t = std::fmax(limit, std::fmin(1. - limit, t));
    }
    return NAN;
  }
  // Brent-Dekker, b the best point, c the enclosing opposite (reference refinement.hpp:41-63).
  double c = b, fc = fb, d = b - a, e = d;
  for (int j = 0; j < 160; ++j) {
    if ((fb < 0.) == (fc < 0.)) { c = a; fc = fa; e = d = b - a; }
    if (std::fabs(fc) < std::fabs(fb)) { a = b; fa = fb; b = c; fb = fc; c = a; fc = fa; }
    if (fb == 0.) return b;
    const double mid = .5 * (c - b), step = .5 * stop;
    if (std::fabs(c - b) <= stop) return .5 * (b + c);
    if (std::fabs(e) >= step && std::fabs(fa) > std::fabs(fb)) {
      double s = fb / fa, p, q;
      if (a == c) { p = 2. * mid * s; q = 1. - s; }
      else {
        q = fa / fc;
        const double r = fb / fc;
        p = s * (2. * mid * q * (q - r) - (b - a) * (r - 1.));
        q = (q - 1.) * (r - 1.) * (s - 1.);
      }
      if (p > 0.) q = -q; else p = -p;
      const double old = e;
      e = d;
      if (rfinite(p / q) && 2. * p < std::fmin(3. * mid * q - std::fabs(step * q), std::fabs(old * q))) d = p / q;
      else { d = mid; e = mid; }
    } else { d = mid; e = mid; }
    a = b; fa = fb;
    b += std::fabs(d) > step ? d : (mid > 0. ? step : -step);
    fb = evaluate(m, w, b);
    ++ev;
    if (!rfinite(fb)) return NAN;
  }
  return NAN;
}

// ---------------------------------------------------------------------------
// Love phase counting (reference love_phase.hpp:5-91).
// ---------------------------------------------------------------------------
inline double phase_unwrap(double raw, double previous) {
  return previous + std::atan2(std::sin(raw - previous), std::cos(raw - previous));
}

inline double love_phase(const Medium &m, double omega, double c) {
  const double *h = m.h, *vs = m.vs, *rho = m.rho;
  const int last = m.L - 1, top = vs[0] == 0. ? 1 : 0;
  const double k = omega / c, khalf = m.kb ? m.kb[last] : omega / vs[last];
  if (!(k > khalf)) return NAN;
  double scale = rho[last] * vs[last] * vs[last] * std::sqrt((k + khalf) * (k - khalf));
  double theta = two_pi() * .125;
  for (int j = last - 1; j >= top; --j) {
    if (h[j] == 0.) continue;
    const double mu = rho[j] * vs[j] * vs[j], kb = m.kb ? m.kb[j] : omega / vs[j];
    const double r = std::sqrt(std::fabs((k + kb) * (k - kb))), next = mu * (r > 0. ? r : 1. / h[j]);
    if (!(scale > 0.) || !(next > 0.) || !rfinite(scale / next)) return NAN;
    theta = phase_unwrap(std::atan2((scale / next) * std::sin(theta), std::cos(theta)), theta);
    scale = next;
    if (k < kb) theta -= r * h[j];
    else {
      const double x = std::sin(theta), y = std::cos(theta);
      double u, t;
      if (r == 0.) { t = x; u = y + x * h[j] * scale / mu; }
      else {
        const double e = std::exp(-2. * r * h[j]), cq = .5 * (1. + e), sn = .5 * (1. - e);
        t = x * cq + y * sn;
        u = y * cq + x * sn;
      }
      if (std::fmax(std::fabs(t), std::fabs(u)) < 1e-14) return NAN;
      theta = phase_unwrap(std::atan2(t, u), theta);
    }
    if (!rfinite(theta) || std::fabs(theta) > 1048576.) return NAN;
  }
  return theta;
}

inline double love_phase_root(const Medium &m, double omega, double low, double high, double pl, double ph,
                              double level, double phase_tol, int &evaluations) {
  double a = low, b = high;
  if (!(pl >= level && ph < level)) return NAN;
  const double stop = refinement_target(a, b, phase_tol);
  for (int j = 0; j < 160 && b - a > stop; ++j) {
    const double mid = a + .5 * (b - a), p = love_phase(m, omega, mid);
    ++evaluations;
    if (!rfinite(p)) return NAN;
    if (p > level) a = mid; else b = mid;
  }
  if (b - a > stop) return NAN;
  const double fa = evaluate(m, omega, a), fb = evaluate(m, omega, b);
  evaluations += 2;
  if (!rfinite(fa) || !rfinite(fb) || (fa != 0. && fb != 0. && ((fa < 0.) == (fb < 0.)))) return NAN;
  const double root = refine_root(m, omega, a, b, fa, fb, phase_tol, evaluations);
  if (!rfinite(root) || root < low || root > high) return NAN;
  return root;
}

inline bool love_phase_roots(const Medium &m, double omega, double low, double high, int K, double tol,
                             double *roots, int &nr, int &evaluations) {
  nr = 0;
  if (!(high > low)) return false;
  if (!sh_has_guide(m)) return true;
  const double pi = two_pi() * .5, pl = love_phase(m, omega, low), ph = love_phase(m, omega, high);
  evaluations += 2;
  if (!rfinite(pl) || !rfinite(ph) || ph > pl || std::fabs(pl) > 1048576. || std::fabs(ph) > 1048576.)
    return false;
  const double phase_tol = std::fmin(tol, 1e-8);
  for (int n = static_cast<int>(std::floor(pl / pi)); n * pi > ph && nr < K; --n) {
    const double root = love_phase_root(m, omega, low, high, pl, ph, n * pi, phase_tol, evaluations);
    if (!rfinite(root) || (nr && root <= roots[nr - 1])) return false;
    roots[nr++] = root;
  }
  return true;
}

}  // namespace mdo
