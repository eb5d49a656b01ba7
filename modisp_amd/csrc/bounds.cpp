// Automatic search window of the modisp-rocm backend (host C++17).
//
// Semantics follow the reference's bounds (cpp/cpu/solver.cpp:781-790):
// lower = 1/8 of the slowest reference surface-wave speed of any layer
// (cpp/common/physics.hpp:158-175): the half-space Rayleigh speed of each
// solid layer from the Rayleigh cubic, a Scholte speed when the model has a
// water top, or the shear speed for Love; upper = min(requested or Vs of the
// half-space, the largest double below Vs of the half-space).
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "../../include/modisp_rocm.h"

namespace {

// Rayleigh speed of a homogeneous half-space: the root x in (0,1) of
// x^3 - 8x^2 + (24 - 16 r)x - 16(1 - r) = 0, r = (vs/vp)^2, c = vs sqrt(x),
// by 56 bisection steps (the cubic is negative at 0 and positive at 1).
double rayleigh_speed(double vp, double vs) {
  const double r = (vs / vp) * (vs / vp);
  double lo = 0., hi = 1.;
  for (int it = 0; it < 56; ++it) {
    const double x = (lo + hi) * .5;
    const double f = ((x - 8.) * x + 24. - 16. * r) * x - 16. * (1. - r);
    (f < 0. ? lo : hi) = x;
  }
  return vs * std::sqrt((lo + hi) * .5);
}

// Scholte speed of a fluid (vf, rf) over a solid half-space (vp, vs, rs):
// the Rayleigh-Scholte equation divided by x^2 = (c/vs)^2, bisected below
// min(vf, vs).
double scholte_speed(double vf, double rf, double vp, double vs, double rs) {
  const double r = (vs / vp) * (vs / vp), q = (vs / vf) * (vs / vf);
  double lo = 0., hi = std::fmax(0., std::fmin(1., vf / vs) * (1. - 1e-12));
  for (int it = 0; it < 64; ++it) {
    const double x = (lo + hi) * .5, x2 = x * x;
    const double A = std::sqrt(std::fmax(0., 1. - r * x2)), B = std::sqrt(std::fmax(0., 1. - x2));
    const double f = 4. * (1. + r - r * x2) / (1. + A * B) - 4. + x2 +
                     (rf / rs) * x2 * A / std::sqrt(std::fmax(1e-30, 1. - q * x2));
    (f < 0. ? lo : hi) = x;
  }
  return vs * (lo + hi) * .5;
}

double automatic_lower(const double *h, const double *vp, const double *vs, const double *rho, int L, int wave) {
  double slowest = 1e300;
  const int top = vs[0] == 0. ? 1 : 0;
  for (int j = top; j < L; ++j) {
    if (h[j] == 0. && j < L - 1) continue;  // zero-thickness padding rows
    // With (vs/vp)^2 <= 3/4 the Rayleigh speed exceeds 0.68 vs, so a dry
    // layer with 0.68 vs above the current minimum cannot lower it.
    if (!wave && !top && vs[j] > 0. && .68 * vs[j] > slowest) {
      const double ratio = vs[j] / vp[j];
      if (ratio > 0. && ratio * ratio <= .75) continue;
    }
    double x = wave ? vs[j] : rayleigh_speed(vp[j], vs[j]);
    if (!wave && top) x = std::fmin(x, scholte_speed(vp[0], rho[0], vp[j], vs[j], rho[j]));
    slowest = std::fmin(slowest, x);
  }
  return .125 * slowest;
}

}  // namespace

extern "C" int md_rocm_bounds(const double *thickness, const double *vp, const double *vs, const double *rho, int B,
                              int L, int wave, int auto_low, double low, int auto_high, double high, double *out) {
  if (!thickness || !vp || !vs || !rho || !out || B < 1 || L < 1) return 2;
  auto window = [&](int b0, int b1) {
    for (int i = b0; i < b1; ++i) {
      const long long p = static_cast<long long>(i) * L;
      out[2 * i] = auto_low ? automatic_lower(thickness + p, vp + p, vs + p, rho + p, L, wave) : low;
      const double cap = std::nextafter(vs[p + L - 1], 0.);
      out[2 * i + 1] = std::fmin(auto_high ? vs[p + L - 1] : high, cap);
    }
  };
  // Models are independent: large batches are split over host threads.
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int nthreads = static_cast<int>(std::min<long long>({16, hw, std::max(1LL, 1LL * B * L / 16384)}));
  if (nthreads <= 1) {
    window(0, B);
    return 0;
  }
  std::vector<std::thread> pool;
  for (int t = 0; t < nthreads; ++t) {
    const int b0 = static_cast<int>(1LL * B * t / nthreads), b1 = static_cast<int>(1LL * B * (t + 1) / nthreads);
    try {
      pool.emplace_back(window, b0, b1);
    } catch (...) {  // a thread that cannot start: its models are done here
      window(b0, b1);
    }
  }
  for (std::thread &th : pool) th.join();
  return 0;
}
