// Probe driver for the reference checker library oracle/_ref/libmdref.so.
// TEST INFRASTRUCTURE ONLY. Compiled by `make -C oracle ref` against the
// reference's own header-only physics where it lies (REF/cpp/common/*.hpp);
// this file adds nothing but C-ABI entry points that call those headers, so
// the oracle restatement can be compared with the reference functions
// themselves. Never shipped, never loaded by modisp_amd.
#include <vector>
#include "common/physics.hpp"
#include "common/group_velocity.hpp"

namespace {
md::Medium medium(const double *h, const double *vp, const double *vs, const double *rho, int L, int wave,
                  std::vector<double> &ivp, std::vector<double> &ivs, int refinement) {
  ivp.resize(L);
  ivs.resize(L);
  for (int j = 0; j < L; ++j) {
    ivp[j] = 1. / (vp[j] * vp[j]);
    ivs[j] = vs[j] == 0. ? 0. : 1. / (vs[j] * vs[j]);
  }
  md::Medium m{h, vp, vs, rho, ivp.data(), ivs.data(), L, wave};
  m.refinement = refinement;
  return m;
}
}  // namespace

extern "C" {

// md::evaluate (released, secular=0) or the fast recurrence (secular=1).
void mdr_secular(const double *h, const double *vp, const double *vs, const double *rho, int L, int wave,
                 int secular, double period, const double *c, int n, double *f, double *g) {
  std::vector<double> ivp, ivs, irho(L), irho2(L);
  md::Medium m = medium(h, vp, vs, rho, L, wave, ivp, ivs, 1);
  for (int j = 0; j < L; ++j) { irho[j] = 1. / rho[j]; irho2[j] = irho[j] * irho[j]; }
  if (secular == 1) { m.irho = irho.data(); m.irho2 = irho2.data(); m.secular = 1; }
  const double omega = md::two_pi<double>() / period;
  for (int j = 0; j < n; ++j) {
    if (g) {
      const auto s = md::shape_evaluate(m, omega, c[j]);
      f[j] = s.f;
      g[j] = s.g;
    } else
      f[j] = md::evaluate(m, omega, c[j]);
  }
}

double mdr_refine(const double *h, const double *vp, const double *vs, const double *rho, int L, int wave,
                  double period, double a, double b, double tol, int method, int *evaluations) {
  std::vector<double> ivp, ivs;
  const md::Medium m = medium(h, vp, vs, rho, L, wave, ivp, ivs, method);
  const double omega = md::two_pi<double>() / period;
  const double fa = md::evaluate(m, omega, a), fb = md::evaluate(m, omega, b);
  int ev = 2;
  const double r = md::refine_root(m, omega, a, b, fa, fb, tol, ev);
  if (evaluations) *evaluations = ev;
  return r;
}

double mdr_love_phase(const double *h, const double *vp, const double *vs, const double *rho, int L, double period,
                      double c) {
  std::vector<double> ivp, ivs;
  const md::Medium m = medium(h, vp, vs, rho, L, 1, ivp, ivs, 1);
  return md::love_phase(m, md::two_pi<double>() / period, c);
}

int mdr_love_phase_roots(const double *h, const double *vp, const double *vs, const double *rho, int L,
                         double period, double low, double high, int K, double tol, double *roots, int *nr) {
  std::vector<double> ivp, ivs;
  const md::Medium m = medium(h, vp, vs, rho, L, 1, ivp, ivs, 1);
  int ev = 0;
  return md::love_phase_roots(m, md::two_pi<double>() / period, low, high, K, tol, roots, *nr, ev) ? 1 : 0;
}

double mdr_group(const double *h, const double *vp, const double *vs, const double *rho, int L, int wave,
                 double period, double root, double lower, double upper, double rtol, unsigned *flags) {
  std::vector<double> ivp, ivs;
  const md::Medium m = medium(h, vp, vs, rho, L, wave, ivp, ivs, 1);
  const auto g = md::group_velocity(m, md::two_pi<double>() / period, root, lower, upper, rtol, .005);
  if (flags) *flags = g.flags;
  return g.velocity;
}

double mdr_automatic_lower(const double *h, const double *vp, const double *vs, const double *rho, int L, int wave) {
  std::vector<double> ivp, ivs;
  const md::Medium m = medium(h, vp, vs, rho, L, wave, ivp, ivs, 1);
  return md::automatic_lower(m);
}

}  // extern "C"
