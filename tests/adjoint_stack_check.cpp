// The device adjoint's per-root sweeps (modisp_amd/csrc/adjoint_root.hpp:
// forward sweep in k-duals with a tape, reverse sweep with the hand-derived
// layer derivative, half-space and water-top terms) compiled as host code
// over whole layer stacks, for a build with AddressSanitizer and
// UndefinedBehaviorSanitizer: every row, tape slot and accumulator the
// sweeps touch is bounds-checked here (the tape is sized to exactly the rows
// a sweep may use), and the accumulators start as NaN so that a derivative
// the sweep never writes shows up.
//
// Input (stdin): the number of roots, then per root "wave L top omega c"
// and four lines of L values (h, vp, vs, rho; row L-1 the half-space).
// Output: per root, 4 L values dc/d{h, vp, vs, rho}[j] (parameter-major),
// the weight-1 sums the device kernel accumulates for that root.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../modisp_amd/csrc/secular.hpp"
#include "../modisp_amd/csrc/adjoint_root.hpp"

namespace {
struct Row {  // the device's prepared row (PrepRow, make_prep_row)
  double h, ivp, ivs, g2, rho, irho, irho2, imu;
};
Row make_row(double h, double vp, double vs, double rho) {
  Row r;
  r.h = h;
  r.ivp = 1. / vp;
  r.ivs = vs > 0. ? 1. / vs : 0.;
  r.g2 = 2. * vs * vs;
  r.rho = rho;
  r.irho = 1. / rho;
  r.irho2 = r.irho * r.irho;
  r.imu = vs > 0. ? 1. / (rho * vs * vs) : 0.;
  return r;
}
}  // namespace

int main() {
  int n = 0;
  if (std::scanf("%d", &n) != 1) return 1;
  for (int it = 0; it < n; ++it) {
    int wave, L, top;
    double omega, c;
    if (std::scanf("%d %d %d %lf %lf", &wave, &L, &top, &omega, &c) != 5) return 1;
    std::vector<double> v(4 * static_cast<size_t>(L));
    for (double &x : v)
      if (std::scanf("%lf", &x) != 1) return 1;
    std::vector<Row> rows(L);
    for (int j = 0; j < L; ++j) rows[j] = make_row(v[j], v[L + j], v[2 * L + j], v[3 * L + j]);
    // Tape: rows top..L-2 only (6 slots each); accumulators 4 x L.
    std::vector<double> tape(6 * static_cast<size_t>(L - 1 - top > 0 ? L - 1 - top : 0), NAN);
    std::vector<double> g(4 * static_cast<size_t>(L), 0.);
    std::vector<char> touched(4 * static_cast<size_t>(L), 0);
    auto row = [&](int j) { return rows.at(j); };
    auto add = [&](int p, int j, double x) {
      g.at(static_cast<size_t>(p) * L + j) += x;
      touched.at(static_cast<size_t>(p) * L + j) = 1;
    };
    auto tp = [&](int j, int i) -> double & { return tape.at(static_cast<size_t>(j - top) * 6 + i); };
    if (wave == 0)
      mdr::rayleigh_root_adjoint(row, add, tp, L, top, omega, c, 1.);
    else
      mdr::love_root_adjoint(row, add, tp, L, top, omega, c, 1.);
    for (size_t i = 0; i < g.size(); ++i) std::printf("%.17g ", touched[i] ? g[i] : NAN);
    std::printf("\n");
  }
  return 0;
}
