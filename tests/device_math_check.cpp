// Accuracy of the FP64 phase sine/cosine and the Newton reciprocal square
// root used by the device evaluators (modisp_amd/csrc/secular.hpp), compiled
// for the host with the same source. Prints the worst errors.
#include <cmath>
#include <cstdio>
#include <random>
#include "../modisp_amd/csrc/secular.hpp"

int main() {
  std::mt19937_64 g(7);
  long double es = 0, ec = 0, ey = 0, er = 0;
  for (int i = 0; i < 4000000; ++i) {
    // phases k h of the layers: up to ~1e4 rad, dense near 0 and multiples of pi/2
    const double x = i < 1000000 ? std::ldexp(std::uniform_real_distribution<double>(-1., 1.)(g), -(i % 40))
                   : i < 2000000 ? (i % 20000) * 1.5707963267948966 + std::ldexp(1., -(i % 50)) * ((i & 1) ? 1 : -1)
                                 : std::uniform_real_distribution<double>(-1e4, 1e4)(g);
    double s, c;
    mdr::sincos_phase(x, &s, &c);
    es = std::fmax(es, std::fabs((long double)s - sinl((long double)x)));
    ec = std::fmax(ec, std::fabs((long double)c - cosl((long double)x)));
    const double r2 = std::ldexp(1. + std::uniform_real_distribution<double>(0., 1.)(g), (i % 200) - 100);
    // a 22-bit estimate, worse than the hardware's (24 bits on gfx950)
    const double y0 = (1. / std::sqrt(r2)) * (1. + std::ldexp(std::uniform_real_distribution<double>(-1., 1.)(g), -22));
    const double y = mdr::rsq_newton(r2, y0), r = r2 * y;
    const long double ty = 1.0L / sqrtl((long double)r2), tr = sqrtl((long double)r2);
    ey = std::fmax(ey, std::fabs((y - ty) / ty));
    er = std::fmax(er, std::fabs((r - tr) / tr));
  }
  std::printf("%.3Le %.3Le %.3Le %.3Le\n", es, ec, ey, er);
  return 0;
}
