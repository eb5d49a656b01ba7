// TEST INFRASTRUCTURE ONLY -- the oracle (see md_oracle.hpp header).
//
// Restatement of the reference FP64 CPU backend's per-trace search for
// accuracy="high" with search_strategy="qe_adaptive" (cpp/cpu/solver.cpp):
// model preparation (465-497), per-trace constants (90-94), density-guided
// adaptive sampling (254-327), dense verification grids (329-370), bracket
// collection (221-252), the six-level ladder and its guarded second family
// (517-584), acceptance flags (585-601), Love phase counting (544-550), and
// the group-velocity pass (707-730). Bounds follow cpp/cpu/solver.cpp:781-790.
//
// Exposed through a C ABI and loaded with ctypes by oracle/__init__.py.
#include <atomic>
#include <cstring>
#include <thread>
#include <utility>
#include "md_oracle.hpp"
#include "md_oracle_deriv.hpp"

namespace mdo {
namespace {

constexpr int kCap = 2048;  // reference solver.cpp:24
using Point = std::pair<double, double>;

// Trace-flag bits (reference python/modisp/__init__.py:15-32).
enum : unsigned {
  F_CAPACITY = 8, F_NEAR_CUTOFF = 16, F_NONFINITE = 32, F_REFINE_FAILED = 64, F_REJECTED = 128,
  F_SAMPLING = 512, F_CONFLICT = 1024, F_CLOSE = 4096, F_LOVE_FALLBACK = 16384, F_UNCONVERGED = 4
};

struct Model {
  Medium view;
  std::vector<double> ia, ib, tx, tn;
  double seed = NAN, water_seed = NAN;
  bool sh_guide = true;
};

struct Work {
  std::vector<Point> points, extra, merged;
  std::vector<double> roots, previous, coarse, full, guarded, ka, kb, gammk;
  Medium medium;
  unsigned long long evaluations = 0;
};

double eval(const Medium &m, double omega, double c, Work &w) {
  ++w.evaluations;
  return evaluate(m, omega, c);
}

void unique_points(std::vector<Point> &p) {
  std::sort(p.begin(), p.end(), [](const Point &a, const Point &b) { return a.first < b.first; });
  p.erase(std::unique(p.begin(), p.end(), [](const Point &a, const Point &b) { return a.first == b.first; }),
          p.end());
}

// Bracket collection over sorted samples (reference solver.cpp:221-252).
bool collect_roots(const Medium &m, double omega, Work &w, int K, double tol, bool first = true, bool last = true) {
  auto &p = w.points;
  auto &roots = w.roots;
  for (size_t i = 0; i + 1 < p.size(); ++i) {
    const double a = p[i].first, b = p[i + 1].first, fa = p[i].second, fb = p[i + 1].second;
    if (!finite(fa) || !finite(fb)) return false;
    if (fa == 0.) {
      if (fb == 0. || (i > 0 && p[i - 1].second == 0.)) return false;
      if ((i > 0 && std::signbit(p[i - 1].second) != std::signbit(fb)) || (first && i == 0)) {
        if (static_cast<int>(roots.size()) < K) roots.push_back(a);
      } else
        return false;
    } else if (fb != 0. && std::signbit(fa) != std::signbit(fb)) {
      if (static_cast<int>(roots.size()) < K) {
        int ev = 0;
        const double r = refine_root(m, omega, a, b, fa, fb, tol, ev);
        w.evaluations += ev;
        if (!finite(r)) return false;
        roots.push_back(r);
      }
    }
  }
  if (last && p.size() > 1 && p.back().second == 0.) {
    if (p[p.size() - 2].second == 0.) return false;
    if (static_cast<int>(roots.size()) < K) roots.push_back(p.back().first);
  }
  return true;
}

// Density-guided adaptive sampling (reference solver.cpp:254-327).
bool search_adaptive(const Model &model, double period, double lo, double hi, double delta, int K, double tol,
                     Work &w, unsigned &flags, bool shifted) {
  const Medium &m = w.medium;
  const double freq = 1. / period, omega = two_pi() / period;
  const double nlo = freq * density_base(m, lo), nhi = freq * density_base(m, hi), step = delta * .25;
  const double count = std::ceil(std::fmax(0., nhi - nlo) / step);
  w.roots.clear();
  w.points.clear();
  if (!finite(count) || count > kCap - 64 - 6 * m.L - 1) { flags |= F_CAPACITY; return false; }
  auto point = [&](double c) {
    if (c < lo || c > hi || !finite(c)) return;
    w.points.emplace_back(c, eval(m, omega, c, w));
  };
  for (int j = 0; j < 64; ++j) point(lo + (hi - lo) * (shifted && j > 0 && j < 63 ? (j - .5) / 62. : j / 63.));
  for (int j = 0; j < m.L; ++j)
    for (double v : {m.vp[j], m.vs[j]})
      for (int side = -1; side <= 1; ++side) point(v * (1. + side * 1e-7));
  point(model.seed);
  if (!m.wave && m.vs[0] == 0.) point(model.water_seed);
  for (int j = 1; j <= static_cast<int>(count); ++j) {
    const double target = (nlo + (j - (shifted ? .5 : 0.)) * step) / freq;
    if (target < model.tn.back()) {
      const double c = inverse_density(m, model.tx.data(), model.tn.data(), static_cast<int>(model.tx.size()), target);
      if (!finite(c) || std::fabs(freq * (density_base(m, c) - target)) > delta * 1e-3) {
        flags |= F_SAMPLING;
        return false;
      }
      point(c);
    }
  }
  unique_points(w.points);
  if (w.points.size() > kCap) { flags |= F_CAPACITY; return false; }
  for (int round = 0; round < 3; ++round) {  // parabolic extremum proposals
    w.extra.clear();
    for (size_t i = 0; i + 2 < w.points.size(); ++i) {
      const double x1 = w.points[i].first, x2 = w.points[i + 1].first, x3 = w.points[i + 2].first;
      const double f1 = w.points[i].second, f2 = w.points[i + 1].second, f3 = w.points[i + 2].second;
      if (!finite(f1) || !finite(f2) || !finite(f3)) { flags |= F_NONFINITE; return false; }
      const double scale = std::fmax(std::fabs(f1), std::fmax(std::fabs(f2), std::fabs(f3)));
      if (scale == 0.) { flags |= F_NONFINITE; return false; }
      const double s1 = (f2 / scale - f1 / scale) / (x2 - x1), s2 = (f3 / scale - f2 / scale) / (x3 - x2);
      const double a = (s2 - s1) / (x3 - x1);
      double x = NAN;
      if (finite(a) && std::fabs(a) > 64. * kEps / ((x3 - x1) * (x3 - x1))) x = (x1 + x2) * .5 - s1 / (2. * a);
      if (!finite(x) && std::fabs(f2) < .05) x = (x1 + x3) * .5;
      const double gap = 16. * kEps * std::fmax(1., std::fabs(x2));
      if (!(x > x1 + gap && x < x3 - gap) || std::fabs(x - x2) <= gap) continue;
      w.extra.emplace_back(x, eval(m, omega, x, w));
    }
    if (w.extra.empty()) break;
    if (w.points.size() + w.extra.size() > kCap) { flags |= F_CAPACITY; return false; }
    std::sort(w.extra.begin(), w.extra.end());
    w.merged.clear();
    std::merge(w.points.begin(), w.points.end(), w.extra.begin(), w.extra.end(), std::back_inserter(w.merged));
    w.points.swap(w.merged);
    w.points.erase(std::unique(w.points.begin(), w.points.end(),
                               [](const Point &a, const Point &b) { return a.first == b.first; }),
                   w.points.end());
  }
  if (!w.points.empty() && ambiguous_sh_cutoff(m, w.points.back().first, w.points.back().second)) {
    flags |= F_NEAR_CUTOFF;
    return false;
  }
  if (!collect_roots(m, omega, w, K, tol)) { flags |= F_NONFINITE; return false; }
  return true;
}

// Dense uniform verification grid (reference solver.cpp:329-370), tile 2048.
bool search_dense(double period, double lo, double hi, int factor, int K, double tol, Work &w, unsigned &flags) {
  const Medium &m = w.medium;
  const double need = dense_workload(m, period, lo, hi);
  w.roots.clear();
  if (!finite(need) || need > 524288.) { flags |= F_CAPACITY; return false; }
  int cells = 32768;
  while (cells < need) cells *= 2;
  cells *= factor;
  const double omega = two_pi() / period;
  const int tile = 2048;
  for (int start = 0; start < cells;) {
    w.points.clear();
    int n = std::min(tile, cells - start + 1);
    for (int j = 0; j < n; ++j) {
      const int index = start + j;
      const double c = lo + (hi - lo) * index / cells;
      w.points.emplace_back(c, eval(m, omega, c, w));
    }
    if (n < kCap && start + n < cells + 1 && w.points.back().second == 0.) {
      const double c = lo + (hi - lo) * (start + n) / cells;
      w.points.emplace_back(c, eval(m, omega, c, w));
      ++n;
    }
    if (start + n == cells + 1 && ambiguous_sh_cutoff(m, w.points.back().first, w.points.back().second)) {
      flags |= F_NEAR_CUTOFF;
      return false;
    }
    if (!collect_roots(m, omega, w, K, tol, start == 0, start + n == cells + 1)) {
      flags |= F_NONFINITE;
      return false;
    }
    if (static_cast<int>(w.roots.size()) >= K) break;
    start += n - 1;
  }
  return true;
}

struct Config {
  int B, L, nf, modes, wave, automatic, refinement, love_phase, group_mode;
  double tol, accept, close_rtol, group_rtol;
  int period_stride;
};

struct Outputs {
  double *velocity;
  uint32_t *flags;
  int32_t *levels, *counts;
  double *group;
  uint32_t *group_flags;
};

void prepare_model(Model &q, const double *h, const double *vp, const double *vs, const double *rho,
                   const Config &cfg, double lo, double hi) {
  const int L = cfg.L;
  q.ia.assign(L, 0.);
  q.ib.assign(L, 0.);
  q.view = Medium{};
  q.view.h = h; q.view.vp = vp; q.view.vs = vs; q.view.rho = rho;
  q.view.ivp = q.ia.data(); q.view.ivs = q.ib.data();
  q.view.L = L; q.view.wave = cfg.wave; q.view.refinement = cfg.refinement;
  q.sh_guide = !cfg.wave || sh_has_guide(q.view);
  if (hi <= lo || !q.sh_guide) return;
  for (int j = 0; j < L; ++j) {
    q.ia[j] = 1. / (vp[j] * vp[j]);
    q.ib[j] = vs[j] == 0. ? 0. : 1. / (vs[j] * vs[j]);
  }
  const int top = vs[0] == 0. ? 1 : 0;
  q.seed = cfg.wave ? vs[top] : rayleigh_reference(vp[top], vs[top]);
  q.water_seed = (!cfg.wave && top) ? scholte_reference(vp[0], rho[0], vp[1], vs[1], rho[1]) : NAN;
  q.tx.clear();
  q.tn.clear();
  for (int j = 0; j < 129; ++j) q.tx.push_back(lo + (hi - lo) * j / 128.);
  for (int j = 0; j < L; ++j)
    for (double c : {vp[j], vs[j]})
      if (c > lo && c < hi) q.tx.push_back(c);
  std::sort(q.tx.begin(), q.tx.end());
  q.tx.erase(std::unique(q.tx.begin(), q.tx.end()), q.tx.end());
  for (double c : q.tx) q.tn.push_back(density_base(q.view, c));
}

void solve_trace(const Model &model, double lo, double hi, double period, const Config &cfg, Work &w,
                 size_t out_off, size_t trace, const Outputs &out) {
  const int K = cfg.modes + 2, nf = cfg.nf, L = cfg.L;
  const double tol = cfg.tol;
  const bool range_empty = hi <= lo, analytical_empty = hi > lo && !model.sh_guide;
  if (range_empty || analytical_empty) {
    for (int m = 0; m < cfg.modes; ++m) {
      out.velocity[out_off + static_cast<size_t>(m) * nf] = NAN;
      if (out.group) {
        out.group[out_off + static_cast<size_t>(m) * nf] = NAN;
        out.group_flags[out_off + static_cast<size_t>(m) * nf] = GROUP_MISSING_ROOT;
      }
    }
    out.flags[trace] = 0;
    out.levels[trace] = range_empty ? -4 : -3;
    out.counts[trace] = 0;
    return;
  }
  unsigned word = 0;
  int level = -1;
  bool previous_ok = false;
  const double omega = two_pi() / period;
  w.ka.resize(L); w.kb.resize(L); w.gammk.resize(L);
  for (int j = 0; j < L; ++j) {  // reference solver.cpp:90-94
    w.ka[j] = omega / model.view.vp[j];
    w.kb[j] = model.view.vs[j] == 0. ? 0. : omega / model.view.vs[j];
    const double z = model.view.vs[j] / omega;
    w.gammk[j] = 2. * z * z;
  }
  w.medium = model.view;
  w.medium.ka = w.ka.data(); w.medium.kb = w.kb.data(); w.medium.gammk = w.gammk.data();
  bool shifted_companion = false;
  // Six-level ladder: adaptive densities 0..3, dense grids 4..5 (reference solver.cpp:517-542).
  auto search = [&](bool guarded, int begin) -> int {
    previous_ok = false;
    w.previous.clear();
    auto &coarse = w.coarse;
    coarse.clear();
    bool coarse_ok = false;
    for (int lev = begin; lev < 6; ++lev) {
      const double low = cfg.automatic && lev < 2 ? std::fmin((lev == 0 ? 4. : 2.) * lo, (lo + hi) * .5) : lo;
      double adaptive_hi = hi;
      if (guarded && lev > 0 && lev < 4 && previous_ok && static_cast<int>(w.previous.size()) == K &&
          !(word & (8 | 32 | 64 | 512)))
        adaptive_hi = std::fmin(hi, w.previous.back() * (1. + 4. * tol));
      const bool ok = lev < 4 ? search_adaptive(model, period, low, adaptive_hi, .5 / (1 << lev), K, tol, w, word,
                                                guarded && shifted_companion)
                              : search_dense(period, lo, hi, lev == 4 ? 1 : 2, K, tol, w, word);
      bool same = ok && previous_ok && w.roots.size() == w.previous.size();
      for (size_t m = 0; m < w.roots.size() && same; ++m) {
        same = roots_agree(w.roots[m], w.previous[m], tol);
        if (m && !(w.roots[m] > w.roots[m - 1])) same = false;
      }
      if (lev == 0 && ok) { coarse = w.roots; coarse_ok = true; }
      if (lev == 3 && (word & (8 | 32 | 64 | 512))) same = false;
      if (same && (lev == 3 || lev == 5) && coarse_ok &&
          !contains_witnesses(w.roots.data(), static_cast<int>(w.roots.size()), coarse.data(),
                              static_cast<int>(coarse.size()), K, tol)) {
        same = false;
        word |= F_CONFLICT;
      }
      if (same && (lev == 3 || lev == 5)) return lev;
      previous_ok = ok;
      w.previous = w.roots;
    }
    return -1;
  };
  bool phase_ok = false;
  if (cfg.wave && cfg.love_phase) {  // reference solver.cpp:544-550
    w.roots.assign(K, 0.);
    int nr = 0, ev = 0;
    phase_ok = love_phase_roots(w.medium, omega, lo, hi, K, tol, w.roots.data(), nr, ev);
    w.evaluations += ev;
    if (phase_ok) {
      w.roots.resize(nr);
      level = 11;
      if (has_close_roots(w.roots.data(), nr, cfg.close_rtol)) word |= F_CLOSE;
    } else {
      w.roots.clear();
      word |= F_LOVE_FALLBACK;
    }
  }
  if (!phase_ok) level = search(false, 0);
  if (level >= 1 && level <= 3 && !(word & (8 | 32 | 64 | 512))) {  // guarded family (solver.cpp:569-584)
    auto &full = w.full;
    full = w.roots;
    shifted_companion = static_cast<int>(full.size()) < K;
    const int check_level = search(true, 0);
    bool same = check_level >= 0 && full.size() == w.roots.size();
    for (size_t m = 0; m < full.size() && same; ++m) same = roots_agree(full[m], w.roots[m], tol);
    if (same) level = std::max(level, check_level);
    else {
      word |= F_CONFLICT;
      if (check_level < 0) level = -1;
      else {
        auto &guarded = w.guarded;
        guarded = w.roots;
        level = search(false, 4);
        if (level >= 0 && (!contains_witnesses(w.roots.data(), static_cast<int>(w.roots.size()), full.data(),
                                               static_cast<int>(full.size()), K, tol) ||
                           !contains_witnesses(w.roots.data(), static_cast<int>(w.roots.size()), guarded.data(),
                                               static_cast<int>(guarded.size()), K, tol)))
          level = -1;
      }
    }
  }
  const bool ready = level >= 0;
  if (!ready) word |= F_UNCONVERGED;
  bool cutoff = false;
  const int nr = static_cast<int>(w.roots.size());
  for (int m = 0; m < cfg.modes; ++m) {
    double root = ready && m < nr ? w.roots[m] : NAN;
    if (finite(root) && root >= .99 * model.view.vs[L - 1]) cutoff = true;
    if (finite(root) && cfg.accept < 1.) {
      const double value = eval(w.medium, omega, root, w);
      if (!finite(value) || std::fabs(value) > cfg.accept) { root = NAN; word |= F_REJECTED; }
    }
    const size_t cell = out_off + static_cast<size_t>(m) * nf;
    out.velocity[cell] = root;
    if (out.group) {  // group pass (reference solver.cpp:707-730)
      if (!finite(root)) { out.group[cell] = NAN; out.group_flags[cell] = GROUP_MISSING_ROOT; continue; }
      double glo, ghi;
      unsigned gflag;
      group_window(w.roots.data(), nr, m, lo, hi, cfg.close_rtol, glo, ghi, gflag);
      if (gflag) { out.group[cell] = NAN; out.group_flags[cell] = gflag; continue; }
      if (cfg.group_mode == 1) {
        const GroupResult g = group_velocity_fd(model.view, omega, root, glo, ghi, cfg.group_rtol);
        w.evaluations += g.evaluations;
        out.group[cell] = g.velocity;
        out.group_flags[cell] = g.flags;
      } else {
        unsigned gf = 0;
        out.group[cell] = group_result_analytic(model.view.h, model.view.vp, model.view.vs, model.view.rho, L,
                                                cfg.wave, omega, root, glo, ghi, cfg.group_rtol, gf);
        out.group_flags[cell] = gf;
      }
    }
  }
  if (cutoff) word |= F_NEAR_CUTOFF;
  out.flags[trace] = word;
  out.levels[trace] = level;
  out.counts[trace] = ready ? nr : -1;
}

template <class F>
void parallel_for(size_t n, int threads, F &&fn) {
  threads = std::max(1, std::min<int>(threads, static_cast<int>(std::min<size_t>(n, 256))));
  std::atomic<size_t> next{0};
  auto body = [&](int id) {
    for (size_t i; (i = next.fetch_add(1)) < n;) fn(i, id);
  };
  if (threads == 1) { body(0); return; }
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) pool.emplace_back(body, t);
  for (auto &t : pool) t.join();
}

Medium plain_medium(const double *h, const double *vp, const double *vs, const double *rho, int L, int wave) {
  Medium m;
  m.h = h; m.vp = vp; m.vs = vs; m.rho = rho; m.L = L; m.wave = wave;
  return m;
}

}  // namespace
}  // namespace mdo

using namespace mdo;

extern "C" {

// Search bounds [B,2] (reference cpp/cpu/solver.cpp:781-790).
void mdo_bounds(const double *h, const double *vp, const double *vs, const double *rho, int B, int L, int wave,
                int auto_low, double low, int auto_high, double high, double *out) {
  for (int i = 0; i < B; ++i) {
    const size_t p = static_cast<size_t>(i) * L;
    const Medium v = plain_medium(h + p, vp + p, vs + p, rho + p, L, wave);
    out[2 * i] = auto_low ? automatic_lower(v) : low;
    out[2 * i + 1] = std::fmin(auto_high ? vs[p + L - 1] : high, std::nextafter(vs[p + L - 1], 0.));
  }
}

// Released secular function (Rayleigh or Love) at n velocities; optional shape feature.
void mdo_secular(const double *h, const double *vp, const double *vs, const double *rho, int L, int wave,
                 double period, const double *c, int n, double *f, double *g) {
  const Medium m = plain_medium(h, vp, vs, rho, L, wave);
  const double omega = two_pi() / period;
  for (int j = 0; j < n; ++j) {
    double shape = NAN;
    f[j] = wave == 1 ? love_secular(m, omega, c[j], g ? &shape : nullptr)
                     : rayleigh_secular(m, omega, c[j], g ? &shape : nullptr);
    if (g) g[j] = shape;
  }
}

double mdo_love_phase(const double *h, const double *vp, const double *vs, const double *rho, int L, double period,
                      double c) {
  const Medium m = plain_medium(h, vp, vs, rho, L, 1);
  return love_phase(m, two_pi() / period, c);
}

// The Love roots by the unwrapped phase (reference love_phase.hpp:76-91):
// returns 1 on success (nr roots), 0 when the reference would fall back.
int mdo_love_phase_roots(const double *h, const double *vp, const double *vs, const double *rho, int L,
                         double period, double low, double high, int K, double tol, double *roots, int *nr) {
  const Medium m = plain_medium(h, vp, vs, rho, L, 1);
  int n = 0, ev = 0;
  const bool ok = love_phase_roots(m, two_pi() / period, low, high, K, tol, roots, n, ev);
  *nr = n;
  return ok ? 1 : 0;
}

// One bracket refinement; method 0 legacy, 1 Brent, 2 Chandrupatla.
double mdo_refine(const double *h, const double *vp, const double *vs, const double *rho, int L, int wave,
                  double period, double a, double b, double tol, int method, int *evaluations) {
  Medium m = plain_medium(h, vp, vs, rho, L, wave);
  m.refinement = method;
  const double omega = two_pi() / period;
  const double fa = evaluate(m, omega, a), fb = evaluate(m, omega, b);
  int ev = 2;
  const double r = refine_root(m, omega, a, b, fa, fb, tol, ev);
  if (evaluations) *evaluations = ev;
  return r;
}

double mdo_group_fd(const double *h, const double *vp, const double *vs, const double *rho, int L, int wave,
                    double period, double root, double lower, double upper, double rtol, unsigned *flags) {
  const Medium m = plain_medium(h, vp, vs, rho, L, wave);
  const GroupResult g = group_velocity_fd(m, two_pi() / period, root, lower, upper, rtol);
  if (flags) *flags = g.flags;
  return g.velocity;
}

double mdo_group_analytic(const double *h, const double *vp, const double *vs, const double *rho, int L, int wave,
                          double period, double c) {
  return group_velocity_analytic(h, vp, vs, rho, L, wave, two_pi() / period, c);
}

// dF/d(which) at (period, c) with the positive attenuation and normalization
// factors held constant (the derivative of F itself at a root); which as in
// secular_direction.
double mdo_secular_derivative(const double *h, const double *vp, const double *vs, const double *rho, int L, int wave,
                              double period, double c, int which) {
  return secular_direction(h, vp, vs, rho, L, wave, two_pi() / period, c, which).d;
}

void mdo_phase_gradient(const double *h, const double *vp, const double *vs, const double *rho, int L, int wave,
                        double period, double c, double *out) {
  phase_gradient(h, vp, vs, rho, L, wave, two_pi() / period, c, out);
}

// Full multimode solve over B models x nf periods. periods is [nf] when
// period_stride == 0, else [B, period_stride] per-model rows. Returns the
// number of secular evaluations. group_mode: 0 none, 1 reference FD, 2 analytic.
unsigned long long mdo_solve(const double *h, const double *vp, const double *vs, const double *rho, int B, int L,
                             const double *periods, int nf, int period_stride, const double *bounds, int modes,
                             int wave, int automatic, double tol, double accept, int refinement, int love_phase,
                             double close_rtol, int threads, int group_mode, double group_rtol, double *velocity,
                             uint32_t *flags, int32_t *levels, int32_t *counts, double *group,
                             uint32_t *group_flags) {
  Config cfg{B, L, nf, modes, wave, automatic, refinement, love_phase, group_mode,
             tol, accept, close_rtol, group_rtol, period_stride};
  Outputs out{velocity, flags, levels, counts, group_mode ? group : nullptr, group_mode ? group_flags : nullptr};
  std::vector<Model> models(B);
  parallel_for(B, threads, [&](size_t i, int) {
    const size_t off = i * L;
    prepare_model(models[i], h + off, vp + off, vs + off, rho + off, cfg, bounds[2 * i], bounds[2 * i + 1]);
  });
  const int nthreads = std::max(1, threads);
  std::vector<Work> work(nthreads);
  parallel_for(static_cast<size_t>(B) * nf, threads, [&](size_t t, int id) {
    const int i = static_cast<int>(t / nf), f = static_cast<int>(t % nf);
    const double period = period_stride ? periods[static_cast<size_t>(i) * period_stride + f] : periods[f];
    const size_t out_off = static_cast<size_t>(i) * modes * nf + f;
    solve_trace(models[i], bounds[2 * i], bounds[2 * i + 1], period, cfg, work[id], out_off, t, out);
  });
  unsigned long long ev = 0;
  for (auto &w : work) ev += w.evaluations;
  return ev;
}

}  // extern "C"
