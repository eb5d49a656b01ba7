// gfx950 kernels of the multimode root pipeline.
//
// A trace is one (model, period) pair. For every trace the pipeline finds the
// first K = modes + 2 phase-velocity roots of the secular function F(c) in the
// trace's window [lo, hi] (the reference's internal guard roots,
// cpp/cpu/solver.cpp:418), refines the first `modes`, and optionally their
// group velocities:
//
//   scan_kernel<T,S>   one 256-lane block per trace; nc = 256*S uniform
//                      c-samples, one lane per sample (S per lane), layer
//                      constants staged in LDS; block-wide prefix scans
//                      classify and compact the ascending events up to the
//                      K-th certain sign change into the trace's event list:
//                      certain brackets, FP32-uncertain nodes (|F| < tau, or
//                      inside the branch-point and low-velocity guards) and
//                      same-sign troughs; uncertain nodes and troughs are
//                      appended to device work queues.
//   rescan_kernel      FP64 rescan (persistent) of FP32 traces whose event
//                      list overflowed: no node of an FP64 grid is uncertain.
//   recheck_kernel     FP64 re-evaluation of every queued uncertain node,
//                      one lane per node (dense across traces).
//   trough_kernel      FP64 golden-section probe of each queued trough for a
//                      hidden root pair (one lane per trough).
//   assemble_kernel    one lane per trace: merges certain FP32 signs with FP64
//                      rechecks and trough splits into the ordered bracket
//                      list, the root count and trace flags.
//   refine_kernel      one lane per bracket: safeguarded Brent-Dekker in FP64
//                      (reference cpp/common/refinement.hpp:41-63 stop rule).
//   group_kernel       one lane per root: analytic implicit group velocity
//                      U = -dF/dk / dF/dw by forward-mode dual numbers.
//
// The persistent kernels (recheck, trough, refine, group) read their work
// counts from device memory, so the whole sequence is launched back to back on
// one stream with no host synchronisation and can be captured in a hipGraph.
#include <hip/hip_runtime.h>
#include <cstdint>
#include "pipeline.hpp"
#include "secular.hpp"

namespace mdr {

// ---------------------------------------------------------------------------
// Layer tables.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ Layer<T> make_layer(const PrepRow &r, double omega, double iw2, int wave) {
  Layer<T> l;
  l.h = T(r.h);
  l.kb = T(omega * r.ivs);
  if (wave == 0) {
    l.ka = T(omega * r.ivp);
    l.gk = T(r.g2 * iw2);
    l.rho = T(r.rho);
    l.irho = T(r.irho);
    l.irho2 = T(r.irho2);
  } else {
    l.ka = T(0);
    l.gk = T(0);
    l.rho = T(0.5 * r.rho * r.g2);  // mu = rho vs^2
    l.irho = T(r.imu);
    l.irho2 = T(0);
  }
  l.pad = T(0);
  return l;
}

// Secular F from prepared rows held in global memory (FP64 lane-per-item
// work); optionally the summed log2 normalization scale.
__device__ __forceinline__ double secular_rows(const PrepRow *rows, int L, int top, int wave, double omega,
                                               double c, float *log_scale = nullptr) {
  const double k = omega / c, iw2 = 1. / (omega * omega);
  float ls = 0.f;
  double F;
  if (wave == 0) {
    const double k2 = k * k;
    double e[5];
    double m = rayleigh_halfspace(make_layer<double>(rows[L - 1], omega, iw2, 0), k, k2, e);
    if (log_scale) ls += scale_log2(m);
    for (int j = L - 2; j >= top; --j) {
      m = rayleigh_layer(make_layer<double>(rows[j], omega, iw2, 0), k, k2, e);
      if (log_scale) ls += scale_log2(m);
    }
    F = top ? rayleigh_water_top(make_layer<double>(rows[0], omega, iw2, 0), k, e) : e[0];
  } else {
    const Layer<double> hs = make_layer<double>(rows[L - 1], omega, iw2, 1);
    double e[2] = {hs.rho * m_sqrt((k + hs.kb) * m_fabs(k - hs.kb)), 1.};
    double m = rescale<2>(e);
    if (log_scale) ls += scale_log2(m);
    for (int j = L - 2; j >= top; --j) {
      m = love_layer(make_layer<double>(rows[j], omega, iw2, 1), k, e);
      if (log_scale) ls += scale_log2(m);
    }
    F = e[0];
  }
  if (log_scale) *log_scale = ls;
  return F;
}

// ---------------------------------------------------------------------------
// Wave-level helpers (one wavefront owns one trace; no block barriers).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ uint64_t lanes_below(int lane) { return lane ? (~0ull >> (64 - lane)) : 0ull; }
__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  return v;
}

// Secular F with layer constants formed on the fly from global rows (models
// too deep for an LDS layer table).
template <class T>
__device__ __forceinline__ T secular_global(const PrepRow *rows, int L, int top, int wave, double omega,
                                            double iw2, T k, float *log_scale) {
  float ls = 0.f;
  if (wave == 0) {
    const T k2 = k * k;
    T e[5];
    ls += scale_log2(rayleigh_halfspace(make_layer<T>(rows[L - 1], omega, iw2, 0), k, k2, e));
    for (int j = L - 2; j >= top; --j)
      ls += scale_log2(rayleigh_layer(make_layer<T>(rows[j], omega, iw2, 0), k, k2, e));
    *log_scale = ls;
    return top ? rayleigh_water_top(make_layer<T>(rows[0], omega, iw2, 0), k, e) : e[0];
  }
  const Layer<T> hs = make_layer<T>(rows[L - 1], omega, iw2, 1);
  T e[2] = {hs.rho * m_sqrt((k + hs.kb) * m_fabs(k - hs.kb)), T(1)};
  ls += scale_log2(rescale<2>(e));
  for (int j = L - 2; j >= top; --j) ls += scale_log2(love_layer(make_layer<T>(rows[j], omega, iw2, 1), k, e));
  *log_scale = ls;
  return e[0];
}

// ---------------------------------------------------------------------------
// Scan: one wavefront per trace.
//
// Nodes. Uniform grid: nc nodes lo + j (hi - lo)/(nc - 1). Adaptive grid:
// a table of kFloor uniform points plus every layer speed (vs, and vp for
// Rayleigh) inside the window, each table cell subdivided so that it holds
// at most 1/rho of the WKB mode count N(c) = (1/pi) sum_j h_j (sqrt(kb_j^2 -
// k^2)_+ + sqrt(ka_j^2 - k^2)_+) (the reference's density_base / T,
// cpp/common/physics.hpp:130-138); a cell that starts at a layer speed gets
// quadratically spaced nodes, matching N ~ sqrt(c - v) there.
//
// Nodes are evaluated 64 at a time (62 new per batch, 2 overlap so every
// classified lane sees both neighbours). Ballots classify certain sign
// changes, FP32-uncertain nodes and same-sign troughs; popcount prefixes
// compact them into the trace's ordered event list. The scan stops after the
// K-th certain sign change, when no later node can change the first K roots.
// ---------------------------------------------------------------------------
constexpr int kWaves = kBlock / 64;
constexpr int kTabMax = kFloor + 2 * kAdaptRows;

template <class T>
struct WaveScratch {
  Layer<T> lay[kAdaptRows];
  double tc[kTabMax];        // table coordinates, ascending
  double ks[2 * kAdaptRows];  // layer speeds inside the window (unsorted)
  float tn[kTabMax];         // N at table points
  int pre[kTabMax];          // node index of table point m
  unsigned char kink[kTabMax];
};

template <class T>
__device__ void scan_wave(const PipeArgs &a, int t, WaveScratch<T> &sh, int lane) {
  const TraceInfo ti = a.traces[t];
  const ModelInfo mi = a.models[ti.model];
  const int tl = t - a.trace_begin;
  Event *ev = a.events + static_cast<size_t>(tl) * kEventCap;
  TraceScan *ts = a.scan + tl;
  constexpr bool fp32 = sizeof(T) == 4;
  if (mi.empty) {
    if (lane == 0) *ts = TraceScan{0, 0u, mi.empty, 0, 0, 0};
    return;
  }
  const double lo = mi.lo, hi = mi.hi;
  // Far below the largest shear speed the FP32 compound matrix cancels
  // (reference limit: max Vs / 16); such windows go straight to the FP64 grid.
  if (fp32 && lo < mi.stable_low) {
    if (lane == 0) {
      *ts = TraceScan{0, kTraceOverflow, 0, 0, 0, 0};
      a.overflow_list[atomicAdd(a.overflow_count, 1)] = tl;
    }
    return;
  }
  const int L = mi.L, top = mi.top, K = mi.K;
  const double omega = ti.omega, iw2 = 1. / (omega * omega);
  const PrepRow *rows = a.rows + mi.row0;
  const bool in_lds = L <= kAdaptRows;
  if (in_lds)
    for (int j = lane; j < L; j += 64) sh.lay[j] = make_layer<T>(rows[j], omega, iw2, a.wave);
  wave_sync();
  const bool adaptive = a.adaptive && in_lds;
  int total = a.nc, ntab = 0;
  const double dc = (hi - lo) / double(a.nc - 1);
  if (adaptive) {
    // Layer speeds strictly inside the window (row `lane`; L <= 64).
    double v0 = 0., v1 = 0.;
    int nmine = 0;
    if (lane < L - 1 && rows[lane].h > 0.) {
      const PrepRow r = rows[lane];
      if (r.ivs > 0.) {
        const double v = 1. / r.ivs;
        if (v > lo && v < hi) { v0 = v; ++nmine; }
      }
      if (a.wave == 0) {
        const double v = 1. / r.ivp;
        if (v > lo && v < hi) {
          if (nmine) v1 = v; else v0 = v;
          ++nmine;
        }
      }
    }
    const int incl = wave_incl_scan(nmine, lane);
    const int nk = __shfl(incl, 63, 64);
    const int off = incl - nmine;
    if (nmine > 0) sh.ks[off] = v0;
    if (nmine > 1) sh.ks[off + 1] = v1;
    wave_sync();
    const double dt = (hi - lo) / double(kFloor - 1);
    auto floor_point = [&](int m) { return m == kFloor - 1 ? hi : lo + dt * double(m); };
    auto density = [&](double c) {
      const float k = float(omega / c), k2 = k * k;
      float n = 0.f;
      for (int j = top; j < L - 1; ++j) {
        const float h = float(sh.lay[j].h);
        const float kb = float(sh.lay[j].kb), ka = float(sh.lay[j].ka);
        n += h * sqrtf(fmaxf(0.f, kb * kb - k2));
        if (a.wave == 0) n += h * sqrtf(fmaxf(0.f, ka * ka - k2));
      }
      if (top && a.wave == 0) {  // acoustic water layer: P term only
        const float h = float(sh.lay[0].h), ka = float(sh.lay[0].ka);
        n += h * sqrtf(fmaxf(0.f, ka * ka - k2));
      }
      return n * 0.3183098861837907f;  // 1/pi
    };
    // Merge the kFloor uniform points with the layer speeds by rank.
    {
      const double tm = floor_point(lane);
      int below = 0;
      for (int r = 0; r < nk; ++r) below += sh.ks[r] < tm;
      sh.tc[lane + below] = tm;
      sh.kink[lane + below] = 0;
      sh.tn[lane + below] = density(tm);
    }
    for (int q = 0; q < nmine; ++q) {
      const double v = q ? v1 : v0;
      const int me = off + q;
      int rank = 0;
      for (int r = 0; r < nk; ++r) {
        const double u = sh.ks[r];
        rank += (u < v) || (u == v && r < me);
      }
      int ubelow = 0;
      for (int m = 0; m < kFloor; ++m) ubelow += floor_point(m) <= v;
      sh.tc[rank + ubelow] = v;
      sh.kink[rank + ubelow] = 1;
      sh.tn[rank + ubelow] = density(v);
    }
    wave_sync();
    // Density nodes per table cell; node index of every table point.
    ntab = kFloor + nk;
    int carry = 0;
    for (int base = 0; base < ntab; base += 64) {
      const int m = base + lane;
      int n = 0;
      if (m + 1 < ntab && sh.tc[m + 1] > sh.tc[m]) {
        const float want = ceilf(a.rho * (sh.tn[m + 1] - sh.tn[m])) - 1.f;
        n = want > 0.f ? (want < float(kCellMax) ? int(want) : kCellMax) : 0;
      }
      const int inc = wave_incl_scan(n, lane);
      if (m < ntab) sh.pre[m] = m + carry + inc - n;
      carry += __shfl(inc, 63, 64);
    }
    wave_sync();
    total = sh.pre[ntab - 1] + 1;
  }
  // Coordinate of node j (ascending in j).
  auto node_c = [&](int j) -> double {
    if (!adaptive) return lo + dc * double(j);
    int l = 0, h = ntab - 1;  // largest m with pre[m] <= j
    while (l < h) {
      const int mid = (l + h + 1) >> 1;
      if (sh.pre[mid] <= j) l = mid; else h = mid - 1;
    }
    const int i = j - sh.pre[l];
    if (i == 0) return sh.tc[l];
    const int n = sh.pre[l + 1] - sh.pre[l];  // interior nodes + 1
    const double s = double(i) / double(n);
    return sh.tc[l] + (sh.tc[l + 1] - sh.tc[l]) * (sh.kink[l] ? s * s : s);
  };
  const float tau = fp32 ? ti.tau : 0.f;
  const double hi_guard = mi.cutoff_guard;
  const float theta = kTroughMax;
  int nb = 0, ne = 0;
  bool overflow = false;
  for (int base = 0;; base += 62) {
    const int j = base + lane;
    const bool active = j < total;
    double c = 0.;
    float f = 0.f, sc = 0.f;
    bool u = false;
    if (active) {
      c = node_c(j);
      const T k = T(omega / c);
      T F;
      if (in_lds)
        F = a.wave == 0 ? rayleigh_F_scale<T>(sh.lay, L, top, k, &sc) : love_F_scale<T>(sh.lay, L, top, k, &sc);
      else
        F = secular_global<T>(rows, L, top, a.wave, omega, iw2, k, &sc);
      f = float(F);
      u = fp32 && (!(fabsf(f) >= tau) || c >= hi_guard);
    }
    const float fl = __shfl(f, (lane + 63) & 63, 64), fr = __shfl(f, (lane + 1) & 63, 64);
    const float sl = __shfl(sc, (lane + 63) & 63, 64), sr = __shfl(sc, (lane + 1) & 63, 64);
    const double cl = __shfl(c, (lane + 63) & 63, 64), cr = __shfl(c, (lane + 1) & 63, 64);
    const uint64_t um = __ballot(u);
    const bool ul = (um >> ((lane + 63) & 63)) & 1, ur = (um >> ((lane + 1) & 63)) & 1;
    const bool has_left = j >= 1, has_right = j + 1 < total;
    // Lanes 1..62 see both neighbours; lane 0 only at the first node.
    const bool classify = active && (lane >= 1 || base == 0) && lane <= 62;
    int kind = kNone;
    if (classify) {
      if (u) kind = kUncertain;
      else if (has_right && !ur && signbit(f) != signbit(fr)) kind = kBracket;
      else if (has_left && has_right && !ul && !ur && signbit(fl) == signbit(f) && signbit(fr) == signbit(f)) {
        // Same-sign window: a hidden root pair shows as a small local minimum
        // of |F| (trough) or as a dip of the log2 normalization scale (a
        // collapsing state vector between two very close roots).
        const float af = fabsf(f);
        if (af < fabsf(fl) && af < fabsf(fr) && af < theta) kind = kTrough;
        else if (sc < sl && sc < sr && 0.5f * (sl + sr) - sc > kDipDepth) kind = kDip;
      }
    }
    const uint64_t bm = __ballot(kind == kBracket);
    uint64_t em = __ballot(kind != kNone);
    bool done = total - 1 - base <= 62;
    const int nbb = __popcll(bm);
    if (nb + nbb >= K) {  // cut at the K-th certain sign change
      uint64_t m = bm;
      for (int r = nb + 1; r < K; ++r) m &= m - 1;
      const int cut = __builtin_ctzll(m);
      em &= cut == 63 ? ~0ull : ((2ull << cut) - 1ull);
      done = true;
    }
    nb += nbb;
    const int nnew = __popcll(em);
    if (ne + nnew > kEventCap) {
      overflow = true;
      break;
    }
    const bool mine = (em >> lane) & 1;
    const uint64_t uqm = em & __ballot(kind == kUncertain), tqm = em & __ballot(kind == kTrough || kind == kDip);
    int qu = 0, qt = 0;
    if (lane == 0) {
      if (uqm) qu = atomicAdd(a.unc_count, __popcll(uqm));
      if (tqm) qt = atomicAdd(a.trough_count, __popcll(tqm));
    }
    qu = __shfl(qu, 0, 64);
    qt = __shfl(qt, 0, 64);
    if (mine) {
      const int slot = ne + __popcll(em & lanes_below(lane));
      Event x;
      x.c0 = has_left ? cl : NAN;
      x.c1 = c;
      x.c2 = has_right ? cr : NAN;
      x.v = 0.;
      x.x = NAN;
      x.fl = fl;
      x.fc = f;
      x.fr = fr;
      x.node = j;
      x.kind = kind;
      x.sc = sc;
      ev[slot] = x;
      const uint32_t g = static_cast<uint32_t>(tl * kEventCap + slot);
      if (kind == kUncertain) a.unc_queue[qu + __popcll(uqm & lanes_below(lane))] = g;
      if (kind == kTrough || kind == kDip) a.trough_queue[qt + __popcll(tqm & lanes_below(lane))] = g;
    }
    ne += nnew;
    if (done) break;
  }
  if (lane == 0) {
    *ts = TraceScan{overflow ? 0 : ne, overflow ? kTraceOverflow : 0u, 0, nb, 0, total};
    // An FP32 trace whose events overflow is rescanned on the FP64 grid,
    // which has no uncertain nodes.
    if (overflow && fp32) a.overflow_list[atomicAdd(a.overflow_count, 1)] = tl;
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void scan_kernel(PipeArgs a) {
  __shared__ WaveScratch<T> scratch[kWaves];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int t = a.trace_begin + blockIdx.x * kWaves + w;
  if (t < a.trace_end) scan_wave<T>(a, t, scratch[w], lane);
}

// FP64 rescan of the traces whose FP32 event list overflowed (persistent,
// one wavefront per listed trace).
__global__ __launch_bounds__(kBlock) void rescan_kernel(PipeArgs a) {
  __shared__ WaveScratch<double> scratch[kWaves];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int n = *a.overflow_count;
  for (int q = blockIdx.x * kWaves + w; q < n; q += gridDim.x * kWaves)
    scan_wave<double>(a, a.overflow_list[q] + a.trace_begin, scratch[w], lane);
}

// ---------------------------------------------------------------------------
// FP64 recheck of uncertain nodes (dense lanes over the device queue).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void recheck_kernel(PipeArgs a) {
  const int n = min(*a.unc_count, a.queue_cap);
  for (int q = blockIdx.x * kBlock + threadIdx.x; q < n; q += gridDim.x * kBlock) {
    const uint32_t g = a.unc_queue[q];
    const int tl = g / kEventCap;
    Event &x = a.events[g];
    const TraceInfo ti = a.traces[tl + a.trace_begin];
    const ModelInfo mi = a.models[ti.model];
    x.v = secular_rows(a.rows + mi.row0, mi.L, mi.top, a.wave, ti.omega, x.c1);
  }
}

// ---------------------------------------------------------------------------
// FP64 probe of same-sign troughs: golden-section search for the minimum of
// sign*F over [c_{i-1}, c_{i+1}], stopping at the first opposite-sign point.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void trough_kernel(PipeArgs a) {
  const int n = min(*a.trough_count, a.queue_cap);
  for (int q = blockIdx.x * kBlock + threadIdx.x; q < n; q += gridDim.x * kBlock) {
    const uint32_t g = a.trough_queue[q];
    const int tl = g / kEventCap;
    Event &x = a.events[g];
    if (x.kind != kTrough && x.kind != kDip) continue;
    const bool dip = x.kind == kDip;
    const TraceInfo ti = a.traces[tl + a.trace_begin];
    const ModelInfo mi = a.models[ti.model];
    const PrepRow *rows = a.rows + mi.row0;
    double lo = x.c0, hi = x.c2;
    const double sign = x.fc < 0.f ? -1. : 1.;
    // Golden-section minimization of sign*F (trough) or of the log2 scale
    // (dip), stopping at the first point where F takes the opposite sign.
    auto probe = [&](double c, float &key) {
      float ls;
      const double f = sign * secular_rows(rows, mi.L, mi.top, a.wave, ti.omega, c, &ls);
      key = dip ? ls : float(f);
      return f;
    };
    constexpr double phi = 0.6180339887498949;
    double p = hi - phi * (hi - lo), r = lo + phi * (hi - lo);
    float kp, kr;
    double fp = probe(p, kp), fr = probe(r, kr);
    float kmin = fminf(kp, kr);
    double split = NAN, fsplit = 0.;
    for (int it = 0; it < kTroughIters; ++it) {
      if (fp <= 0.) { split = p; fsplit = sign * fp; break; }
      if (fr <= 0.) { split = r; fsplit = sign * fr; break; }
      // A dip that does not deepen like log|c - c*| is not a collapsing resonance.
      if (dip && it >= kDipStopIter && kmin > x.sc - kDipStopDrop) break;
      if (kp < kr) {
        hi = r; r = p; fr = fp; kr = kp;
        p = hi - phi * (hi - lo);
        fp = probe(p, kp);
        kmin = fminf(kmin, kp);
      } else {
        lo = p; p = r; fp = fr; kp = kr;
        r = lo + phi * (hi - lo);
        fr = probe(r, kr);
        kmin = fminf(kmin, kr);
      }
    }
    x.x = split;
    x.v = fsplit;
  }
}

// ---------------------------------------------------------------------------
// Assemble ordered brackets per trace.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void assemble_kernel(PipeArgs a) {
  const int tl = blockIdx.x * kBlock + threadIdx.x;
  const int t = tl + a.trace_begin;
  if (t >= a.trace_end) return;
  const TraceInfo ti = a.traces[t];
  const ModelInfo mi = a.models[ti.model];
  TraceScan &ts = a.scan[tl];
  Bracket *out = a.brackets + static_cast<size_t>(tl) * a.kmax;
  const int K = mi.K;
  int nb = 0;
  const unsigned flags = ts.flags;
  if (ts.status == 0 && !(flags & kTraceOverflow)) {
    const Event *ev = a.events + static_cast<size_t>(tl) * kEventCap;
    auto emit = [&](double lo, double hi, double flo, double fhi) {
      if (nb < a.kmax) out[nb] = Bracket{lo, hi, flo, fhi};
      ++nb;
    };
    int prev_unc = -2;
    double prev_v = 0.;
    for (int e = 0; e < ts.nev && nb < K; ++e) {
      const Event x = ev[e];
      if (x.kind == kBracket) {
        emit(x.c1, x.c2, x.fc, x.fr);
      } else if (x.kind == kTrough || x.kind == kDip) {
        if (x.x == x.x) {  // hidden pair split at x
          emit(x.c0, x.x, x.fl, x.v);
          if (nb < K) emit(x.x, x.c2, x.v, x.fr);
        }
      } else {  // uncertain node with its FP64 value; owns its left cell when
                // that neighbour is certain and always its right cell
        const double v = x.v;
        if (x.c0 == x.c0) {
          const double fl = prev_unc == x.node - 1 ? prev_v : double(x.fl);
          if (signbit(fl) != signbit(v) && nb < K) emit(x.c0, x.c1, fl, v);
        }
        const bool right_unc = e + 1 < ts.nev && ev[e + 1].kind == kUncertain && ev[e + 1].node == x.node + 1;
        if (!right_unc && x.c2 == x.c2 && nb < K) {
          const double fr = double(x.fr);
          if (signbit(v) != signbit(fr)) emit(x.c1, x.c2, v, fr);
        }
        prev_unc = x.node;
        prev_v = v;
      }
    }
  }
  const int count = nb < K ? nb : K;
  ts.count = (flags & kTraceOverflow) ? -1 : (ts.status ? 0 : count);
}

// ---------------------------------------------------------------------------
// Brent-Dekker refinement in FP64, one lane per (trace, root) slot.
// ---------------------------------------------------------------------------
__device__ double brent(const PrepRow *rows, int L, int top, int wave, double omega, double a, double b, double fa,
                        double fb, double tol, int &evals) {
  if (fa == 0.) return a;
  if (fb == 0.) return b;
  const double stop = fmax(16. * 2.220446049250313e-16 * fmax(fabs(a), fabs(b)),
                           fmin((b - a) / 64., .125 * tol * fmax(fabs(a), fabs(b))));
  double c = b, fc = fb, d = b - a, e = d;
  for (int j = 0; j < 160; ++j) {
    if ((fb < 0.) == (fc < 0.)) { c = a; fc = fa; e = d = b - a; }
    if (fabs(fc) < fabs(fb)) { a = b; fa = fb; b = c; fb = fc; c = a; fc = fa; }
    if (fb == 0.) return b;
    const double mid = .5 * (c - b), step = .5 * stop;
    if (fabs(c - b) <= stop) return .5 * (b + c);
    if (fabs(e) >= step && fabs(fa) > fabs(fb)) {
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
      const double ratio = p / q;
      if (ratio == ratio && fabs(ratio) < INFINITY && 2. * p < fmin(3. * mid * q - fabs(step * q), fabs(old * q)))
        d = ratio;
      else { d = mid; e = mid; }
    } else { d = mid; e = mid; }
    a = b; fa = fb;
    b += fabs(d) > step ? d : (mid > 0. ? step : -step);
    fb = secular_rows(rows, L, top, wave, omega, b);
    ++evals;
    if (!(fb == fb) || fabs(fb) == INFINITY) return NAN;
  }
  return NAN;
}

__global__ __launch_bounds__(kBlock) void refine_kernel(PipeArgs a) {
  const int ntl = a.trace_end - a.trace_begin;
  const int total = ntl * a.kmax;
  for (int q = blockIdx.x * kBlock + threadIdx.x; q < total; q += gridDim.x * kBlock) {
    const int tl = q / a.kmax, m = q % a.kmax;
    const TraceScan &ts = a.scan[tl];
    const TraceInfo ti = a.traces[tl + a.trace_begin];
    const ModelInfo mi = a.models[ti.model];
    double root = NAN;
    if (m < ts.count && m < mi.K) {
      const Bracket br = a.brackets[q];
      int ev = 0;
      root = brent(a.rows + mi.row0, mi.L, mi.top, a.wave, ti.omega, br.lo, br.hi, br.flo, br.fhi, a.tol, ev);
    }
    a.roots[q] = root;
  }
}

// ---------------------------------------------------------------------------
// Finalize: trace flags, counts, levels, and the phase-velocity output.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void finalize_kernel(PipeArgs a) {
  const int tl = blockIdx.x * kBlock + threadIdx.x;
  const int t = tl + a.trace_begin;
  if (t >= a.trace_end) return;
  const TraceInfo ti = a.traces[t];
  const ModelInfo mi = a.models[ti.model];
  const TraceScan ts = a.scan[tl];
  const double *roots = a.roots + static_cast<size_t>(tl) * a.kmax;
  unsigned flags = 0;
  int count = ts.count, level = a.fp32 ? kLevelScan32 : kLevelScan64;
  if (ts.status) { level = ts.status == 1 ? -4 : -3; count = 0; }
  if (ts.flags & kTraceOverflow) { flags |= kFlagUnconverged; level = -1; count = -1; }
  bool cutoff = false, close = false;
  int nfinite = 0;
  for (int m = 0; m < count && m < mi.K; ++m) {
    const double r = roots[m];
    if (!(r == r)) { flags |= kFlagRefineFailed; continue; }
    ++nfinite;
    if (m > 0 && roots[m - 1] == roots[m - 1] && r - roots[m - 1] <= a.close_rtol * r) close = true;
  }
  for (int m = 0; m < mi.modes; ++m) {
    double r = m < count ? roots[m] : NAN;
    if (r == r && r >= .99 * mi.vs_half) cutoff = true;
    const size_t cell = static_cast<size_t>(mi.out0) + static_cast<size_t>(m) * mi.T + ti.f;
    if (a.velocity64) a.velocity64[cell] = r;
    if (a.velocity32) a.velocity32[cell] = float(r);
  }
  (void)nfinite;
  if (cutoff) flags |= kFlagNearCutoff;
  if (close) flags |= kFlagCloseRoot;
  const size_t tc = static_cast<size_t>(mi.trace_out0) + ti.f;
  a.flags[tc] = flags;
  a.levels[tc] = level;
  a.counts[tc] = count;
}

// ---------------------------------------------------------------------------
// Analytic group velocity U = dw/dk = -F_k / F_w at each refined root.
// ---------------------------------------------------------------------------
__device__ double group_at(const PrepRow *rows, int L, int top, int wave, double omega, double c) {
  const double iw = 1. / omega, iw2 = iw * iw;
  const Dual2 k(omega / c, 1., 0.), w(omega, 0., 1.);
  auto layer = [&](const PrepRow &r) {
    Layer<Dual2> l;
    l.h = Dual2(r.h);
    l.kb = Dual2(omega * r.ivs, 0., r.ivs);
    if (wave == 0) {
      l.ka = Dual2(omega * r.ivp, 0., r.ivp);
      l.gk = Dual2(r.g2 * iw2, 0., -2. * r.g2 * iw2 * iw);
      l.rho = Dual2(r.rho);
      l.irho = Dual2(r.irho);
      l.irho2 = Dual2(r.irho2);
    } else {
      l.rho = Dual2(0.5 * r.rho * r.g2);
      l.irho = Dual2(r.imu);
    }
    return l;
  };
  Dual2 F;
  if (wave == 0) {
    const Dual2 k2 = k * k;
    Dual2 e[5];
    rayleigh_halfspace(layer(rows[L - 1]), k, k2, e);
    for (int j = L - 2; j >= top; --j) rayleigh_layer(layer(rows[j]), k, k2, e);
    F = top ? rayleigh_water_top(layer(rows[0]), k, e) : e[0];
  } else {
    const Layer<Dual2> hs = layer(rows[L - 1]);
    Dual2 e[2] = {hs.rho * m_sqrt((k + hs.kb) * m_fabs(k - hs.kb)), Dual2(1.)};
    for (int j = L - 2; j >= top; --j) love_layer(layer(rows[j]), k, e);
    F = e[0];
  }
  (void)w;
  return -F.a / F.b;
}

__global__ __launch_bounds__(kBlock) void group_kernel(PipeArgs a) {
  const int ntl = a.trace_end - a.trace_begin;
  const int total = ntl * a.kmax;
  for (int q = blockIdx.x * kBlock + threadIdx.x; q < total; q += gridDim.x * kBlock) {
    const int tl = q / a.kmax, m = q % a.kmax;
    const TraceInfo ti = a.traces[tl + a.trace_begin];
    const ModelInfo mi = a.models[ti.model];
    if (m >= mi.modes) continue;
    const TraceScan &ts = a.scan[tl];
    const size_t cell = static_cast<size_t>(mi.out0) + static_cast<size_t>(m) * mi.T + ti.f;
    const double *roots = a.roots + static_cast<size_t>(tl) * a.kmax;
    const double c = m < ts.count ? roots[m] : NAN;
    double u = NAN;
    unsigned gf = 0;
    if (!(c == c)) gf = kGroupMissing;
    else {
      const bool close = (m > 0 && c - roots[m - 1] <= a.close_rtol * c) ||
                         (m + 1 < ts.count && roots[m + 1] - c <= a.close_rtol * roots[m + 1]);
      if (close) gf = kGroupClose;
      else {
        u = group_at(a.rows + mi.row0, mi.L, mi.top, a.wave, ti.omega, c);
        if (!(u == u) || fabs(u) == INFINITY) { u = NAN; gf = kGroupNonfinite; }
      }
    }
    a.group[cell] = u;
    a.group_flags[cell] = gf;
  }
}

// Reset the device work counters (a kernel, so the sequence can be graph-captured).
__global__ void reset_kernel(int *counters, int n) {
  if (threadIdx.x < n) counters[threadIdx.x] = 0;
}

// Accumulate this chunk's work counters into the run statistics:
// [0] traces, [1] FP64 rechecks, [2] trough/dip probes, [3] FP64 rescans.
__global__ void stats_kernel(const int *counters, unsigned long long *stats, int traces) {
  if (threadIdx.x == 0) {
    stats[0] += static_cast<unsigned long long>(traces);
    stats[1] += static_cast<unsigned long long>(counters[0]);
    stats[2] += static_cast<unsigned long long>(counters[1]);
    stats[3] += static_cast<unsigned long long>(counters[2]);
  }
}

}  // namespace mdr

// ---------------------------------------------------------------------------
// Launchers used by the host pipeline.
// ---------------------------------------------------------------------------
namespace mdr {

template <class T>
static hipError_t launch_scan_t(const PipeArgs &a, hipStream_t s) {
  const int n = a.trace_end - a.trace_begin;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL((scan_kernel<T>), dim3((n + kWaves - 1) / kWaves), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_pipeline(const ScanArgs &sa, const PipeArgs &pa, bool fp32, bool group, int persistent_blocks,
                           hipStream_t s, hipEvent_t *marks) {
  const int n = sa.trace_end - sa.trace_begin;
  if (n <= 0) return hipSuccess;
  auto mark = [&](int i) {
    if (marks) (void)hipEventRecord(marks[i], s);
  };
  hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, s, sa.counters, 4);
  mark(0);
  hipError_t err = fp32 ? launch_scan_t<float>(sa, s) : launch_scan_t<double>(sa, s);
  if (err != hipSuccess) return err;
  mark(1);
  const int pb = persistent_blocks;
  if (fp32) hipLaunchKernelGGL(rescan_kernel, dim3(pb), dim3(kBlock), 0, s, pa);
  mark(2);
  hipLaunchKernelGGL(recheck_kernel, dim3(pb), dim3(kBlock), 0, s, pa);
  hipLaunchKernelGGL(trough_kernel, dim3(pb), dim3(kBlock), 0, s, pa);
  mark(3);
  hipLaunchKernelGGL(assemble_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, pa);
  mark(4);
  hipLaunchKernelGGL(refine_kernel, dim3(pb), dim3(kBlock), 0, s, pa);
  hipLaunchKernelGGL(finalize_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, pa);
  mark(5);
  if (group) hipLaunchKernelGGL(group_kernel, dim3(pb), dim3(kBlock), 0, s, pa);
  mark(6);
  if (pa.stats) hipLaunchKernelGGL(stats_kernel, dim3(1), dim3(64), 0, s, pa.counters, pa.stats, n);
  return hipGetLastError();
}

}  // namespace mdr

// ---------------------------------------------------------------------------
// Device-side preparation for rectangular batches resident in HBM.
// ---------------------------------------------------------------------------
namespace mdr {

__global__ void prep_models_kernel(const double *h, const double *vp, const double *vs, const double *rho,
                                   const double *bounds, int B, int L, int T, int modes, int wave, PrepRow *rows,
                                   ModelInfo *models) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= B) return;
  const size_t off = static_cast<size_t>(m) * L;
  for (int j = 0; j < L; ++j) rows[off + j] = make_prep_row(h[off + j], vp[off + j], vs[off + j], rho[off + j]);
  double slow;
  models[m] = make_model_info(h + off, vp + off, vs + off, L, bounds[2 * m], bounds[2 * m + 1], wave, modes, T,
                              static_cast<long long>(off), static_cast<long long>(m) * modes * T,
                              static_cast<long long>(m) * T, &slow);
}

__global__ void prep_traces_kernel(const double *periods, int row_periods, int B, int T, const ModelInfo *models,
                                   PrepParams p, TraceInfo *traces) {
  const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<long long>(B) * T) return;
  const int m = static_cast<int>(t / T), f = static_cast<int>(t % T);
  const double period = row_periods ? periods[t] : periods[f];
  traces[t] = make_trace(m, f, period, models[m].slowness, p);
}

hipError_t launch_prep(const double *h, const double *vp, const double *vs, const double *rho, const double *bounds,
                       const double *periods, int row_periods, int B, int L, int T, int modes, const PrepParams &p,
                       PrepRow *rows, ModelInfo *models, TraceInfo *traces, hipStream_t s) {
  hipLaunchKernelGGL(prep_models_kernel, dim3((B + 255) / 256), dim3(256), 0, s, h, vp, vs, rho, bounds, B, L, T,
                     modes, p.wave, rows, models);
  const long long n = static_cast<long long>(B) * T;
  hipLaunchKernelGGL(prep_traces_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, periods,
                     row_periods, B, T, models, p, traces);
  return hipGetLastError();
}

}  // namespace mdr

// ---------------------------------------------------------------------------
// Diagnostic grid probe: F on the scan grid with the scan's own evaluators.
// ---------------------------------------------------------------------------
namespace mdr {

template <class T>
__global__ __launch_bounds__(kBlock) void probe_kernel(const TraceInfo *traces, const ModelInfo *models,
                                                       const PrepRow *rows_all, int ntraces, int nc, int wave,
                                                       double *out) {
  __shared__ Layer<T> lay[kMaxRows];
  const int t = blockIdx.x;
  if (t >= ntraces) return;
  const TraceInfo ti = traces[t];
  const ModelInfo mi = models[ti.model];
  const PrepRow *rows = rows_all + mi.row0;
  const double omega = ti.omega, iw2 = 1. / (omega * omega);
  for (int j = threadIdx.x; j < mi.L; j += kBlock) lay[j] = make_layer<T>(rows[j], omega, iw2, wave);
  __syncthreads();
  const double dc = (mi.hi - mi.lo) / double(nc - 1);
  for (int i = threadIdx.x; i < nc; i += kBlock) {
    const double c = mi.lo + dc * double(i);
    const T k = T(omega / c);
    const T f = wave == 0 ? rayleigh_F<T>(lay, mi.L, mi.top, k) : love_F<T>(lay, mi.L, mi.top, k);
    out[static_cast<size_t>(t) * nc + i] = double(f);
  }
}

hipError_t launch_probe(const TraceInfo *traces, const ModelInfo *models, const PrepRow *rows, int ntraces, int nc,
                        int wave, int fp32, double *out, hipStream_t s) {
  if (fp32)
    hipLaunchKernelGGL(probe_kernel<float>, dim3(ntraces), dim3(kBlock), 0, s, traces, models, rows, ntraces, nc,
                       wave, out);
  else
    hipLaunchKernelGGL(probe_kernel<double>, dim3(ntraces), dim3(kBlock), 0, s, traces, models, rows, ntraces, nc,
                       wave, out);
  return hipGetLastError();
}

}  // namespace mdr
