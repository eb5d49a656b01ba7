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

// Secular F from prepared rows held in global memory (FP64 lane-per-item work).
__device__ __forceinline__ double secular_rows(const PrepRow *rows, int L, int top, int wave, double omega,
                                               double c) {
  const double k = omega / c, iw2 = 1. / (omega * omega);
  if (wave == 0) {
    const double k2 = k * k;
    double e[5];
    rayleigh_halfspace(make_layer<double>(rows[L - 1], omega, iw2, 0), k, k2, e);
    for (int j = L - 2; j >= top; --j) rayleigh_layer(make_layer<double>(rows[j], omega, iw2, 0), k, k2, e);
    if (top) {
      Layer<double> wl = make_layer<double>(rows[0], omega, iw2, 0);
      return rayleigh_water_top(wl, k, e);
    }
    return e[0];
  }
  const Layer<double> hs = make_layer<double>(rows[L - 1], omega, iw2, 1);
  double e[2] = {hs.rho * m_sqrt((k + hs.kb) * m_fabs(k - hs.kb)), 1.};
  rescale<2>(e);
  for (int j = L - 2; j >= top; --j) love_layer(make_layer<double>(rows[j], omega, iw2, 1), k, e);
  return e[0];
}

// ---------------------------------------------------------------------------
// Block-wide inclusive scan of four counters (256 lanes = 4 wave64s).
// ---------------------------------------------------------------------------
struct Count4 {
  int ev, br, un, tr;
};
__device__ __forceinline__ Count4 operator+(Count4 a, Count4 b) {
  return {a.ev + b.ev, a.br + b.br, a.un + b.un, a.tr + b.tr};
}
__device__ __forceinline__ Count4 shfl_up4(Count4 v, int d) {
  return {__shfl_up(v.ev, d, 64), __shfl_up(v.br, d, 64), __shfl_up(v.un, d, 64), __shfl_up(v.tr, d, 64)};
}
// Returns the inclusive prefix; total is written to *total (all lanes).
__device__ __forceinline__ Count4 block_scan4(Count4 v, Count4 *wave_tot, Count4 *total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const Count4 u = shfl_up4(v, d);
    if (lane >= d) v = v + u;
  }
  if (lane == 63) wave_tot[wave] = v;
  __syncthreads();
  Count4 off{0, 0, 0, 0}, tot{0, 0, 0, 0};
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    const Count4 t = wave_tot[w];
    if (w < wave) off = off + t;
    tot = tot + t;
  }
  __syncthreads();
  *total = tot;
  return v + off;
}

// ---------------------------------------------------------------------------
// Scan: FP32 (or FP64) uniform grid, classification and event compaction.
// ---------------------------------------------------------------------------
template <class T, int S>
struct ScanShared {
  Layer<T> lay[kMaxRows];
  float fval[kBlock * S];
  unsigned char unc[kBlock * S];
  Count4 wave_tot[kBlock / 64];
  int cut_node, q_unc, q_tr;
};

// Scans trace t with the whole block. Every return is block-uniform.
template <class T, int S>
__device__ void scan_trace(const PipeArgs &a, int t, ScanShared<T, S> &sh) {
  constexpr int NC = kBlock * S;
  Layer<T> *lay = sh.lay;
  float *fval = sh.fval;
  unsigned char *unc = sh.unc;
  Count4 *wave_tot = sh.wave_tot;
  const TraceInfo ti = a.traces[t];
  const ModelInfo mi = a.models[ti.model];
  const int L = mi.L, top = mi.top, K = mi.K;
  const int tid = threadIdx.x;
  const double omega = ti.omega, iw2 = 1. / (omega * omega);
  const PrepRow *rows = a.rows + mi.row0;
  for (int j = tid; j < L; j += kBlock) lay[j] = make_layer<T>(rows[j], omega, iw2, a.wave);
  Event *ev = a.events + static_cast<size_t>(t - a.trace_begin) * kEventCap;
  TraceScan *ts = a.scan + (t - a.trace_begin);
  if (mi.empty) {
    if (tid == 0) {
      ts->nev = 0;
      ts->flags = 0;
      ts->status = mi.empty;
    }
    return;
  }
  __syncthreads();

  const double lo = mi.lo, dc = (mi.hi - mi.lo) / double(NC - 1);
  const bool fp32 = sizeof(T) == 4;
  const float tau = fp32 ? ti.tau : 0.f;
  const double hi_guard = mi.cutoff_guard;
  // Far below the largest shear speed the FP32 compound matrix cancels
  // (reference limit: max Vs / 16); such windows go straight to the FP64 grid.
  if (fp32 && lo < mi.stable_low) {
    if (tid == 0) {
      ts->nev = 0;
      ts->flags = kTraceOverflow;
      ts->status = 0;
      a.overflow_list[atomicAdd(a.overflow_count, 1)] = t - a.trace_begin;
    }
    return;
  }
  // Evaluate: sample i = s*kBlock + tid (consecutive lanes, consecutive samples).
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int i = s * kBlock + tid;
    const double c = lo + dc * double(i);
    const T k = T(omega / c);
    const T f = a.wave == 0 ? rayleigh_F<T>(lay, L, top, k) : love_F<T>(lay, L, top, k);
    const float fv = float(f);
    fval[i] = fv;
    unc[i] = (fp32 && (!(fabsf(fv) >= tau) || c >= hi_guard)) ? 1 : 0;
  }
  __syncthreads();

  // Classify the contiguous nodes [tid*S, tid*S + S) of this lane.
  const float theta = kTroughMax;
  unsigned char kind[S];
  Count4 cnt{0, 0, 0, 0};
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int i = tid * S + s;
    const float fi = fval[i];
    const bool ui = unc[i];
    unsigned char k8 = kNone;
    if (ui) k8 = kUncertain;
    else if (i + 1 < NC && !unc[i + 1] && (signbit(fi) != signbit(fval[i + 1]))) k8 = kBracket;
    else if (i >= 1 && i + 1 < NC && !unc[i - 1] && !unc[i + 1]) {
      const float fl = fval[i - 1], fr = fval[i + 1];
      const float ai = fabsf(fi);
      if (signbit(fl) == signbit(fi) && signbit(fr) == signbit(fi) && ai < fabsf(fl) && ai < fabsf(fr) &&
          ai < theta)
        k8 = kTrough;
    }
    kind[s] = k8;
    cnt.ev += k8 != kNone;
    cnt.br += k8 == kBracket;
    cnt.un += k8 == kUncertain;
    cnt.tr += k8 == kTrough;
  }
  Count4 total;
  const Count4 incl = block_scan4(cnt, wave_tot, &total);
  const Count4 excl{incl.ev - cnt.ev, incl.br - cnt.br, incl.un - cnt.un, incl.tr - cnt.tr};
  // Cut after the K-th certain bracket (the first K roots lie at or before it).
  if (tid == 0) sh.cut_node = NC;
  __syncthreads();
  if (excl.br < K && incl.br >= K) {
    int b = excl.br;
#pragma unroll
    for (int s = 0; s < S; ++s)
      if (kind[s] == kBracket && ++b == K) sh.cut_node = tid * S + s;
  }
  __syncthreads();
  const int cut = sh.cut_node;
  // Events at or before the cut, and how many of them are uncertain / troughs.
  Count4 kept{0, 0, 0, 0};
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int i = tid * S + s;
    const bool keep = kind[s] != kNone && i <= cut;
    kept.ev += keep;
    kept.un += keep && kind[s] == kUncertain;
    kept.tr += keep && kind[s] == kTrough;
  }
  Count4 ktot;
  const Count4 kin = block_scan4(kept, wave_tot, &ktot);
  const bool overflow = ktot.ev > kEventCap;
  if (tid == 0) {
    sh.q_unc = (!overflow && ktot.un) ? atomicAdd(a.unc_count, ktot.un) : 0;
    sh.q_tr = (!overflow && ktot.tr) ? atomicAdd(a.trough_count, ktot.tr) : 0;
    ts->nev = overflow ? 0 : ktot.ev;
    ts->flags = overflow ? kTraceOverflow : 0u;
    ts->status = 0;
    ts->certain = total.br;
    // An FP32 trace whose events overflow is rescanned on the FP64 grid,
    // which has no uncertain nodes.
    if (overflow && fp32) a.overflow_list[atomicAdd(a.overflow_count, 1)] = t - a.trace_begin;
  }
  __syncthreads();
  if (overflow) return;
  const int q_unc = sh.q_unc, q_tr = sh.q_tr;
  int e = kin.ev - kept.ev, u = kin.un - kept.un, r = kin.tr - kept.tr;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int i = tid * S + s;
    if (kind[s] == kNone || i > cut) continue;
    Event x;
    x.node = i;
    x.kind = kind[s];
    x.fl = i > 0 ? fval[i - 1] : 0.f;
    x.fc = fval[i];
    x.fr = i + 1 < NC ? fval[i + 1] : 0.f;
    x.v = 0.;
    x.x = 0.;
    if (kind[s] == kBracket) x.fr = fval[i + 1];
    ev[e] = x;
    const uint32_t gidx = static_cast<uint32_t>((t - a.trace_begin) * kEventCap + e);
    if (kind[s] == kUncertain && q_unc + u < a.queue_cap) a.unc_queue[q_unc + u] = gidx;
    if (kind[s] == kTrough && q_tr + r < a.queue_cap) a.trough_queue[q_tr + r] = gidx;
    u += kind[s] == kUncertain;
    r += kind[s] == kTrough;
    ++e;
  }
}

template <class T, int S>
__global__ __launch_bounds__(kBlock) void scan_kernel(PipeArgs a) {
  __shared__ ScanShared<T, S> sh;
  const int t = blockIdx.x + a.trace_begin;
  if (t < a.trace_end) scan_trace<T, S>(a, t, sh);
}

// FP64 rescan of the traces whose FP32 event list overflowed (persistent).
template <int S>
__global__ __launch_bounds__(kBlock) void rescan_kernel(PipeArgs a) {
  __shared__ ScanShared<double, S> sh;
  const int n = *a.overflow_count;
  for (int q = blockIdx.x; q < n; q += gridDim.x) {
    scan_trace<double, S>(a, a.overflow_list[q] + a.trace_begin, sh);
    __syncthreads();
  }
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
    const double dc = (mi.hi - mi.lo) / double(a.nc - 1);
    const double c = mi.lo + dc * double(x.node);
    x.v = secular_rows(a.rows + mi.row0, mi.L, mi.top, a.wave, ti.omega, c);
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
    const TraceInfo ti = a.traces[tl + a.trace_begin];
    const ModelInfo mi = a.models[ti.model];
    const PrepRow *rows = a.rows + mi.row0;
    const double dc = (mi.hi - mi.lo) / double(a.nc - 1);
    double lo = mi.lo + dc * double(x.node - 1), hi = mi.lo + dc * double(x.node + 1);
    const double sign = x.fc < 0.f ? -1. : 1.;
    constexpr double phi = 0.6180339887498949;
    double p = hi - phi * (hi - lo), r = lo + phi * (hi - lo);
    double fp = sign * secular_rows(rows, mi.L, mi.top, a.wave, ti.omega, p);
    double fr = sign * secular_rows(rows, mi.L, mi.top, a.wave, ti.omega, r);
    double split = NAN, fsplit = 0.;
    for (int it = 0; it < kTroughIters; ++it) {
      if (fp <= 0.) { split = p; fsplit = sign * fp; break; }
      if (fr <= 0.) { split = r; fsplit = sign * fr; break; }
      if (fp < fr) {
        hi = r; r = p; fr = fp;
        p = hi - phi * (hi - lo);
        fp = sign * secular_rows(rows, mi.L, mi.top, a.wave, ti.omega, p);
      } else {
        lo = p; p = r; fp = fr;
        r = lo + phi * (hi - lo);
        fr = sign * secular_rows(rows, mi.L, mi.top, a.wave, ti.omega, r);
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
  unsigned flags = ts.flags;
  if (ts.status == 0 && !(flags & kTraceOverflow)) {
    const Event *ev = a.events + static_cast<size_t>(tl) * kEventCap;
    const double dc = (mi.hi - mi.lo) / double(a.nc - 1);
    auto node_c = [&](int i) { return mi.lo + dc * double(i); };
    auto emit = [&](double lo, double hi, double flo, double fhi) {
      if (nb < a.kmax) out[nb] = Bracket{lo, hi, flo, fhi};
      ++nb;
    };
    int prev_unc = -2;
    double prev_v = 0.;
    const bool fp32 = a.fp32;
    for (int e = 0; e < ts.nev && nb < K; ++e) {
      const Event x = ev[e];
      const int i = x.node;
      if (x.kind == kBracket) {
        emit(node_c(i), node_c(i + 1), x.fc, x.fr);
      } else if (x.kind == kTrough) {
        if (x.x == x.x) {  // split found
          emit(node_c(i - 1), x.x, x.fl, x.v);
          if (nb < K) emit(x.x, node_c(i + 1), x.v, x.fr);
        }
      } else {  // uncertain node with FP64 value x.v
        const double v = x.v;
        if (i > 0) {
          const bool left_unc = prev_unc == i - 1;
          const double fl = left_unc ? prev_v : double(x.fl);
          if (signbit(fl) != signbit(v) && nb < K) emit(node_c(i - 1), node_c(i), fl, v);
        }
        const bool right_unc = e + 1 < ts.nev && ev[e + 1].kind == kUncertain && ev[e + 1].node == i + 1;
        if (!right_unc && i + 1 < a.nc && nb < K) {
          const double fr = double(x.fr);
          if (signbit(v) != signbit(fr)) emit(node_c(i), node_c(i + 1), v, fr);
        }
        prev_unc = i;
        prev_v = v;
      }
    }
    (void)fp32;
  }
  const int count = nb < K ? nb : K;
  ts.count = (flags & kTraceOverflow) ? -1 : (ts.status ? 0 : count);
  ts.flags = flags;
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

}  // namespace mdr

// ---------------------------------------------------------------------------
// Launchers used by the host pipeline.
// ---------------------------------------------------------------------------
namespace mdr {

template <class T>
static hipError_t launch_scan_t(const PipeArgs &a, int nc, hipStream_t s) {
  const int n = a.trace_end - a.trace_begin;
  if (n <= 0) return hipSuccess;
  switch (nc) {
    case 256: hipLaunchKernelGGL((scan_kernel<T, 1>), dim3(n), dim3(kBlock), 0, s, a); break;
    case 512: hipLaunchKernelGGL((scan_kernel<T, 2>), dim3(n), dim3(kBlock), 0, s, a); break;
    case 1024: hipLaunchKernelGGL((scan_kernel<T, 4>), dim3(n), dim3(kBlock), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

static hipError_t launch_rescan(const PipeArgs &a, int nc, int blocks, hipStream_t s) {
  switch (nc) {
    case 256: hipLaunchKernelGGL((rescan_kernel<1>), dim3(blocks), dim3(kBlock), 0, s, a); break;
    case 512: hipLaunchKernelGGL((rescan_kernel<2>), dim3(blocks), dim3(kBlock), 0, s, a); break;
    case 1024: hipLaunchKernelGGL((rescan_kernel<4>), dim3(blocks), dim3(kBlock), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
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
  hipError_t err = fp32 ? launch_scan_t<float>(sa, pa.nc, s) : launch_scan_t<double>(sa, pa.nc, s);
  if (err != hipSuccess) return err;
  mark(1);
  const int pb = persistent_blocks;
  if (fp32 && (err = launch_rescan(pa, pa.nc, pb, s)) != hipSuccess) return err;
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
