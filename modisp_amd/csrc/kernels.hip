// gfx950 kernels of the multimode root pipeline.
//
// A trace is one (model, period) pair. For every trace the pipeline finds the
// first K = modes + 2 phase-velocity roots of the secular function F(c) in the
// trace's window [lo, hi] (the reference's internal guard roots,
// cpp/cpu/solver.cpp:418), refines them in FP64, and optionally computes the
// group velocities of the first `modes`:
//
//   table_kernel       one wavefront per model: the adaptive grid's table (the
//                      window's floor points merged with the layer speeds, with
//                      the frequency-free mode density).
//   scan_kernel        one wavefront per trace, W traces per block in an
//                      XCD-aware block order; layer constants in LDS; nodes
//                      (density-adaptive or uniform) evaluated 64 at a time in
//                      FP32 or FP64; ballots classify certain sign changes,
//                      FP32-uncertain nodes and hidden-pair candidates
//                      (troughs, scale dips) into the trace's event list, up to
//                      the K-th certain sign change.
//   rescan_kernel      FP64 rescan (persistent) of FP32 traces whose event list
//                      overflowed or whose window needs FP64.
//   compact_kernel     recheck and probe queues from the per-trace event masks
//                      (one atomic per block).
//   recheck_kernel     FP64 value of every uncertain node (dense lanes).
//   unc_probe_kernel   trough test at the rechecked nodes.
//   trough_kernel      golden-section probes of the candidates, one search per
//                      lane as a state machine, idle lanes fed the next probes
//                      of the wave's runs in order: an FP32 pass, and an FP64
//                      pass over the searches it handed over (or every probe of
//                      the FP64 path); a point of opposite sign splits the cell.
//   assemble_kernel    one lane per trace: ordered brackets and the root count.
//   love_range_kernel  Love: the exact phase count of each trace's window (a
//                      trace with one takes no pair probes);
//   love_count_kernel  Love: the exact phase count verifies each trace;
//   love_phase_kernel  a disagreeing trace is re-bracketed by phase bisection.
//   refine_kernel      one bracket per lane as a state machine: FP64
//                      Brent-Dekker (default), Chandrupatla or the legacy
//                      regula falsi (reference cpp/common/refinement.hpp:11-65);
//                      tiles of 64 traces taken mode-major, idle lanes fed in
//                      order, the tile's rows staged in LDS; guard roots whose
//                      close-root tests their brackets decide are not refined.
//   finalize_kernel    outputs in the reference layout, trace flags.
//   group_kernel       mode-major lanes: analytic implicit group velocity
//                      U = -dF/dk / dF/dw by forward-mode dual numbers.
//   slope/anchor_flag/continuation/gather kernels: anchor continuation.
//
// Work counts live in device memory and the consumer kernels are persistent
// kernels that read them, so the sequence runs back to back on one stream with
// no host synchronisation and can be captured in a hipGraph.
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

// Running product of normalization scales as mant 2^ex, mant in [1, 2) in
// FP32, from the bits of each scale m (a normal positive double): its
// exponent exactly, its leading 20 mantissa bits into the product. Integer
// and FP32 operations only; one log2 at the end. The dropped low mantissa
// bits perturb the product by < 1e-6 relative, smoothly enough for bracket
// interpolation.
struct ScaleProduct {
  float mant = 1.f;
  int ex = 0;
  __device__ __forceinline__ void add(double m) {
    const int hi = __double2hiint(m);
    ex += (hi >> 20) - 1023;
    mant *= __int_as_float(((hi & 0xfffff) << 3) | 0x3f800000);
    const int carry = (__float_as_int(mant) >> 23) - 127;  // 0 or 1: mant in [1, 4)
    ex += carry;
    mant = __int_as_float(__float_as_int(mant) - (carry << 23));
  }
};

// Secular F from prepared rows held in global memory (FP64 lane-per-item
// work); optionally the summed log2 normalization scale (float or double sum).

__device__ __forceinline__ double secular_rows(const PrepRow *rows, int L, int top, int wave, double omega,
                                               double c, float *log_scale = nullptr, ScaleProduct *product = nullptr) {
  const double k = omega / c, iw2 = 1. / (omega * omega);
  float ls = 0.f;
  double F;
  if (wave == 0) {
    const double k2 = k * k;
    double e[5];
    double m = rayleigh_halfspace(make_layer<double>(rows[L - 1], omega, iw2, 0), k, k2, e);
    if (log_scale) ls += scale_log2(m);
    if (product) product->add(m);
    for (int j = L - 2; j >= top; --j) {
      m = rayleigh_layer(make_layer<double>(rows[j], omega, iw2, 0), k, k2, e);
      if (log_scale) ls += scale_log2(m);
      if (product) product->add(m);
    }
    F = top ? rayleigh_water_top(make_layer<double>(rows[0], omega, iw2, 0), k, e) : e[0];
  } else {
    const Layer<double> hs = make_layer<double>(rows[L - 1], omega, iw2, 1);
    double e[2] = {hs.rho * m_sqrt((k + hs.kb) * m_fabs(k - hs.kb)), 1.};
    double m = rescale<2>(e);
    if (log_scale) ls += scale_log2(m);
    if (product) product->add(m);
    for (int j = L - 2; j >= top; --j) {
      m = love_layer(make_layer<double>(rows[j], omega, iw2, 1), k, e);
      if (log_scale) ls += scale_log2(m);
      if (product) product->add(m);
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
// Adds a per-lane count to a run statistic with one atomic per wavefront
// (every lane of the wave must call it).
__device__ __forceinline__ void wave_count(unsigned long long *stat, unsigned n) {
  unsigned long long v = n;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(stat, v);
}
__device__ __forceinline__ uint64_t lanes_below(int lane) { return lane ? (~0ull >> (64 - lane)) : 0ull; }
// The value held by lane - 1 (from_left) or lane + 1 (from_right), modulo 64:
// one DPP wave rotate each (wave_ror:1 / wave_rol:1), no LDS round trip.
// Every lane of the wave must be active.
__device__ __forceinline__ int from_left(int x) { return __builtin_amdgcn_mov_dpp(x, 0x13C, 0xF, 0xF, false); }
__device__ __forceinline__ int from_right(int x) { return __builtin_amdgcn_mov_dpp(x, 0x134, 0xF, 0xF, false); }
__device__ __forceinline__ float from_left(float x) { return __int_as_float(from_left(__float_as_int(x))); }
__device__ __forceinline__ float from_right(float x) { return __int_as_float(from_right(__float_as_int(x))); }
__device__ __forceinline__ double from_left(double x) {
  return __hiloint2double(from_left(__double2hiint(x)), from_left(__double2loint(x)));
}
__device__ __forceinline__ double from_right(double x) {
  return __hiloint2double(from_right(__double2hiint(x)), from_right(__double2loint(x)));
}
__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  return v;
}

// The phase velocity whose wavenumber omega/c is exactly representable in
// FP32: evaluating F in FP32 at k = float(omega / c) evaluates it at this c.
// The result is kept inside [lo, hi] (one FP32 step of k inward).
__device__ __forceinline__ double snap32(double omega, double c, double lo, double hi, float *k_out = nullptr) {
  float k = float(omega / c);
  double s = omega / double(k);
  if (s > hi) {
    k = __int_as_float(__float_as_int(k) + 1);  // k > 0: next float up
    s = omega / double(k);
  } else if (s < lo) {
    k = __int_as_float(__float_as_int(k) - 1);  // next float down
    s = omega / double(k);
  }
  // float(omega / s) rounds back to k exactly (the double quotient is within
  // 2^-52 of k, far inside half a float ulp), so callers may take k as is.
  if (k_out) *k_out = k;
  return s;
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

// Per-wavefront LDS scratch for models of at most R rows.
template <class T, int R>
struct WaveScratch {
  static constexpr int kTab = kFloor + 2 * R;
  Layer<T> lay[R];
  double tc[kTab];   // table coordinates, ascending
  float tn[kTab];    // N at table points
  int pre[kTab];     // node index of table point m
  unsigned char kink[kTab];
};

template <class T, int R>
__device__ void scan_wave(const PipeArgs &a, int t, WaveScratch<T, R> &sh, int lane) {
  const TraceInfo ti = a.traces[t];
  const ModelInfo mi = a.models[ti.model];
  const int tl = t - a.trace_begin;
  Event *ev = a.events + static_cast<size_t>(tl) * kEventCap;
  TraceScan *ts = a.scan + tl;
  constexpr bool fp32 = sizeof(T) == 4;
  if (mi.empty) {
    if (lane == 0) *ts = TraceScan{0, 0u, mi.empty, 0, 0, 0, 0ull, 0ull, 0u};
    return;
  }
  const double lo = mi.lo, hi = mi.hi;
  // Far below the largest shear speed the FP32 compound matrix cancels
  // (reference limit: max Vs / 16). Where no root can lie down there the
  // nodes below the limit are FP64-rechecked like any uncertain node;
  // otherwise the window goes straight to the FP64 grid.
  if (fp32 && lo < mi.stable_low && !mi.low_free) {
    if (lane == 0) {
      *ts = TraceScan{0, kTraceOverflow, 0, 0, 0, 0, 0ull, 0ull, 0u};
      a.overflow_list[atomicAdd(a.overflow_count, 1)] = tl;
    }
    return;
  }
  const int L = mi.L, top = mi.top, K = mi.K;  // L <= R (the launcher picks R)
  const double omega = ti.omega, iw2 = 1. / (omega * omega);
  const PrepRow *rows = a.rows + mi.row0;
  for (int j = lane; j < L; j += 64) sh.lay[j] = make_layer<T>(rows[j], omega, iw2, a.wave);
  wave_sync();
  const bool adaptive = a.adaptive;
  int total = a.nc, ntab = 0;
  const double dc = (hi - lo) / double(a.nc - 1);
  if (adaptive) {
    // The model's table (window floor points and layer speeds with the
    // frequency-free mode density D, built once per model by table_kernel):
    // N(c) = omega D(c) at this trace's frequency.
    const TablePoint *tab = a.table + static_cast<size_t>(ti.model) * a.table_stride;
    ntab = a.ntab[ti.model];
    for (int m = lane; m < ntab; m += 64) {
      const TablePoint tp = tab[m];
      sh.tc[m] = tp.c;
      sh.tn[m] = float(omega * double(tp.d));
      sh.kink[m] = static_cast<unsigned char>(tp.kink);
    }
    wave_sync();
    // Density nodes per table cell; node index of every table point.
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
  uint64_t umask = 0ull, pmask = 0ull;
  bool overflow = false;
  int evaluated = 0;
  for (int base = 0;; base += 62) {
    const int j = base + lane;
    const bool active = j < total;
    evaluated += total - base < 64 ? total - base : 64;
    double c = 0.;
    float f = 0.f, sc = 0.f;
    bool u = false;
    if (active) {
      c = node_c(j);
      // FP32: an interior node moves to the phase velocity whose wavenumber is
      // the FP32 k actually evaluated, so a node's sign belongs to its
      // coordinate. The window ends stay exact and are always rechecked in FP64.
      const bool end = j == 0 || j == total - 1;
      float k32 = 0.f;
      if (fp32 && !end) c = snap32(omega, c, lo, hi, &k32);
      const T k = fp32 && !end ? T(k32) : T(omega / c);
      const T F = a.wave == 0 ? rayleigh_F_scale<T>(sh.lay, L, top, k, &sc) : love_F_scale<T>(sh.lay, L, top, k, &sc);
      f = float(F);
      u = fp32 && (!(fabsf(f) >= tau) || c >= hi_guard || end || c < mi.stable_low);
    }
    const float fl = from_left(f), fr = from_right(f);
    const float sl = from_left(sc), sr = from_right(sc);
    const double cl = from_left(c), cr = from_right(c);
    // Nodes closer than 1e-6 (relative) to a neighbour -- the quadratic
    // spacing after a layer speed packs them there when many modes crowd
    // above it -- are below what FP32 resolves (its k steps by 6e-8; snapped
    // nodes can even coincide, and at the speed itself FP32 takes the r -> 0
    // limit where FP64 does not): they are rechecked.
    if (fp32 && active && ((lane >= 1 && j >= 1 && c - cl < 1e-6 * c) || (lane <= 62 && j + 1 < total && cr - c < 1e-6 * c)))
      u = true;
    const uint64_t um = __ballot(u);
    const bool ul = (um >> ((lane + 63) & 63)) & 1, ur = (um >> ((lane + 1) & 63)) & 1;
    const bool has_left = j >= 1, has_right = j + 1 < total;
    // Lanes 1..62 see both neighbours; lane 0 only at the first node.
    const bool classify = active && (lane >= 1 || base == 0) && lane <= 62;
    // A sign change between two nodes outside the FP32 band. FP32 rounding of
    // the layer constants moves a root by up to ~1e-7 relative, so a node that
    // close to a root can show the wrong sign. Where the FP32 values cannot
    // place the root away from both ends, both ends are rechecked in FP64
    // (uncertain) and the assembly takes the root from whichever cell the
    // FP64 signs put it in (the change itself, or the cell left or right of
    // it). That is: a root close to one end (the smaller |F| below 1/50 of
    // the larger); a step (deep evanescent stacks: F jumps between +-A within
    // far less than a node spacing, so both ends show nearly the full |F| of
    // the nodes beyond them, and the step's position is unknown); a cell
    // narrower than 1e-4 relative (dense nodes next to a layer speed); and a
    // change at the batch edge, whose far neighbour is not at hand.
    const bool change = active && has_right && !u && !ur && lane <= 62 && signbit(f) != signbit(fr);
    const uint64_t cm = __ballot(change);
    const bool change_l = lane >= 1 && ((cm >> (lane - 1)) & 1);
    const float f2 = from_right(fr);  // value two nodes to the right
    bool verify = false;
    if (fp32 && change) {
      const float a1 = fabsf(f), a2 = fabsf(fr), amin = fminf(a1, a2), amax = fmaxf(a1, a2);
      const float outer = fmaxf(has_left ? fabsf(fl) : 0.f, j + 2 < total ? fabsf(f2) : 0.f);
      verify = lane >= 62 || amin < 0.02f * amax || amin >= 0.25f * outer || cr - c < 1e-4 * c;
    }
    // A change beside a scale dip (the state collapses there, so F is as
    // sensitive to rounding as at a step) is rechecked too.
    const bool dip_here = has_left && has_right && sc < sl && sc < sr && 0.5f * (sl + sr) - sc > kDipDepth;
    const bool dip_right = (__ballot(dip_here) >> ((lane + 1) & 63)) & 1;
    verify = verify || (fp32 && change && (dip_here || dip_right));
    const uint64_t vm = __ballot(verify);
    const bool verify_l = lane >= 1 && ((vm >> (lane - 1)) & 1);
    // A same-sign trough needs its neighbours' signs: next to an FP32-uncertain
    // node a local minimum of |F| below the probe threshold is rechecked too,
    // and the trough test is repeated on the FP64 values (unc_probe_kernel).
    // Otherwise a hidden pair beside the cutoff guard or the band would be
    // probed only by the FP64 path.
    const bool trough_pending = fp32 && !u && (ul || ur) && has_left && has_right && !(cl < mi.stable_low) &&
                                fabsf(f) < fabsf(fl) && fabsf(f) < fabsf(fr) &&
                                (fabsf(f) < theta || fabsf(f) < kTroughRel * fminf(fabsf(fl), fabsf(fr)));
    // A change whose right node is rechecked belongs to that node's event
    // (its left cell), not to a certain bracket here.
    // A node at a scale dip is rechecked as well: where the state collapses
    // F steps, so its FP32 sign decides which cells a pair probe searches.
    const bool recheck = u || verify || verify_l || trough_pending || (fp32 && dip_here);
    const bool recheck_r = (__ballot(recheck) >> ((lane + 1) & 63)) & 1;
    int kind = kNone, probe = 0;
    if (classify) {
      if (recheck) kind = kUncertain;
      else if (change && !recheck_r) kind = kBracket;
      // No probe next to FP32 values below the stable limit (meaningless
      // there, and the band holds no root).
      if (has_left && has_right && !(fp32 && cl < mi.stable_low)) {
        // A hidden root pair in a same-sign cell shows as a small local
        // minimum of |F| (trough) or as a dip of the log2 normalization scale
        // (a state vector that collapses between two very close roots). The
        // scale needs no sign, so a dip is flagged next to uncertain nodes too;
        // the probe later searches only the cells whose end signs agree.
        const float af = fabsf(f);
        const bool certain3 = !u && !ul && !ur;
        const bool same_l = signbit(fl) == signbit(f), same_r = signbit(fr) == signbit(f);
        // A same-sign local minimum of |F| that is small, or 10% below both
        // neighbours: F falling toward zero and rising again without a sign
        // change is the trace of two roots closer than the node spacing, and
        // a pair in a broad trough can leave no node below theta. (Steps of F
        // in deep evanescent stacks vary by far less between nodes.)
        // A scale dip, where there is one, leads the search unless |F| itself
        // is below theta: the scale follows a pair more surely than a shallow
        // trough of F does.
        const bool low3 = certain3 && same_l && same_r && af < fabsf(fl) && af < fabsf(fr);
        if (low3 && af < theta)
          probe = kTrough;
        else if (sc < sl && sc < sr && 0.5f * (sl + sr) - sc > kDipDepth && !(certain3 && !same_l && !same_r))
          probe = kDip;  // both wave types: a Love trace whose phase count is unavailable relies on it too
        else if (low3 && af < kTroughRel * fminf(fabsf(fl), fabsf(fr)))
          probe = kTrough;
      } else if ((has_left != has_right) && !fp32) {
        // A window end: a pair in the end cell shows only as |F| falling
        // toward the end (one side). The FP32 ends are rechecked in FP64 and
        // take this test on those values (unc_probe_kernel).
        const float an = has_left ? fabsf(fl) : fabsf(fr);
        const bool same = has_left ? signbit(fl) == signbit(f) : signbit(fr) == signbit(f);
        const float af = fabsf(f);
        if (same && af < an && (af < theta || af < kTroughRel * an)) probe = kTrough;
      }
      if (probe && kind == kNone) kind = kProbe;
    }
    // Sign changes counted toward the cut at K. FP64: every certain one.
    // FP32: a change whose neighbouring cells hold none (a lone node of the
    // wrong sign makes two adjacent changes that its recheck removes; a
    // change at lane 62 is left to be counted later), so the cut never comes
    // before K roots the FP64 signs confirm.
    const bool counted = fp32 ? (change && classify && !change_l && lane <= 61 && !((cm >> (lane + 1)) & 1))
                              : kind == kBracket;
    const uint64_t bm = __ballot(counted);
    uint64_t em = __ballot(kind != kNone);
    bool done = total - 1 - base <= 62;
    const int nbb = __popcll(bm);
    if (nb + nbb >= K) {  // cut at the K-th counted sign change (keeping its right node's event)
      uint64_t m = bm;
      for (int r = nb + 1; r < K; ++r) m &= m - 1;
      const int cut = __builtin_ctzll(m) + (fp32 ? 1 : 0);
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
    {
      // Event bits of this batch for the queue compaction: the lanes of em in
      // order are slots ne, ne + 1, ...; compress the uncertain / probed lane
      // ballots onto those slots (wave-uniform, scalar work).
      const uint64_t ub = __ballot(kind == kUncertain) & em, pb = __ballot(probe != 0) & em;
      uint64_t cu = 0ull, cp = 0ull;
      int i = 0;
      for (uint64_t m = em; m; m &= m - 1ull, ++i) {
        const uint64_t low = m & (~m + 1ull);
        if (ub & low) cu |= 1ull << i;
        if (pb & low) cp |= 1ull << i;
      }
      umask |= cu << ne;
      pmask |= cp << ne;
    }
    if (mine) {
      const int slot = ne + __popcll(em & lanes_below(lane));
      Event x;
      x.c0 = has_left ? cl : NAN;
      x.c1 = c;
      x.c2 = has_right ? cr : NAN;
      x.v = 0.;
      x.x = NAN;
      x.vx = 0.;
      x.fl = fl;
      x.fc = f;
      x.fr = fr;
      x.sl = sl;
      x.sc = sc;
      x.sr = sr;
      x.node = j;
      x.kind = kind;
      x.probe = probe;
      x.pad = 0;
      ev[slot] = x;
    }
    ne += nnew;
    if (done) break;
  }
  if (lane == 0) {
    *ts = TraceScan{overflow ? 0 : ne, overflow ? kTraceOverflow : 0u, 0, nb, 0, evaluated, umask, pmask, 0u};
    // An FP32 trace whose events overflow is rescanned on the FP64 grid,
    // which has no uncertain nodes.
    if (overflow && fp32) a.overflow_list[atomicAdd(a.overflow_count, 1)] = tl;
  }
}

// R: rows the LDS scratch holds; W: wavefronts (traces) per block.
// ---------------------------------------------------------------------------
// Per-model scan table (the adaptive grid's skeleton): the kFloor uniform
// points of the model's window merged by rank with every layer speed strictly
// inside it (vs, and vp for Rayleigh), each with the frequency-free mode
// density D(c) = (1/pi) sum_j h_j (sqrt(1/vs_j^2 - 1/c^2)+ + sqrt(1/vp_j^2 -
// 1/c^2)+), so that N(c) = omega D(c) for every period of the model. One
// wavefront per model; the traces of a model read the table from L2.
// ---------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(kBlock) void table_kernel(const ModelInfo *models, const PrepRow *rows_all, int nmodels,
                                                       int wave, int stride, TablePoint *table, int *ntab_out) {
  __shared__ double ks_all[kWaves][2 * R];
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int model = blockIdx.x * kWaves + w;
  if (model >= nmodels) return;
  const ModelInfo mi = models[model];
  double *ks = ks_all[w];
  TablePoint *tab = table + static_cast<size_t>(model) * stride;
  if (mi.empty || !(mi.hi > mi.lo)) {
    if (lane == 0) ntab_out[model] = 0;
    return;
  }
  const double lo = mi.lo, hi = mi.hi;
  const PrepRow *rows = rows_all + mi.row0;
  const int L = mi.L, top = mi.top;
  // Layer speeds strictly inside the window, compacted in row order.
  int nk = 0;
  for (int base = 0; base < L - 1; base += 64) {
    const int j = base + lane;
    double v0 = 0., v1 = 0.;
    int nmine = 0;
    if (j < L - 1 && rows[j].h > 0.) {
      const PrepRow r = rows[j];
      if (r.ivs > 0.) {
        const double v = 1. / r.ivs;
        if (v > lo && v < hi) { v0 = v; ++nmine; }
      }
      if (wave == 0) {
        const double v = 1. / r.ivp;
        if (v > lo && v < hi) {
          if (nmine) v1 = v; else v0 = v;
          ++nmine;
        }
      }
    }
    const int incl = wave_incl_scan(nmine, lane);
    const int off = nk + incl - nmine;
    if (nmine > 0) ks[off] = v0;
    if (nmine > 1) ks[off + 1] = v1;
    nk += __shfl(incl, 63, 64);
  }
  wave_sync();
  const double dt = (hi - lo) / double(kFloor - 1);
  auto floor_point = [&](int m) { return m == kFloor - 1 ? hi : lo + dt * double(m); };
  auto density = [&](double c) {
    const double is2 = 1. / (c * c);
    double n = 0.;
    for (int j = top; j < L - 1; ++j) {
      const PrepRow r = rows[j];
      n += r.h * sqrt(fmax(0., r.ivs * r.ivs - is2));
      if (wave == 0) n += r.h * sqrt(fmax(0., r.ivp * r.ivp - is2));
    }
    if (top && wave == 0) n += rows[0].h * sqrt(fmax(0., rows[0].ivp * rows[0].ivp - is2));  // water: P only
    return n * 0.3183098861837907;  // 1/pi
  };
  {
    const double tm = floor_point(lane);
    int below = 0;
    for (int r = 0; r < nk; ++r) below += ks[r] < tm;
    tab[lane + below] = TablePoint{tm, float(density(tm)), 0};
  }
  for (int q = lane; q < nk; q += 64) {
    const double v = ks[q];
    int rank = 0;
    for (int r = 0; r < nk; ++r) {
      const double u = ks[r];
      rank += (u < v) || (u == v && r < q);
    }
    // Floor points at or below v: the closed form, corrected against the same
    // floor_point values the merge above compares (v < hi).
    int m0 = min(kFloor - 2, max(0, int((v - lo) / dt)));
    while (m0 < kFloor - 2 && floor_point(m0 + 1) <= v) ++m0;
    while (m0 >= 0 && floor_point(m0) > v) --m0;
    tab[rank + m0 + 1] = TablePoint{v, float(density(v)), 1};
  }
  if (lane == 0) ntab_out[model] = kFloor + nk;
}

hipError_t launch_tables(const ModelInfo *models, const PrepRow *rows, int nmodels, int max_rows, int wave, int stride,
                         TablePoint *table, int *ntab, hipStream_t s) {
  if (nmodels <= 0) return hipSuccess;
  const dim3 grid((nmodels + kWaves - 1) / kWaves);
  if (max_rows <= kAdaptRows)
    hipLaunchKernelGGL(table_kernel<kAdaptRows>, grid, dim3(kBlock), 0, s, models, rows, nmodels, wave, stride, table,
                       ntab);
  else
    hipLaunchKernelGGL(table_kernel<kMaxRows>, grid, dim3(kBlock), 0, s, models, rows, nmodels, wave, stride, table,
                       ntab);
  return hipGetLastError();
}

// XCD-aware block order: the dispatcher deals blocks round-robin to the 8
// XCDs (block b runs on XCD b % 8), each with its own L2. Renumbering so that
// every XCD gets one contiguous range of blocks keeps the traces of a model
// (consecutive, sharing its layer rows and scan table) on one XCD, so each
// model's data is fetched from HBM into one L2 instead of eight. A bijection
// of [0, nb) for any nb.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  constexpr int kXcd = 8;
  const int q = nb / kXcd, r = nb % kXcd, x = b % kXcd, i = b / kXcd;
  return x * q + min(x, r) + i;
}

template <class T, int R, int W>
__global__ __launch_bounds__(64 * W) void scan_kernel(PipeArgs a) {
  __shared__ WaveScratch<T, R> scratch[W];
  // The wave index is uniform across the wavefront; readfirstlane lets the
  // compiler keep every per-trace quantity in scalar registers.
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int t = a.trace_begin + xcd_block(blockIdx.x, gridDim.x) * W + w;
  if (t < a.trace_end) scan_wave<T, R>(a, t, scratch[w], lane);
}

// FP64 rescan of the traces whose FP32 event list overflowed (persistent,
// one wavefront per listed trace).
template <int R, int W>
__global__ __launch_bounds__(64 * W) void rescan_kernel(PipeArgs a) {
  __shared__ WaveScratch<double, R> scratch[W];
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int n = *a.overflow_count;
  for (int q = blockIdx.x * W + w; q < n; q += gridDim.x * W)
    scan_wave<double, R>(a, __builtin_amdgcn_readfirstlane(a.overflow_list[q]) + a.trace_begin, scratch[w], lane);
}

// ---------------------------------------------------------------------------
// Work-queue compaction: one lane per trace counts its uncertain and probed
// events, a block-wide prefix sum places them, and one atomic per block
// reserves the block's range in each queue (the scan itself takes no
// device-wide atomics).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void compact_kernel(PipeArgs a) {
  __shared__ int part[2][kBlock / 64];
  __shared__ int base[2];
  const int tl = blockIdx.x * kBlock + threadIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const bool live = tl < a.trace_end - a.trace_begin;
  uint64_t um = 0ull, pm = 0ull;
  if (live) {
    const TraceScan ts = a.scan[tl];
    if (ts.status == 0 && !(ts.flags & kTraceOverflow)) {
      um = ts.umask;
      // A Love trace whose phase count is known needs no pair probe: the
      // count verifies its brackets and re-brackets it by phase where a pair
      // hides (love_count_kernel, love_phase_kernel).
      pm = (ts.flags & kScanPhaseReady) ? 0ull : ts.pmask;
    }
  }
  const int nu = __popcll(um), np = __popcll(pm);
  const int iu = wave_incl_scan(nu, lane), ip = wave_incl_scan(np, lane);
  if (lane == 63) { part[0][w] = iu; part[1][w] = ip; }
  __syncthreads();
  if (threadIdx.x == 0) {
    int su = 0, sp = 0;
    for (int k = 0; k < kBlock / 64; ++k) {
      const int u = part[0][k], p = part[1][k];
      part[0][k] = su; part[1][k] = sp;
      su += u; sp += p;
    }
    base[0] = su ? atomicAdd(a.unc_count, su) : 0;
    base[1] = sp ? atomicAdd(a.trough_count, sp) : 0;
  }
  __syncthreads();
  int qu = base[0] + part[0][w] + iu - nu, qp = base[1] + part[1][w] + ip - np;
  const uint32_t g0 = static_cast<uint32_t>(tl) * kEventCap;
  for (; um; um &= um - 1ull) a.unc_queue[qu++] = g0 + __builtin_ctzll(um);
  for (; pm; pm &= pm - 1ull) a.trough_queue[qp++] = g0 + __builtin_ctzll(pm);
}

// ---------------------------------------------------------------------------
// FP64 recheck of uncertain nodes (dense lanes over the device queue).
// ---------------------------------------------------------------------------
template <int Wave>  // per wave type: the Love instantiation needs far fewer registers
__global__ __launch_bounds__(kBlock) void recheck_kernel(PipeArgs a) {
  const int n = min(*a.unc_count, a.queue_cap);
  for (int q = blockIdx.x * kBlock + threadIdx.x; q < n; q += gridDim.x * kBlock) {
    const uint32_t g = a.unc_queue[q];
    const int tl = g / kEventCap;
    Event &x = a.events[g];
    const TraceInfo ti = a.traces[tl + a.trace_begin];
    const ModelInfo mi = a.models[ti.model];
    x.v = secular_rows(a.rows + mi.row0, mi.L, mi.top, Wave, ti.omega, x.c1);
  }
}

// ---------------------------------------------------------------------------
// Best available value at the three nodes of event e: FP64 where a node was
// rechecked (an uncertain neighbour is always the adjacent event), else the scan's.
// ---------------------------------------------------------------------------
struct NodeValues {
  double l, c, r;
  bool lu, ru;
};
// The fields of an event that its neighbours' values depend on.
struct NodeFields {
  int node, kind;
  double v;
};
__device__ inline NodeFields node_fields(const Event &x) { return NodeFields{x.node, x.kind, x.v}; }
// Node fields of the slots beside event e, clamped into the trace's kEventCap
// slots so that they can be read together with the event itself (a slot at
// or past the trace's event count is ignored by node_values).
__device__ inline void neighbour_fields(const Event *ev, int e, NodeFields &l, NodeFields &r) {
  l = node_fields(ev[e > 0 ? e - 1 : e]);
  r = node_fields(ev[e + 1 < kEventCap ? e + 1 : e]);
}
__device__ inline NodeValues node_values(const Event &x, int e, int nev, const NodeFields &l, const NodeFields &r) {
  NodeValues n;
  n.lu = e > 0 && l.node == x.node - 1 && l.kind == kUncertain;
  n.ru = e + 1 < nev && r.node == x.node + 1 && r.kind == kUncertain;
  n.l = n.lu ? l.v : double(x.fl);
  n.r = n.ru ? r.v : double(x.fr);
  n.c = x.kind == kUncertain ? x.v : double(x.fc);
  return n;
}

// Trough test at rechecked nodes: once an uncertain node has its FP64 value,
// a same-sign local minimum of |F| is probed like a certain one, so that an
// FP32-uncertain node never masks a pair of close roots. (Dips need no sign
// and are flagged by the scan itself.)
__global__ __launch_bounds__(kBlock) void unc_probe_kernel(PipeArgs a) {
  const int n = min(*a.unc_count, a.queue_cap);
  const int lane = threadIdx.x & 63;
  // Uniform trip count so that every lane of a wave reaches the ballot.
  const int stride = gridDim.x * kBlock, first = blockIdx.x * kBlock + (threadIdx.x & ~63);
  for (int q0 = first; q0 < n; q0 += stride) {
    const int q = q0 + lane;
    bool want = false;
    uint32_t g = 0;
    if (q < n) {
      g = a.unc_queue[q];
      const int tl = g / kEventCap, e = g % kEventCap;
      const Event *ev = a.events + static_cast<size_t>(tl) * kEventCap;
      // Every read this lane needs is issued at once: the event, the node
      // fields of the slots beside it (in bounds of the trace's kEventCap
      // slots; a slot at or past nev is ignored below), the trace's scan
      // record and the model's FP32 limit.
      const Event x = ev[e];
      NodeFields xl, xr;
      neighbour_fields(ev, e, xl, xr);
      const unsigned tflags = a.scan[tl].flags;
      const int nev = a.scan[tl].nev;
      // The FP32 probe pass cannot work below the FP32 limit (a root-free band).
      const double low = a.fp32 ? a.models[a.traces[tl + a.trace_begin].model].stable_low : 0.;
      const int kind = x.probe & 0xff;
      const bool has_l = x.c0 == x.c0, has_r = x.c2 == x.c2;
      const bool phase_ready = (tflags & kScanPhaseReady) != 0;  // no pair probes (compact_kernel)
      if (!phase_ready && kind != kTrough && (has_l || has_r) && !(has_l && x.c0 < low) && !(!has_l && x.c1 < low)) {
        const NodeValues v = node_values(x, e, nev, xl, xr);
        // Both neighbours, or the one neighbour of a window end.
        const double al = has_l ? fabs(v.l) : INFINITY, ar = has_r ? fabs(v.r) : INFINITY;
        const bool falls = (!has_l || signbit(v.l) == signbit(v.c)) && (!has_r || signbit(v.r) == signbit(v.c)) &&
                           fabs(v.c) < al && fabs(v.c) < ar;
        const bool trough_abs = falls && fabs(v.c) < double(kTroughMax);
        // A scale dip the scan flagged at an uncertain node is already queued;
        // where the FP64 values show |F| below theta there it is searched as a
        // trough (on sign * F, as the FP64 scan would have flagged it).
        if (trough_abs && kind == kDip) a.events[g].probe = (x.probe & ~0xff) | kTrough;
        // A scale dip the scan left unprobed because the FP32 signs changed on
        // both sides (a lone FP32 sign error there) is probed once the FP64
        // signs say otherwise, as the FP64 scan would have flagged it; a
        // relative trough only where there is no dip to follow.
        const bool same_l = signbit(v.l) == signbit(v.c), same_r = signbit(v.r) == signbit(v.c);
        const bool dip = kind == 0 && !trough_abs && has_l && has_r && x.sc < x.sl && x.sc < x.sr &&
                         0.5f * (x.sl + x.sr) - x.sc > kDipDepth && (same_l || same_r);
        const bool trough_rel = kind == 0 && !trough_abs && !dip && falls &&
                                fabs(v.c) < double(kTroughRel) * fmin(al, ar);
        if (dip) a.events[g].probe = (x.probe & ~0xff) | kDip;
        want = kind == 0 && (trough_abs || dip || trough_rel);
      }
    }
    const uint64_t m = __ballot(want);
    if (!m) continue;
    int slot = 0;
    if (lane == 0) slot = atomicAdd(a.trough_count, __popcll(m));
    slot = __shfl(slot, 0, 64) + __popcll(m & lanes_below(lane));
    if (want) {
      if ((a.events[g].probe & 0xff) == 0) a.events[g].probe = kTrough;
      if (slot < a.queue_cap) a.trough_queue[slot] = g;
    }
  }
}

// ---------------------------------------------------------------------------
// FP64 probe of same-sign troughs: golden-section search for the minimum of
// sign*F over [c_{i-1}, c_{i+1}], stopping at the first opposite-sign point.
// ---------------------------------------------------------------------------
// Golden-section search as a per-lane state machine: minimise sign*F
// (trough) or the log2 scale (dip) over the probe region, stopping at the
// first point where F takes the opposite sign.
struct GoldenState {
  double lo, hi, p, r, fp, fr, sign, pending;
  float kp, kr, kmin, sc;
  int it, stage, which;  // stage 0/1: first two points; 2: iterating. which: 0 p, 1 r
  bool dip;
};

__device__ __forceinline__ void golden_init(GoldenState &s, double lo, double hi, double sign, bool dip, float sc) {
  constexpr double phi = 0.6180339887498949;
  s.lo = lo; s.hi = hi; s.sign = sign; s.dip = dip; s.sc = sc;
  s.p = hi - phi * (hi - lo);
  s.r = lo + phi * (hi - lo);
  s.it = 0; s.stage = 0; s.which = 0;
  s.fp = s.fr = 0.; s.kp = s.kr = 0.f;
  s.kmin = INFINITY;
}

// 0: F is needed at s.pending; 1: finished with `split` (NaN: no pair found).
__device__ __forceinline__ int golden_step(GoldenState &s, double &split, double &fsplit) {
  constexpr double phi = 0.6180339887498949;
  if (s.stage < 2) { s.which = s.stage; s.pending = s.stage ? s.r : s.p; return 0; }
  if (s.it >= kTroughIters) { split = NAN; return 1; }
  if (s.fp <= 0.) { split = s.p; fsplit = s.sign * s.fp; return 1; }
  if (s.fr <= 0.) { split = s.r; fsplit = s.sign * s.fr; return 1; }
  // A dip that does not deepen like log|c - c*| is not a collapsing resonance.
  if (s.dip && s.it >= kDipStopIter && s.kmin > s.sc - kDipStopDrop) { split = NAN; return 1; }
  if (s.kp < s.kr) {
    s.hi = s.r; s.r = s.p; s.fr = s.fp; s.kr = s.kp;
    s.p = s.hi - phi * (s.hi - s.lo);
    s.which = 0; s.pending = s.p;
  } else {
    s.lo = s.p; s.p = s.r; s.fp = s.fr; s.kp = s.kr;
    s.r = s.lo + phi * (s.hi - s.lo);
    s.which = 1; s.pending = s.r;
  }
  ++s.it;
  return 0;
}

__device__ __forceinline__ void golden_take(GoldenState &s, double x, double f, float key) {
  if (s.which == 0) { s.p = x; s.fp = f; s.kp = key; } else { s.r = x; s.fr = f; s.kr = key; }
  if (s.stage < 2) {
    if (++s.stage == 2) s.kmin = fminf(s.kmin, fminf(s.kp, s.kr));
  } else {
    s.kmin = fminf(s.kmin, key);
  }
}

// Event::probe layout: bits 0-7 the probe kind; bit 8 set when the FP32 pass
// handed the search to the FP64 pass; bits 16-23 the iterations done by then.
// While handed over, Event::x/vx hold the current bracket and Event::pad the
// best key (float bits), until the FP64 pass writes the split.
constexpr int kProbeNeeds64 = 0x100;

// Probes of same-sign troughs and scale dips, one search per lane as a state
// machine. T = float: the FP32 pass of the FP32 path -- when a point falls
// inside the FP32 uncertainty band (or the cutoff guard, or the bracket has
// narrowed to the FP32 spacing of k) the search is handed, with its bracket,
// iteration count and best key, to the FP64 pass. T = double: every probe of
// the FP64 path, or the handed-over probes of the FP32 path, in FP64; a
// resumed search restarts the golden section on the handed-over bracket (one
// extra evaluation: the retained interior point sits at a golden point of it).
//
// Work order: the probe list is cut into runs of 64 consecutive entries (the
// events of neighbouring traces, in ascending c); a wave takes the runs wave,
// wave + waves, ... as one stream and hands its entries out in order to
// whichever lanes are idle (ballot + popcount prefix), so the searches in
// flight in a wave stay within about one run: similar traces and speeds, so
// the branches of the wave's evaluations mostly agree.
template <class T, int Wave>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void trough_kernel(PipeArgs a) {
  constexpr bool fp32 = sizeof(T) == 4;
  // The FP64 pass of the FP32 path takes the compact list of handed-over probes.
  const bool from_list = !fp32 && a.fp32;
  const int n = min(from_list ? *a.handoff_count : *a.trough_count, a.queue_cap);
  const int lane = threadIdx.x & 63;
  const int wave_id = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  // Runs of 64 entries, dealt round-robin over the waves: at least four runs
  // per wave where the list is long enough to keep every resident wave busy
  // (a wave that holds a longer stream refills its lanes instead of ending
  // each run with a tail), and one wave per run below that.
  const int nruns = (n + 63) / 64;
  const int nwaves =
      min(static_cast<int>(gridDim.x) * (kBlock / 64), max(a.probe_waves[fp32 ? 0 : 1], (nruns + 3) / 4));
  const int stream = wave_id < nruns && wave_id < nwaves ? ((nruns - 1 - wave_id) / nwaves + 1) * 64 : 0;
  int cursor = 0;  // wave-uniform: next entry of the wave's stream
  GoldenState s{};
  bool active = false;
  uint32_t g = 0;
  const PrepRow *rows = nullptr;
  int L = 0, top = 0;
  double omega = 0., iw2 = 0., guard = 0.;
  float tau = 0.f;
  unsigned evals = 0;
  // A finished search: its split (NaN: no pair), or, in the FP32 pass, an
  // opposite-sign point, which the FP64 pass must confirm. An FP32 point of
  // opposite sign can sit on the wrong side of a root that close (FP32 moves
  // a step of F by ~1e-7): the FP64 pass repeats the search on the current
  // bracket, so the split it keeps is an FP64 sign change. Rare, so one
  // atomic per lane.
  auto finish = [&](double split, double fsplit) {
    if (fp32 && split == split) {
      Event &x = a.events[g];
      x.x = s.lo;
      x.vx = s.hi;
      x.pad = __float_as_int(s.stage == 2 ? s.kmin : INFINITY);
      x.probe |= kProbeNeeds64 | (s.it << 16);
      const int slot = atomicAdd(a.handoff_count, 1);
      if (slot < a.queue_cap) a.handoff_list[slot] = g;
    } else {
      a.events[g].x = split;
      a.events[g].vx = fsplit;
    }
    active = false;
  };
  // Each pass: lanes with an evaluated point take a step; idle lanes are
  // handed the next entries (and take steps) until every lane needs one
  // evaluation of F or the stream is done; then the wave evaluates together.
  // Every loop condition is wave-uniform.
  for (;;) {
    if (active) {
      double split, fsplit = 0.;
      if (golden_step(s, split, fsplit)) finish(split, fsplit);
    }
    for (;;) {
      const uint64_t idle = __ballot(!active);
      if (!idle || cursor >= stream) break;
      if (!active) {
        const int v = cursor + __popcll(idle & lanes_below(lane));
        const int q = (wave_id + (v >> 6) * nwaves) * 64 + (v & 63);
        if (v < stream && q < n) {
          g = from_list ? a.handoff_list[q] : a.trough_queue[q];
          const int tl = g / kEventCap, e = g % kEventCap;
          // The event, the node fields of the slots beside it (in bounds of
          // the trace's slots; one at or past nev is ignored), the trace's
          // event count and its TraceInfo are read together.
          const Event *ev = a.events + static_cast<size_t>(tl) * kEventCap;
          const Event x = ev[e];
          NodeFields xl, xr;
          neighbour_fields(ev, e, xl, xr);
          const int nev = a.scan[tl].nev;
          const TraceInfo ti = a.traces[tl + a.trace_begin];
          const int kind = x.probe & 0xff;
          const bool resumed = !fp32 && a.fp32;
          if ((kind == kTrough || kind == kDip) && !(resumed && !(x.probe & kProbeNeeds64))) {
            // Search only the neighbouring cells whose end values share the node's sign.
            const NodeValues nv = node_values(x, e, nev, xl, xr);
            const bool left = x.c0 == x.c0 && signbit(nv.l) == signbit(nv.c);
            const bool right = x.c2 == x.c2 && signbit(nv.r) == signbit(nv.c);
            if (left || right) {
              const ModelInfo mi = a.models[ti.model];
              rows = a.rows + mi.row0;
              L = mi.L;
              top = mi.top;
              omega = ti.omega;
              iw2 = 1. / (omega * omega);
              guard = mi.cutoff_guard;
              tau = ti.tau;
              const double sign = signbit(nv.c) ? -1. : 1.;
              if (resumed) {
                golden_init(s, x.x, x.vx, sign, kind == kDip, x.sc);
                s.it = (x.probe >> 16) & 0xff;
                s.kmin = __int_as_float(x.pad);
              } else {
                golden_init(s, left ? x.c0 : x.c1, right ? x.c2 : x.c1, sign, kind == kDip, x.sc);
              }
              active = true;
              double split, fsplit = 0.;
              if (golden_step(s, split, fsplit)) finish(split, fsplit);
            }
          }
        }
      }
      cursor += __popcll(idle);
    }
    if (!__ballot(active)) break;
    if (!active) continue;
    ++evals;
    float ls;
    if constexpr (fp32) {
      // FP32 only while the interval is far wider than the FP32 spacing of k.
      float k32;
      const double cs = snap32(omega, s.pending, s.lo, s.hi, &k32);
      bool handed = !(cs < guard && s.hi - s.lo > 2e-6 * s.hi && cs > s.lo && cs < s.hi);
      if (!handed) {
        const float f = secular_global<float>(rows, L, top, Wave, omega, iw2, k32, &ls);
        handed = !(fabsf(f) >= tau);
        if (!handed) {
          const double sf = s.sign * double(f);
          golden_take(s, cs, sf, s.dip ? ls : float(sf));
        }
      }
      if (handed) {
        Event &x = a.events[g];
        x.x = s.lo;
        x.vx = s.hi;
        x.pad = __float_as_int(s.stage == 2 ? s.kmin : INFINITY);
        x.probe |= kProbeNeeds64 | (s.it << 16);
        active = false;
      }
      // Append the handed-over probes to the FP64 pass's list (one atomic per
      // wave; the lanes evaluating here are the wave's active lanes).
      const uint64_t hm = __ballot(handed);
      if (hm) {
        const int leader = __ffsll(static_cast<unsigned long long>(hm)) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(a.handoff_count, __popcll(hm));
        base = __shfl(base, leader, 64);
        if (handed) {
          const int slot = base + __popcll(hm & lanes_below(lane));
          if (slot < a.queue_cap) a.handoff_list[slot] = g;
        }
      }
    } else {
      const double f = s.sign * secular_rows(rows, L, top, Wave, omega, s.pending, &ls);
      golden_take(s, s.pending, f, s.dip ? ls : float(f));
    }
  }
  if (a.stats) wave_count(a.stats + (fp32 ? 7 : 8), evals);
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
    auto emit = [&](double lo, double hi, double flo, double fhi, float slo = NAN, float shi = NAN) {
      if (nb < K) out[nb] = Bracket{lo, hi, flo, fhi, slo, shi};
      ++nb;
    };
    int last_split = -2;  // left node of the cell that holds the last split pair
    // Both roots of a split pair lie in the same-sign cell [lo, hi] that holds the split.
    auto emit_pair = [&](const Event &x, int cell, double lo, double hi, double flo, double fhi) {
      if (cell == last_split || !(x.x > lo && x.x < hi) || signbit(flo) != signbit(fhi) ||
          signbit(x.vx) == signbit(flo))
        return;
      last_split = cell;
      emit(lo, x.x, flo, x.vx);
      if (nb < K) emit(x.x, hi, x.vx, fhi);
    };
    // A rolling window over the events: each event is read once, and the
    // node fields of event e - 1 are kept from the previous step.
    const int nev = ts.nev;
    Event x{}, y{};
    if (nev > 0) x = ev[0];
    NodeFields prev{-2, kNone, 0.};
    for (int e = 0; e < nev && nb < K; ++e, x = y) {
      if (e + 1 < nev) y = ev[e + 1];
      const NodeValues v = node_values(x, e, nev, prev, node_fields(y));
      prev = node_fields(x);
      const bool split = x.probe && x.x == x.x, split_left = split && x.x < x.c1;
      // Left cell [c0, c1]: an uncertain node owns it when that neighbour is certain.
      if (split_left) emit_pair(x, x.node - 1, x.c0, x.c1, v.l, v.c);
      else if (x.kind == kUncertain && !v.lu && x.c0 == x.c0 && signbit(v.l) != signbit(v.c))
        emit(x.c0, x.c1, v.l, v.c, x.sl, x.sc);
      if (nb >= K) break;
      // Right cell [c1, c2]: a certain sign change, or an uncertain node's.
      if (split && !split_left) emit_pair(x, x.node, x.c1, x.c2, v.c, v.r);
      else if (x.kind == kBracket || (x.kind == kUncertain && x.c2 == x.c2 && signbit(v.c) != signbit(v.r)))
        emit(x.c1, x.c2, v.c, v.r, x.sc, x.sr);
    }
  }
  const int count = nb < K ? nb : K;
  ts.count = (flags & kTraceOverflow) ? -1 : (ts.status ? 0 : count);
}

// ---------------------------------------------------------------------------
// Love waves: exact root count from the unwrapped SH phase (reference
// cpp/common/love_phase.hpp:9-36). Roots are the c where the phase crosses
// n*pi; the phase decreases with c.
// ---------------------------------------------------------------------------
// vs_half: the half-space shear speed, so that the half-space wavenumber is
// omega / vs rounded once, as the reference forms it: whether k > khalf holds
// at the top of the window (nextafter(vs_half, 0)) decides if the phase exists.
__device__ double love_phase_dev(const PrepRow *rows, int L, int top, double omega, double c, double vs_half) {
  constexpr double pi = 3.141592653589793;
  const int last = L - 1;
  const double k = omega / c, khalf = omega / vs_half;
  if (!(k > khalf)) return NAN;
  double scale = 0.5 * rows[last].rho * rows[last].g2 * sqrt((k + khalf) * (k - khalf));
  double theta = .25 * pi;
  // prev + the representative of raw - prev in [-pi, pi]: the same branch as
  // atan2(sin(raw - prev), cos(raw - prev)), by a two-term 2 pi reduction.
  auto unwrap = [](double raw, double prev) {
    const double d = raw - prev, n = rint(d * 0.15915494309189535);
    return prev + fma(-n, 2.4492935982947064e-16, fma(-n, 6.283185307179586, d));
  };
  for (int j = last - 1; j >= top; --j) {
    const PrepRow r = rows[j];
    if (r.h == 0.) continue;
    const double mu = 0.5 * r.rho * r.g2, kb = omega * r.ivs;
    const double q = sqrt(fabs((k + kb) * (k - kb))), next = mu * (q > 0. ? q : 1. / r.h);
    if (!(scale > 0.) || !(next > 0.) || !(fabs(scale / next) < INFINITY)) return NAN;
    double sn, cs;
    sincos_phase(theta, &sn, &cs);
    theta = unwrap(atan2((scale / next) * sn, cs), theta);
    scale = next;
    if (k < kb) theta -= q * r.h;
    else {
      sincos_phase(theta, &sn, &cs);
      double t, u;
      if (q == 0.) { t = sn; u = cs + sn * r.h * scale / mu; }
      else {
        const double ex = exp(-2. * q * r.h), ch = .5 * (1. + ex), sh = .5 * (1. - ex);
        t = sn * ch + cs * sh;
        u = cs * ch + sn * sh;
      }
      if (fmax(fabs(t), fabs(u)) < 1e-14) return NAN;
      theta = unwrap(atan2(t, u), theta);
    }
    if (!(fabs(theta) <= 1048576.)) return NAN;
  }
  return theta;
}

// Number of roots in [lo, c]: multiples of pi in (phase(c), phase(lo)].
__device__ inline int love_roots_between(double plo, double pc) {
  constexpr double pi = 3.141592653589793;
  const double top = floor(plo / pi);
  const double n = top - floor(pc / pi);
  return n > 0. ? int(n) : 0;
}

// One lane per Love trace, right after the scan (and the FP64 rescans): the
// exact phase count of the trace's window, where the phase exists at both
// ends. Such a trace is verified by the count after assembly, so it takes no
// pair probes.
__global__ __launch_bounds__(kBlock) void love_range_kernel(PipeArgs a) {
  const int tl = blockIdx.x * kBlock + threadIdx.x;
  const int t = tl + a.trace_begin;
  if (t >= a.trace_end) return;
  TraceScan &ts = a.scan[tl];
  if (ts.status || (ts.flags & kTraceOverflow)) return;
  const TraceInfo ti = a.traces[t];
  const ModelInfo mi = a.models[ti.model];
  const PrepRow *rows = a.rows + mi.row0;
  const double pl = love_phase_dev(rows, mi.L, mi.top, ti.omega, mi.lo, mi.vs_half);
  const double ph = love_phase_dev(rows, mi.L, mi.top, ti.omega, mi.hi, mi.vs_half);
  if (!(pl == pl) || !(ph == ph) || ph > pl) return;
  ts.exact = love_roots_between(pl, ph);
  ts.phase_lo = pl;
  ts.flags |= kScanPhaseReady;
}

// One lane per Love trace: compares the assembled brackets with the exact
// phase count and queues the trace for a phase search when they disagree.
__global__ __launch_bounds__(kBlock) void love_count_kernel(PipeArgs a) {
  const int tl = blockIdx.x * kBlock + threadIdx.x;
  const int t = tl + a.trace_begin;
  if (t >= a.trace_end) return;
  TraceScan &ts = a.scan[tl];
  if (ts.status || (ts.flags & kTraceOverflow)) return;
  if (!(ts.flags & kScanPhaseReady)) {  // no phase at a window end (love_range_kernel)
    ts.flags |= kFlagLoveFallback;
    return;
  }
  const TraceInfo ti = a.traces[t];
  const ModelInfo mi = a.models[ti.model];
  const PrepRow *rows = a.rows + mi.row0;
  const double pl = ts.phase_lo;
  const int exact = ts.exact, K = mi.K;
  const int want = exact < K ? exact : K;
  bool ok = ts.count == want;
  if (ok && want == K && exact > K) {
    // Capped: the K-th bracket must end below the (K+1)-th phase crossing.
    const Bracket b = a.brackets[static_cast<size_t>(tl) * a.kmax + K - 1];
    const double pb = love_phase_dev(rows, mi.L, mi.top, ti.omega, b.hi, mi.vs_half);
    ok = pb == pb && love_roots_between(pl, pb) == K;
  }
  if (ok) {
    ts.flags |= kScanLovePhase;
    return;
  }
  ts.count = 0;
  ts.flags |= kScanLovePhase;
  a.phase_list[atomicAdd(a.phase_count, 1)] = tl;
}

// Phase search of the queued Love traces (reference love_phase.hpp:48-91),
// one lane per (trace, mode slot): bisection on the phase until the bracket
// holds a single crossing, n*pi, with opposite signs of F at its ends.
__global__ __launch_bounds__(kBlock) void love_phase_kernel(PipeArgs a) {
  constexpr double pi = 3.141592653589793;
  const int n = *a.phase_count;
  for (int q = blockIdx.x * kBlock + threadIdx.x; q < n * a.kmax; q += gridDim.x * kBlock) {
    const int tl = a.phase_list[q / a.kmax], m = q % a.kmax;
    const TraceInfo ti = a.traces[tl + a.trace_begin];
    const ModelInfo mi = a.models[ti.model];
    if (m >= mi.K) continue;
    const PrepRow *rows = a.rows + mi.row0;
    double lo = mi.lo, hi = mi.hi;
    double pa = love_phase_dev(rows, mi.L, mi.top, ti.omega, lo, mi.vs_half), pb = love_phase_dev(rows, mi.L, mi.top, ti.omega, hi, mi.vs_half);
    const int exact = love_roots_between(pa, pb), want = exact < mi.K ? exact : mi.K;
    if (m == 0) a.scan[tl].count = want;
    if (m >= want) continue;
    const double level = (floor(pa / pi) - double(m)) * pi;
    Bracket br{NAN, NAN, 0., 0., NAN, NAN};
    double fa = 0., fb = 0.;
    bool have_f = false;
    for (int it = 0; it < 200; ++it) {
      if (pa < level + pi && pb > level - pi) {
        fa = secular_rows(rows, mi.L, mi.top, 1, ti.omega, lo);
        fb = secular_rows(rows, mi.L, mi.top, 1, ti.omega, hi);
        if (signbit(fa) != signbit(fb)) { have_f = true; break; }
      }
      const double mid = lo + .5 * (hi - lo);
      if (!(mid > lo && mid < hi)) break;
      const double p = love_phase_dev(rows, mi.L, mi.top, ti.omega, mid, mi.vs_half);
      if (!(p == p)) break;
      if (p > level) { lo = mid; pa = p; } else { hi = mid; pb = p; }
    }
    if (have_f) br = Bracket{lo, hi, fa, fb, NAN, NAN};
    a.brackets[static_cast<size_t>(tl) * a.kmax + m] = br;
  }
}

// ---------------------------------------------------------------------------
// FP64 bracket refinement, one lane per (trace, root) slot.
// ---------------------------------------------------------------------------
// Bracket refinement as a per-lane state machine, by the solver's method
// (reference cpp/common/refinement.hpp:11-65 and physics.hpp:224-237, the same
// steps and stop rules): 0 regula falsi with bisection every third step and
// near the ends (legacy), 1 Brent-Dekker (default), 2 Chandrupatla. A lane
// whose root converges takes its next bracket at once, so every FP64 secular
// evaluation runs with nearly all lanes busy instead of waiting for the
// wave's slowest bracket. refine_step returns 0 when F at refine_point is
// needed (refine_store keeps it), 1 when `root` is final.
// Brent: a, b, c (b the best point, c opposite), d, e the last steps.
// Chandrupatla: a = x1, b = x2 (the bracket), c = x3 (the displaced end), d = t.
// Legacy: [a, b] the bracket.
struct RefineState {
  double a, b, c, fa, fb, fc, d, e, stop, x, fx;
  int iter;
  bool fresh;  // fx holds F(x), not yet used
};

__device__ __forceinline__ double refine_stop(double a, double b, double tol) {
  const double mag = fmax(fabs(a), fabs(b));
  return fmax(16. * 2.220446049250313e-16 * mag, fmin((b - a) / 64., .125 * tol * mag));
}

__device__ __forceinline__ void refine_init(RefineState &s, int method, double lo, double hi, double flo, double fhi,
                                            double tol) {
  s.a = lo; s.b = hi; s.fa = flo; s.fb = fhi;
  s.c = hi; s.fc = fhi;
  s.d = method == 2 ? .5 : hi - lo;
  s.e = s.d;
  s.stop = refine_stop(lo, hi, tol);
  s.x = hi; s.fx = 0.;
  s.iter = 0;
  s.fresh = false;
}

// The root from a converged bracket [a, b] with end values of opposite sign:
// the secant point, which inside a bracket this narrow (the refinement stop
// width) is the root to rounding, so that quantities steep in c (the group
// velocity near a body-wave speed) are evaluated at the root itself; the
// midpoint where the secant is unusable.
__device__ __forceinline__ double bracket_root(double a, double b, double fa, double fb) {
  const double x = a - fa * ((b - a) / (fb - fa));
  return x >= fmin(a, b) && x <= fmax(a, b) ? x : .5 * (a + b);
}

__device__ __forceinline__ void refine_take(RefineState &s, double f) {
  s.fx = f;
  s.fresh = true;
}

// Brent evaluates at s.b; refine_point / refine_store route its value to s.fb.
__device__ __forceinline__ int brent_step(RefineState &s, double &root) {
  if ((s.fb < 0.) == (s.fc < 0.)) { s.c = s.a; s.fc = s.fa; s.e = s.d = s.b - s.a; }
  if (fabs(s.fc) < fabs(s.fb)) { s.a = s.b; s.fa = s.fb; s.b = s.c; s.fb = s.fc; s.c = s.a; s.fc = s.fa; }
  if (s.fb == 0.) { root = s.b; return 1; }
  const double mid = .5 * (s.c - s.b), step = .5 * s.stop;
  if (fabs(s.c - s.b) <= s.stop) { root = bracket_root(s.b, s.c, s.fb, s.fc); return 1; }
  if (s.iter >= 160) { root = NAN; return 1; }
  if (fabs(s.e) >= step && fabs(s.fa) > fabs(s.fb)) {
    const double sr = s.fb / s.fa;
    double pp, qq;
    if (s.a == s.c) { pp = 2. * mid * sr; qq = 1. - sr; }
    else {
      const double q0 = s.fa / s.fc, r = s.fb / s.fc;
      pp = sr * (2. * mid * q0 * (q0 - r) - (s.b - s.a) * (r - 1.));
      qq = (q0 - 1.) * (r - 1.) * (sr - 1.);
    }
    if (pp > 0.) qq = -qq; else pp = -pp;
    const double old = s.e;
    s.e = s.d;
    const double ratio = pp / qq;
    if (ratio == ratio && fabs(ratio) < INFINITY && 2. * pp < fmin(3. * mid * qq - fabs(step * qq), fabs(old * qq)))
      s.d = ratio;
    else { s.d = mid; s.e = mid; }
  } else { s.d = mid; s.e = mid; }
  s.a = s.b;
  s.fa = s.fb;
  s.b += fabs(s.d) > step ? s.d : (mid > 0. ? step : -step);
  ++s.iter;
  return 0;
}

__device__ __forceinline__ int chandrupatla_step(RefineState &s, double &root) {
  if (s.fresh) {
    s.fresh = false;
    const double f = s.fx;
    if (f == 0.) { root = s.x; return 1; }
    s.c = s.b; s.fc = s.fb;
    if ((f < 0.) == (s.fa < 0.)) { s.c = s.a; s.fc = s.fa; } else { s.b = s.a; s.fb = s.fa; }
    s.a = s.x; s.fa = f;
    const double dx = fabs(s.b - s.a);
    if (dx <= s.stop) { root = bracket_root(s.a, s.b, s.fa, s.fb); return 1; }
    // Inverse quadratic step when the three points admit it (Chandrupatla's test).
    const double xi = (s.a - s.b) / (s.c - s.b), phi = (s.fa - s.fb) / (s.fc - s.fb);
    double t = .5;
    if (xi > 0. && xi < 1. && (1. - sqrt(1. - xi)) < phi && phi < sqrt(xi)) {
      const double alpha = (s.c - s.a) / (s.b - s.a);
      const double z = s.fa / (s.fa - s.fb) * s.fc / (s.fc - s.fb) - alpha * s.fa / (s.fc - s.fa) * s.fb / (s.fb - s.fc);
      if (z == z && fabs(z) < INFINITY) t = z;
    }
    const double limit = fmin(.5, .5 * s.stop / dx);
    s.d = fmax(limit, fmin(1. - limit, t));
  }
  if (fabs(s.b - s.a) <= s.stop) { root = bracket_root(s.a, s.b, s.fa, s.fb); return 1; }
  if (s.iter >= 160) { root = NAN; return 1; }
  double x = s.a + s.d * (s.b - s.a);
  if (!(x > fmin(s.a, s.b) && x < fmax(s.a, s.b))) x = .5 * (s.a + s.b);
  s.x = x;
  ++s.iter;
  return 0;
}

__device__ __forceinline__ int legacy_step(RefineState &s, double &root) {
  if (s.fresh) {
    s.fresh = false;
    const double f = s.fx;
    if (f == 0.) { root = s.x; return 1; }
    if ((f < 0.) == (s.fa < 0.)) { s.a = s.x; s.fa = f; } else { s.b = s.x; s.fb = f; }
  }
  const double w = s.b - s.a;
  if (!(s.iter < 100 && w > s.stop)) { root = w <= s.stop ? bracket_root(s.a, s.b, s.fa, s.fb) : NAN; return 1; }
  double x = s.a + w * (fabs(s.fa) / (fabs(s.fa) + fabs(s.fb)));
  if (s.iter % 3 == 2 || !(fabs(x) < INFINITY) || x <= s.a + .1 * w || x >= s.b - .1 * w) x = .5 * (s.a + s.b);
  s.x = x;
  ++s.iter;
  return 0;
}

// The pending evaluation point, and where its value goes.
__device__ __forceinline__ double refine_point(const RefineState &s, int method) { return method == 1 ? s.b : s.x; }
__device__ __forceinline__ void refine_store(RefineState &s, int method, double f) {
  if (method == 1) s.fb = f;
  else refine_take(s, f);
}

__device__ __forceinline__ int refine_step(RefineState &s, int method, double &root) {
  if (method == 2) return chandrupatla_step(s, root);
  if (method == 0) return legacy_step(s, root);
  return brent_step(s, root);
}

// The same refinement run serially by one lane (continuation brackets).
__device__ double refine_serial(int method, const PrepRow *rows, int L, int top, int wave, double omega, double a,
                                double b, double fa, double fb, double tol, int &evals) {
  if (method != 0) {
    if (fa == 0.) return a;
    if (fb == 0.) return b;
  } else if (a == b) {
    return a;
  }
  RefineState s;
  refine_init(s, method, a, b, fa, fb, tol);
  double root;
  while (!refine_step(s, method, root)) {
    const double f = secular_rows(rows, L, top, wave, omega, refine_point(s, method));
    ++evals;
    if (!(f == f) || fabs(f) == INFINITY) return NAN;
    refine_store(s, method, f);
  }
  return root;
}

// f 2^d with d clamped to +-600 bits, never rounding a nonzero f to zero (a
// zero ends the refinement as an exact root).
__device__ __forceinline__ double unscale(double f, double d) {
  const double g = f * exp2(fmin(600., fmax(-600., d)));
  return g == 0. && f != 0. ? copysign(4.9406564584124654e-324, f) : g;
}
// f mant 2^(ex - ref) for a scale product: the integer exponent by ldexp,
// the fraction of ref by one FP32 exp2.
__device__ __forceinline__ double unscale(double f, const ScaleProduct &p, float ref) {
  const float ri = floorf(ref);
  const int e = min(600, max(-600, p.ex - int(ri)));
  const double g = ldexp(f * double(p.mant * exp2f(ri - ref)), e);
  return g == 0. && f != 0. ? copysign(4.9406564584124654e-324, f) : g;
}

// A guard root (beyond the reported modes) serves only the close-root tests
// with its neighbours (trace flag, group-velocity CLOSE and window). When the
// brackets alone decide both tests -- r_m - r_{m-1} <= close_rtol r_m is
// certainly true or certainly false for any roots inside the two brackets,
// with a margin far above rounding -- any point of its bracket gives the
// same flags and windows as the refined root, so it is not refined (unless
// every root is wanted, as the anchors of a continuation are).
__device__ __forceinline__ bool close_decided(double lo1, double hi1, double lo2, double hi2, double rtol) {
  const double far = (1. - rtol) * lo2 - hi1, near = (1. - rtol) * hi2 - lo1;
  return far > 1e-9 * lo2 || near < -1e-9 * hi2;
}
__device__ __forceinline__ bool guard_decided(const PipeArgs &a, size_t q, int m, int count, const Bracket &br) {
  const Bracket prev = a.brackets[q - 1];
  if (!close_decided(prev.lo, prev.hi, br.lo, br.hi, a.close_rtol)) return false;
  if (m + 1 >= count) return true;
  const Bracket next = a.brackets[q + 1];
  return close_decided(br.lo, br.hi, next.lo, next.hi, a.close_rtol);
}

// Refinement work order. A tile is kRefineTile consecutive traces of the
// chunk (the next periods of one or a few models); its items are its (mode,
// trace) pairs, mode-major, so that 64 consecutive items are one mode of 64
// neighbouring traces: roots that sit in the same propagating/evanescent
// regime of every layer, so the branches of a wave's evaluation agree. A
// wave takes the tiles wave, wave + waves, ... one at a time and hands a
// tile's items out in order to whichever lanes are idle (ballot + popcount
// prefix). The rows of the tile's models are staged in the wave's LDS when
// they fit (kRefineRows), else read from global memory.
constexpr int kRefineTile = 64;
template <int Wave>
struct RefineRows {
  static constexpr int value = Wave == 0 ? 64 : 32;  // Love: a smaller LDS share keeps 4 blocks per CU
};

// One instantiation per method and wave type (both fixed for a launch): the
// Love instantiation is not held to the Rayleigh evaluation's registers.
template <int Method, int Wave>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void refine_kernel(PipeArgs a) {
  constexpr int kRows = RefineRows<Wave>::value;
  const int ntl = a.trace_end - a.trace_begin;
  const int lane = threadIdx.x & 63, wib = threadIdx.x >> 6;
  const int wave_id = blockIdx.x * (kBlock / 64) + wib;
  const int nwaves = gridDim.x * (kBlock / 64);
  const int ntiles = (ntl + kRefineTile - 1) / kRefineTile;
  const int tile_items = a.kmax * kRefineTile;
  // The lane's state lives in LDS, so that the evaluation of F has the
  // registers to itself.
  __shared__ RefineState state[kBlock];
  __shared__ PrepRow staged_rows[kBlock / 64][kRows];
  RefineState &s = state[threadIdx.x];
  s = RefineState{};
  unsigned evals = 0;
  for (int tile = wave_id; tile < ntiles; tile += nwaves) {
    // The tile's rows: the span of the models of its traces that hold brackets.
    const int t0 = tile * kRefineTile;
    int lo = INT_MAX, hi = INT_MIN;
    if (t0 + lane < ntl && a.scan[t0 + lane].count > 0) {
      const ModelInfo &mi = a.models[a.traces[t0 + lane + a.trace_begin].model];
      lo = mi.row0;
      hi = mi.row0 + mi.L;
    }
    for (int o = 32; o > 0; o >>= 1) {
      lo = min(lo, __shfl_xor(lo, o, 64));
      hi = max(hi, __shfl_xor(hi, o, 64));
    }
    // Wave-uniform. A tile with no brackets keeps lo = INT_MAX > hi = INT_MIN
    // and stages nothing (hi - lo would overflow, so lo <= hi is tested first).
    const bool staged = lo <= hi && hi - lo <= kRows;
    wave_sync();  // the previous tile's evaluations have read the staged rows
    if (staged && lane < hi - lo) staged_rows[wib][lane] = a.rows[lo + lane];
    wave_sync();
    bool active = false;
    size_t q = 0;  // the lane's item: bracket / root index tl * kmax + m
    int row0 = 0, L = 0, top = 0;
    double omega = 0.;
    float ref = NAN;  // reference log2 scale of the unnormalized F; NaN: refine F itself
    int cursor = 0;   // wave-uniform: next item of the tile
    // Each pass: lanes with an evaluated point take a step; idle lanes are
    // handed the next items (and take steps) until every lane needs one
    // evaluation of F or the tile is done; then the wave evaluates together.
    // Every loop condition is wave-uniform.
    for (;;) {
      if (active) {
        double root;
        if (refine_step(s, Method, root)) {
          a.roots[q] = root;
          active = false;
        }
      }
      for (;;) {
        const uint64_t idle = __ballot(!active);
        if (!idle || cursor >= tile_items) break;
        if (!active) {
          const int v = cursor + __popcll(idle & lanes_below(lane));
          const int m = v / kRefineTile, tl = t0 + (v - m * kRefineTile);
          if (v < tile_items && tl < ntl) {
            q = static_cast<size_t>(tl) * a.kmax + m;
            const int count = a.scan[tl].count;
            const TraceInfo ti = a.traces[tl + a.trace_begin];
            const ModelInfo mi = a.models[ti.model];
            const Bracket br = m < count && m < mi.K ? a.brackets[q] : Bracket{NAN, NAN, 0., 0., NAN, NAN};
            if (!(br.lo == br.lo)) a.roots[q] = NAN;
            else if (br.flo == 0.) a.roots[q] = br.lo;
            else if (br.fhi == 0.) a.roots[q] = br.hi;
            else if (m >= mi.modes && !a.all_roots && guard_decided(a, q, m, min(count, mi.K), br))
              a.roots[q] = bracket_root(br.lo, br.hi, br.flo, br.fhi);
            else {
              // Guard roots beyond the requested modes only feed the close-root test.
              const double tol = m < mi.modes ? a.tol : fmax(a.tol, kGuardTol);
              // With both end scales known, refine G = F 2^(ls - ref): the same
              // zeros (a positive factor), and smooth where the normalized F of a
              // deep evanescent stack is a +-1 step that defeats interpolation.
              const bool scaled = fabsf(br.slo) < INFINITY && fabsf(br.shi) < INFINITY;
              ref = scaled ? fmaxf(br.slo, br.shi) : NAN;
              const double glo = scaled ? unscale(br.flo, double(br.slo) - double(ref)) : br.flo;
              const double ghi = scaled ? unscale(br.fhi, double(br.shi) - double(ref)) : br.fhi;
              refine_init(s, Method, br.lo, br.hi, glo, ghi, tol);
              row0 = mi.row0;
              L = mi.L;
              top = mi.top;
              omega = ti.omega;
              double root;
              if (refine_step(s, Method, root)) a.roots[q] = root;
              else active = true;
            }
          }
        }
        cursor += __popcll(idle);
      }
      if (!__ballot(active)) break;
      if (!active) continue;
      ++evals;
      const double x = refine_point(s, Method);
      __asm__ volatile("" ::: "memory");  // keep no copy of the state in registers across the evaluation
      ScaleProduct sp;
      const PrepRow *rows = staged ? &staged_rows[wib][row0 - lo] : a.rows + row0;
      double f = secular_rows(rows, L, top, Wave, omega, x, nullptr, &sp);
      if (ref == ref) f = unscale(f, sp, ref);
      __asm__ volatile("" ::: "memory");  // reload the state from LDS after the evaluation
      if (!(f == f) || fabs(f) == INFINITY) {
        a.roots[q] = NAN;
        active = false;
      } else {
        refine_store(s, Method, f);
      }
    }
  }
  if (a.stats) wave_count(a.stats + 6, evals);
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
  TraceScan &ts = a.scan[tl];
  const double *roots = a.roots + static_cast<size_t>(tl) * a.kmax;
  unsigned flags = 0, rejected = 0;
  int count = ts.count, level = a.level_grid;
  const bool phase = (ts.flags & kScanLovePhase) != 0;
  if (ts.status) { level = ts.status == 1 ? -4 : -3; count = 0; }
  else flags |= a.trace_extra;
  if (phase) level = kLevelLovePhase;
  flags |= ts.flags & kFlagLoveFallback;
  if (ts.flags & kTraceOverflow) { flags |= kFlagUnconverged; level = -1; count = -1; }
  bool cutoff = false, close = false;
  for (int m = 0; m < count && m < mi.K; ++m) {
    const double r = roots[m];
    if (!(r == r)) { flags |= kFlagRefineFailed; continue; }
    if (m > 0 && roots[m - 1] == roots[m - 1] && r - roots[m - 1] <= a.close_rtol * r) close = true;
  }
  for (int m = 0; m < mi.modes; ++m) {
    double r = m < count ? roots[m] : NAN;
    if (r == r && r >= .99 * mi.vs_half) cutoff = true;
    // Optional acceptance test of each reported root on |F| there (reference
    // cpp/cpu/solver.cpp:333): a root that fails it is reported as NaN.
    if (r == r && a.f_accept < 1.) {
      const double v = secular_rows(a.rows + mi.row0, mi.L, mi.top, a.wave, ti.omega, r);
      if (!(fabs(v) <= a.f_accept)) { r = NAN; flags |= kFlagRejected; rejected |= 1u << m; }
    }
    const size_t cell = static_cast<size_t>(mi.out0) + static_cast<size_t>(m) * mi.T + ti.f;
    if (a.velocity64) a.velocity64[cell] = r;
    if (a.velocity32) a.velocity32[cell] = float(r);
  }
  if (a.all_roots) {
    double *dst = a.all_roots + static_cast<size_t>(t) * a.kmax;
    for (int m = 0; m < a.kmax; ++m) dst[m] = m < count ? roots[m] : NAN;
  }
  ts.rejected = rejected;
  if (cutoff) flags |= kFlagNearCutoff;
  if (close && (a.close_grid || phase)) flags |= kFlagCloseRoot;
  const size_t tc = static_cast<size_t>(mi.trace_out0) + ti.f;
  a.flags[tc] = flags;
  a.levels[tc] = level;
  a.counts[tc] = count;
}

// ---------------------------------------------------------------------------
// Analytic group velocity U = dw/dk = -F_k / F_w at each refined root.
//
// GroupDual: two-direction dual numbers (d/dk, d/dw) of their own, so that the
// group kernel's wave factor can take the narrow series region
// (wave_factor_held<GroupDual, true>): lanes of a wave (different traces and
// modes) then rarely straddle the series and closed-form branches.
// ---------------------------------------------------------------------------
struct GroupDual {
  double v, a, b;
  __device__ GroupDual() : v(0.), a(0.), b(0.) {}
  __device__ GroupDual(double x) : v(x), a(0.), b(0.) {}
  __device__ GroupDual(double x, double da, double db) : v(x), a(da), b(db) {}
};
using GD = GroupDual;
__device__ __forceinline__ GD operator+(GD x, GD y) { return {x.v + y.v, x.a + y.a, x.b + y.b}; }
__device__ __forceinline__ GD operator-(GD x, GD y) { return {x.v - y.v, x.a - y.a, x.b - y.b}; }
__device__ __forceinline__ GD operator-(GD x) { return {-x.v, -x.a, -x.b}; }
__device__ __forceinline__ GD operator*(GD x, GD y) { return {x.v * y.v, x.a * y.v + x.v * y.a, x.b * y.v + x.v * y.b}; }
__device__ __forceinline__ GD operator*(GD x, double s) { return {x.v * s, x.a * s, x.b * s}; }
__device__ __forceinline__ GD &operator*=(GD &x, double s) { x = x * s; return x; }
__device__ __forceinline__ bool operator<(GD x, GD y) { return x.v < y.v; }
__device__ __forceinline__ bool operator==(GD x, GD y) { return x.v == y.v; }
__device__ __forceinline__ GD m_sqrt(GD x) {
  const double r = sqrt(x.v), d = 0.5 / r;
  return {r, x.a * d, x.b * d};
}
__device__ __forceinline__ void m_sincos(GD x, GD *s, GD *c) {
  double sn, cs;
  m_sincos(x.v, &sn, &cs);
  *s = GD(sn, cs * x.a, cs * x.b);
  *c = GD(cs, -sn * x.a, -sn * x.b);
}
__device__ __forceinline__ GD m_rcp(GD x) {
  const double r = 1. / x.v, d = -r * r;
  return {r, x.a * d, x.b * d};
}
__device__ __forceinline__ GD m_fabs(GD x) { return x.v < 0. ? -x : x; }
__device__ __forceinline__ GD m_fma(GD x, GD y, GD z) { return x * y + z; }
__device__ __forceinline__ Wave<GD> wave_factor(GD k, GD kx, GD h) { return wave_factor_held<GD, true>(k, kx, h); }
template <int N>
__device__ __forceinline__ GD rescale(GD *e) {  // the scale is held constant
  double m = 0.;
#pragma unroll
  for (int i = 0; i < N; ++i) m = fmax(m, fabs(e[i].v));
  if (m > 1.0e-40) {
    const double inv = 1. / m;
#pragma unroll
    for (int i = 0; i < N; ++i) e[i] *= inv;
    return GD(m);
  }
  return GD(1.);
}

// dF/dk (fixed omega) into *fk and dF/domega (fixed k) into *fw at (omega, c).
__device__ __forceinline__ void secular_partials(const PrepRow *rows, int L, int top, int wave, double omega,
                                                 double c, double *fk, double *fw) {
  const double iw = 1. / omega, iw2 = iw * iw;
  const GD k(omega / c, 1., 0.), w(omega, 0., 1.);
  auto layer = [&](const PrepRow &r) {
    Layer<GD> l;
    l.h = GD(r.h);
    l.kb = GD(omega * r.ivs, 0., r.ivs);
    if (wave == 0) {
      l.ka = GD(omega * r.ivp, 0., r.ivp);
      l.gk = GD(r.g2 * iw2, 0., -2. * r.g2 * iw2 * iw);
      l.rho = GD(r.rho);
      l.irho = GD(r.irho);
      l.irho2 = GD(r.irho2);
    } else {
      l.rho = GD(0.5 * r.rho * r.g2);
      l.irho = GD(r.imu);
    }
    return l;
  };
  GD F;
  if (wave == 0) {
    const GD k2 = k * k;
    GD e[5];
    rayleigh_halfspace(layer(rows[L - 1]), k, k2, e);
    for (int j = L - 2; j >= top; --j) rayleigh_layer(layer(rows[j]), k, k2, e);
    F = top ? rayleigh_water_top(layer(rows[0]), k, e) : e[0];
  } else {
    const Layer<GD> hs = layer(rows[L - 1]);
    GD e[2] = {hs.rho * m_sqrt((k + hs.kb) * m_fabs(k - hs.kb)), GD(1.)};
    for (int j = L - 2; j >= top; --j) love_layer(layer(rows[j]), k, e);
    F = e[0];
  }
  (void)w;
  *fk = F.a;
  *fw = F.b;
}

__device__ __forceinline__ double group_at(const PrepRow *rows, int L, int top, int wave, double omega, double c) {
  double fk, fw;
  secular_partials(rows, L, top, wave, omega, c, &fk, &fw);
  return -fk / fw;
}

// Group velocity of reported root m of a trace with the reference's flag
// contract (cpp/common/group_velocity.hpp:9-12, 26-104; window and close test
// cpp/cpu/solver.cpp:96-103). In its terms A = dF/dln c = -k F_k and
// B = dF/dln omega = k F_k + omega F_w, U = c A / (A + B) = -F_k / F_w:
//   MISSING_ROOT   no root;
//   CLOSE_ROOT     a neighbouring root within close_rtol;
//   POLISH_FAILED  the root is not strictly inside its window (the
//                  neighbour midpoints, else the search window);
//   ILL_CONDITIONED |A + B| <= 64 eps (|A| + |B|);
//   DERIVATIVE_UNSTABLE the first-order rounding bound of U,
//                  c (|B| dA + |A| dB) / (A + B)^2 with dA = g |A|, dB = g |B|
//                  and g = 32 eps (L + 1) for the forward-mode partials,
//                  exceeds 0.25 group_rtol max(c, |U|) (the reference's
//                  acceptance test on its finite-difference error estimate);
//   NONFINITE      a non-finite derivative or U.
// A flagged root has U = NaN.
__device__ __forceinline__ double group_result(const PrepRow *rows, int L, int top, int wave, double omega,
                                               const double *roots, int m, int count, double lower, double upper,
                                               double close_rtol, double group_rtol, unsigned *flags) {
  constexpr double eps = 2.220446049250313e-16;
  const double c = m < count ? roots[m] : NAN;
  *flags = 0u;
  if (!(c == c) || !(c > 0.)) { *flags = kGroupMissing; return NAN; }
  const bool prev = m > 0, next = m + 1 < count;
  if ((prev && c - roots[m - 1] <= close_rtol * c) || (next && roots[m + 1] - c <= close_rtol * roots[m + 1])) {
    *flags = kGroupClose;
    return NAN;
  }
  const double lo = prev ? .5 * (roots[m - 1] + c) : lower, hi = next ? .5 * (c + roots[m + 1]) : upper;
  if (!(lo > 0. && c > lo && c < hi)) { *flags = kGroupPolish; return NAN; }
  double fk, fw;
  secular_partials(rows, L, top, wave, omega, c, &fk, &fw);
  const double k = omega / c, A = -k * fk, B = k * fk + omega * fw, den = A + B;
  if (!isfinite(A) || !isfinite(B)) { *flags = kGroupNonfinite; return NAN; }
  if (!(fabs(den) > 64. * eps * (fabs(A) + fabs(B)))) { *flags = kGroupIll; return NAN; }
  const double u = -fk / fw;
  if (!isfinite(u)) { *flags = kGroupNonfinite; return NAN; }
  const double g = 32. * eps * double(L + 1);
  const double err = c * (2. * g * fabs(A) * fabs(B)) / (den * den);
  if (!(err <= .25 * group_rtol * fmax(c, fabs(u)))) { *flags = kGroupUnstable; return NAN; }
  return u;
}

// Rayleigh (0) and Love (1) get their own register allocations; two waves per
// SIMD bind only the Rayleigh one (a few spills, still faster than one wave).
template <int Wave>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2))) void group_kernel(PipeArgs a) {
  const int ntl = a.trace_end - a.trace_begin;
  // Mode-major: neighbouring lanes take one mode of neighbouring traces (the
  // next periods of one model), whose roots sit in the same layer regimes, so
  // the propagating/evanescent branches of a wave mostly agree. Guard roots
  // (the last two of kmax) have no group velocity and get no lane.
  const int total = ntl * (a.kmax - 2);
  for (int q = blockIdx.x * kBlock + threadIdx.x; q < total; q += gridDim.x * kBlock) {
    const int m = q / ntl, tl = q - m * ntl;
    const TraceInfo ti = a.traces[tl + a.trace_begin];
    const ModelInfo mi = a.models[ti.model];
    if (m >= mi.modes) continue;
    const TraceScan &ts = a.scan[tl];
    const size_t cell = static_cast<size_t>(mi.out0) + static_cast<size_t>(m) * mi.T + ti.f;
    const double *roots = a.roots + static_cast<size_t>(tl) * a.kmax;
    unsigned gf = kGroupMissing;
    double u = NAN;
    if (!((ts.rejected >> m) & 1u))
      u = group_result(a.rows + mi.row0, mi.L, mi.top, Wave, ti.omega, roots, m, ts.count, mi.lo, mi.hi,
                       a.close_rtol, a.group_rtol, &gf);
    a.group[cell] = u;
    a.group_flags[cell] = gf;
  }
}

// Group velocities of one confirmed trace from its ordered roots (the
// continuity review's helper, reference cpp/common/review_groups.hpp:10-31):
// lane m takes reported mode m; modes beyond the known roots are missing.
__global__ void review_group_kernel(const PrepRow *rows, int L, int top, int wave, double omega, const double *roots,
                                    int known, int modes, double lower, double upper, double close_rtol,
                                    double group_rtol, double *out, uint32_t *flags) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= modes) return;
  unsigned f = kGroupMissing;
  double u = NAN;
  if (m < known) u = group_result(rows, L, top, wave, omega, roots, m, known, lower, upper, close_rtol, group_rtol, &f);
  out[m] = u;
  flags[m] = f;
}

hipError_t launch_review_groups(const PrepRow *rows, int L, int top, int wave, double omega, const double *roots,
                                int known, int modes, double lower, double upper, double close_rtol,
                                double group_rtol, double *out, uint32_t *flags, hipStream_t s) {
  hipLaunchKernelGGL(review_group_kernel, dim3((modes + 63) / 64), dim3(64), 0, s, rows, L, top, wave, omega, roots,
                     known, modes, lower, upper, close_rtol, group_rtol, out, flags);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Anchor continuation (the step rules follow cpp/common/continuation.hpp:12-212).
// ---------------------------------------------------------------------------
// Anchor slopes dc/dlnT = c (c/U - 1) from the analytic group velocity U.
__global__ __launch_bounds__(kBlock) void slope_kernel(ContArgs a) {
  const int total = a.nanchors * a.kmax;
  for (int q = blockIdx.x * kBlock + threadIdx.x; q < total; q += gridDim.x * kBlock) {
    const int p = q / a.kmax;
    const double c = a.anchor_roots[q];
    double slope = NAN;
    if (c == c) {
      const TraceInfo ti = a.traces[a.anchors[p]];
      const ModelInfo mi = a.models[ti.model];
      const double u = group_at(a.rows + mi.row0, mi.L, mi.top, a.wave, ti.omega, c);
      if (u == u && fabs(u) < INFINITY && u != 0.) slope = c * (c / u - 1.);
    }
    a.anchor_slopes[q] = slope;
  }
}

__global__ void anchor_flag_kernel(ContArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.nanchors) return;
  const TraceInfo ti = a.traces[a.anchors[p]];
  a.prediction[a.models[ti.model].trace_out0 + ti.f] = kPredAnchor;
}

// Anchor roots usable as a continuation source (continuation.hpp:22-31).
__device__ bool cont_source_safe(const double *roots, int nr, int K, unsigned flags, double close_rtol) {
  constexpr unsigned unusable = 4u | 8u | 16u | 32u | 64u | 128u | 512u | 1024u;
  if (nr < 1 || nr > K || (flags & unusable)) return false;
  for (int j = 0; j < nr; ++j) {
    if (!(roots[j] > 0.) || !(roots[j] < INFINITY)) return false;
    if (j && roots[j] - roots[j - 1] <= close_rtol * roots[j]) return false;
  }
  return true;
}

__device__ double cont_cap(const double *pred, int j, int nr, double lower, double upper) {
  const double c = pred[j];
  double gap = fmin(c - lower, upper - c);
  if (j) gap = fmin(gap, c - pred[j - 1]);
  if (j + 1 < nr) gap = fmin(gap, pred[j + 1] - c);
  return fmin(.02 * c, .2 * gap);
}

// Cubic Hermite prediction in ln T from the two anchors (continuation.hpp:42-71).
__device__ unsigned cont_predict(const double *cl0, const double *cr0, const double *sl0, const double *sr0, int nr,
                                 double dl, double dr, double lower, double upper, double tol, double *pred,
                                 double *dis) {
  if (!(dl > 0. && dr < 0. && upper > lower)) return kPredWindow;
  const double span = dl - dr, u = dl / span, u2 = u * u, u3 = u2 * u;
  double last_l = 0., last_r = 0., last_up = 0.;
  for (int j = 0; j < nr; ++j) {
    const double c0 = cl0[j], c1 = cr0[j], sl = sl0[j], sr = sr0[j];
    if (!(fabs(sl) < INFINITY) || !(fabs(sr) < INFINITY)) return kPredAnchorUnsafe;
    const double pl = c0 + sl * dl, pr = c1 + sr * dr;
    const double c = (2. * u3 - 3. * u2 + 1.) * c0 + (u3 - 2. * u2 + u) * span * sl + (-2. * u3 + 3. * u2) * c1 +
                     (u3 - u2) * span * sr;
    if (!(fabs(c) < INFINITY) || !(fabs(pl) < INFINITY) || !(fabs(pr) < INFINITY) || !(c > lower && c < upper))
      return kPredWindow;
    const double plo = fmin(pl, pr), phi = fmax(pl, pr);
    if (j && (c <= pred[j - 1] || pl <= last_l || pr <= last_r || plo <= last_up)) return kPredWindow | kPredReview;
    last_l = pl;
    last_r = pr;
    last_up = phi;
    pred[j] = c;
    dis[j] = fmax(fabs(pl - pr), fabs(c - ((1. - u) * c0 + u * c1)));
  }
  for (int j = 0; j < nr; ++j) {
    const double cap = cont_cap(pred, j, nr, lower, upper);
    if (!(cap > 16. * tol * pred[j]) || dis[j] > .5 * cap) return kPredWindow;
  }
  return 0u;
}

// Golden-section search for a point of the opposite sign between two
// same-sign samples (an even root pair); true if one is found.
__device__ bool cont_hidden_pair(const PrepRow *rows, int L, int top, int wave, double omega, double lo, double hi,
                                 double sign, int &ev) {
  constexpr double phi = 0.6180339887498949;
  double p = hi - phi * (hi - lo), r = lo + phi * (hi - lo);
  double fp = sign * secular_rows(rows, L, top, wave, omega, p), fr = sign * secular_rows(rows, L, top, wave, omega, r);
  ev += 2;
  for (int it = 0; it < 40; ++it) {
    if (fp <= 0. || fr <= 0.) return true;
    ++ev;
    if (fp < fr) { hi = r; r = p; fr = fp; p = hi - phi * (hi - lo); fp = sign * secular_rows(rows, L, top, wave, omega, p); }
    else { lo = p; p = r; fp = fr; r = lo + phi * (hi - lo); fr = sign * secular_rows(rows, L, top, wave, omega, r); }
  }
  return false;
}

// Audit of one refined root's bracket (continuation.hpp:126-166): Love by the
// phase levels at the bracket ends; Rayleigh by three samples on each open
// side of the root, with a pair search at a same-sign trough.
__device__ unsigned cont_window(const PrepRow *rows, int L, int top, int wave, double omega, double a, double b,
                                double fa, double fb, double root, double tol, double level, double vs_half, int &ev) {
  if (wave) {
    constexpr double pi = 3.141592653589793;
    ev += 2;
    const double left = love_phase_dev(rows, L, top, omega, a, vs_half),
                 right = love_phase_dev(rows, L, top, omega, b, vs_half);
    if (!(left == left) || !(right == right)) return kPredAuditFailed;
    if (floor(left / pi) != level || floor(right / pi) != level - 1.) return kPredAuditFailed | kPredReview;
    return 0u;
  }
  const double margin = fmax(tol, 4. * 2.220446049250313e-16) * root;
  for (int side = 0; side < 2; ++side) {
    const double edge = side ? b : a, sign = (side ? fb : fa) < 0. ? -1. : 1.;
    if (fabs(root - edge) <= 2. * margin) continue;
    const double end = root + copysign(margin, edge - root), step = (end - edge) * .25;
    double mag[3];
    for (int k = 0; k < 3; ++k) {
      ++ev;
      const double f = sign * secular_rows(rows, L, top, wave, omega, edge + (k + 1) * step);
      if (!(f == f)) return kPredAuditFailed;
      if (f <= 0.) return kPredAuditFailed | kPredReview;
      mag[k] = f;
    }
    if (mag[1] < mag[0] && mag[1] < mag[2] && (mag[1] < .5 * fmin(mag[0], mag[2]) || mag[1] < .05)) {
      const double x = edge + step, y = edge + 3. * step;
      if (cont_hidden_pair(rows, L, top, wave, omega, fmin(x, y), fmax(x, y), sign, ev))
        return kPredAuditFailed | kPredReview;
    }
  }
  return 0u;
}

// The gaps between the accepted brackets must hold no further root
// (continuation.hpp:85-124): equal end signs, then samples spaced by the WKB
// mode count (at least 4 per gap) and at every layer speed inside the gap,
// none of opposite sign and none a trough that hides a pair.
__device__ unsigned cont_gaps(const PrepRow *rows, int L, int top, int wave, double omega, double lower,
                              double upper, const double *lo, const double *hi, int nr, int K, int &ev) {
  if (nr < 1 || nr > K) return kPredAuditFailed;
  if (nr == K) upper = hi[nr - 1];
  auto density = [&](double c) {
    const double k = omega / c, k2 = k * k;
    double n = 0.;
    for (int j = top; j < L - 1; ++j) {
      const double kb = omega * rows[j].ivs, ka = omega * rows[j].ivp;
      n += rows[j].h * sqrt(fmax(0., kb * kb - k2));
      if (wave == 0) n += rows[j].h * sqrt(fmax(0., ka * ka - k2));
    }
    return n * 0.3183098861837907;
  };
  for (int g = 0; g < nr + (nr < K); ++g) {
    const double a = g ? hi[g - 1] : lower, b = g < nr ? lo[g] : upper;
    if (!(b > a)) continue;
    const double fa = secular_rows(rows, L, top, wave, omega, a), fb = secular_rows(rows, L, top, wave, omega, b);
    ev += 2;
    if (!(fa == fa) || !(fb == fb)) return kPredAuditFailed;
    if (fa == 0. || fb == 0. || signbit(fa) != signbit(fb)) return kPredAuditFailed | kPredReview;
    const double sign = fa < 0. ? -1. : 1., endpoint = fmin(fabs(fa), fabs(fb));
    auto sample = [&](double c) -> unsigned {
      ++ev;
      const double f = secular_rows(rows, L, top, wave, omega, c);
      if (!(f == f)) return kPredAuditFailed;
      if (f == 0. || sign * f < 0. || (fabs(f) < .05 && fabs(f) < .1 * endpoint)) return kPredAuditFailed | kPredReview;
      return 0u;
    };
    const int n = max(4, min(64, int(ceil(8. * (density(b) - density(a))))));
    for (int i = 1; i <= n; ++i)
      if (unsigned issue = sample(a + (b - a) * double(i) / double(n + 1))) return issue;
    for (int j = 0; j < L; ++j) {
      const double v[2] = {rows[j].ivs > 0. ? 1. / rows[j].ivs : 0., 1. / rows[j].ivp};
      for (int body = 0; body < 2; ++body)
        for (int side = -1; side <= 1; ++side) {
          const double c = v[body] * (1. + side * 1e-7);
          if (c > a && c < b)
            if (unsigned issue = sample(c)) return issue;
        }
    }
  }
  return 0u;
}

// One lane per target trace.
__global__ __launch_bounds__(kBlock) void continuation_kernel(ContArgs a) {
  int ev = 0;  // FP64 secular (or Love phase) evaluations of this lane
  for (int q = blockIdx.x * kBlock + threadIdx.x; q < a.ntargets; q += gridDim.x * kBlock) {
    const TargetPlan tp = a.targets[q];
    const TraceInfo ti = a.traces[tp.trace];
    const ModelInfo mi = a.models[ti.model];
    const PrepRow *rows = a.rows + mi.row0;
    const TraceInfo tl = a.traces[a.anchors[tp.left]], tr = a.traces[a.anchors[tp.right]];
    const long long cl = mi.trace_out0 + tl.f, cr = mi.trace_out0 + tr.f, ct = mi.trace_out0 + ti.f;
    const int nr = a.counts[cl], K = mi.K;
    unsigned pred = ((a.flags[cl] | a.flags[cr]) & kFlagCloseRoot) ? kPredSourceUncertain : 0u;
    unsigned reason = 0u;
    const double *rl = a.anchor_roots + static_cast<size_t>(tp.left) * a.kmax;
    const double *rr = a.anchor_roots + static_cast<size_t>(tp.right) * a.kmax;
    double pv[kContMax], dis[kContMax], lo[kContMax], hi[kContMax], roots[kContMax];
    if (mi.empty) reason = kPredWindow;
    else if (nr <= 0 || nr > K || nr > kContMax || a.counts[cr] != nr) reason = kPredRootCount;
    else if (!cont_source_safe(rl, nr, K, a.flags[cl], a.close_rtol) ||
             !cont_source_safe(rr, nr, K, a.flags[cr], a.close_rtol))
      reason = kPredAnchorUnsafe;
    else
      reason = cont_predict(rl, rr, a.anchor_slopes + static_cast<size_t>(tp.left) * a.kmax,
                            a.anchor_slopes + static_cast<size_t>(tp.right) * a.kmax, nr, tp.dl, tp.dr, mi.lo, mi.hi,
                            a.tol, pv, dis);
    if (!reason) {
      // Local brackets around the predictions, widened up to three times
      // (continuation.hpp:170-212), refined in FP64, then audited.
      double first_level = 0.;
      if (a.wave) {
        const double ph = love_phase_dev(rows, mi.L, mi.top, ti.omega, mi.lo, mi.vs_half);
        ++ev;
        if (!(ph == ph)) reason = kPredAuditFailed;
        first_level = floor(ph / 3.141592653589793);
      }
      for (int j = 0; j < nr && !reason; ++j) {
        const double c = pv[j], cap = cont_cap(pv, j, nr, mi.lo, mi.hi);
        double width = fmin(cap, fmax(16. * a.tol * c, fmax(2. * dis[j], .05 * c * fmin(fabs(tp.dl), fabs(tp.dr)))));
        double x0 = 0., x1 = 0., f0 = 0., f1 = 0.;
        bool bracket = false;
        for (int attempt = 0; attempt < 4; ++attempt) {
          x0 = fmax(mi.lo, c - width);
          x1 = fmin(mi.hi, c + width);
          f0 = secular_rows(rows, mi.L, mi.top, a.wave, ti.omega, x0);
          f1 = secular_rows(rows, mi.L, mi.top, a.wave, ti.omega, x1);
          ev += 2;
          if (f0 == f0 && f1 == f1 && f0 != 0. && f1 != 0. && signbit(f0) != signbit(f1)) { bracket = true; break; }
          if (width >= cap) break;
          width = fmin(cap, 2. * width);
        }
        if (!bracket) { reason = kPredBracket; break; }
        const double root = refine_serial(a.refinement, rows, mi.L, mi.top, a.wave, ti.omega, x0, x1, f0, f1, a.tol,
                                          ev);
        if (!(root >= x0 && root <= x1)) { reason = kPredCorrection; break; }
        if (j && root - roots[j - 1] <= a.close_rtol * fabs(root)) { reason = kPredCorrection | kPredReview; break; }
        if (a.f_accept < 1.) {
          const double fr = secular_rows(rows, mi.L, mi.top, a.wave, ti.omega, root);
          ++ev;
          if (!(fabs(fr) <= a.f_accept)) { reason = kPredCorrection; break; }
        }
        const unsigned issue = cont_window(rows, mi.L, mi.top, a.wave, ti.omega, x0, x1, f0, f1, root, a.tol,
                                           first_level - j, mi.vs_half, ev);
        if (issue) { reason = issue; pred |= kPredAudit; break; }
        lo[j] = x0;
        hi[j] = x1;
        roots[j] = root;
      }
      if (!reason) {
        pred |= kPredAudit;
        reason = cont_gaps(rows, mi.L, mi.top, a.wave, ti.omega, mi.lo, mi.hi, lo, hi, nr, K, ev);
      }
    }
    if (reason) {
      a.prediction[ct] = pred | reason | kPredFallback;
      a.fallback[atomicAdd(a.fallback_count, 1)] = tp.trace;
      continue;
    }
    a.prediction[ct] = pred | kPredAccepted;
    unsigned word = kFlagApproximate | kFlagContinuation;
    for (int m = 0; m < mi.modes; ++m) {
      const double c = m < nr ? roots[m] : NAN;
      if (c == c && c >= .99 * mi.vs_half) word |= kFlagNearCutoff;
      const size_t cell = static_cast<size_t>(mi.out0) + static_cast<size_t>(m) * mi.T + ti.f;
      a.velocity[cell] = c;
      if (a.group) {
        unsigned gf = 0;
        const double u = group_result(rows, mi.L, mi.top, a.wave, ti.omega, roots, m, nr, mi.lo, mi.hi, a.close_rtol,
                                      a.group_rtol, &gf);
        a.group[cell] = u;
        a.group_flags[cell] = gf;
      }
    }
    a.flags[ct] = word;
    a.levels[ct] = kLevelContinuation;
    a.counts[ct] = nr;
  }
  if (a.evaluations) {
    unsigned long long e = static_cast<unsigned long long>(ev);
    for (int o = 32; o > 0; o >>= 1) e += __shfl_down(e, o, 64);
    if ((threadIdx.x & 63) == 0 && e) atomicAdd(a.evaluations, e);
  }
}

__global__ void gather_kernel(const TraceInfo *all, const int *list, int n, TraceInfo *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = all[list[i]];
}

// Reset the device work counters (a kernel, so the sequence can be graph-captured).
__global__ void reset_kernel(int *counters, int n) {
  if (threadIdx.x < n) counters[threadIdx.x] = 0;
}

__global__ void reset_stats_kernel(unsigned long long *stats, int n) {
  if (threadIdx.x < n) stats[threadIdx.x] = 0ull;
}

// Accumulate this chunk's work counters into the run statistics:
// [0] traces, [1] FP64 rechecks, [2] pair probes, [3] FP64 rescans,
// [4] Love phase searches, [5] evaluated scan nodes (nodes_kernel),
// [6] FP64 refinement evaluations, [7] FP32 and [8] FP64 probe evaluations
// (the consumer kernels add these themselves).
__global__ void stats_kernel(const int *counters, unsigned long long *stats, int traces) {
  if (threadIdx.x == 0) {
    stats[0] += static_cast<unsigned long long>(traces);
    stats[1] += static_cast<unsigned long long>(counters[0]);
    stats[2] += static_cast<unsigned long long>(counters[1]);
    stats[3] += static_cast<unsigned long long>(counters[2]);
    stats[4] += static_cast<unsigned long long>(counters[3]);
  }
}

// Sum of the grid nodes every trace of the chunk evaluated (one atomic per block).
__global__ __launch_bounds__(kBlock) void nodes_kernel(const TraceScan *scan, int n, unsigned long long *stats) {
  unsigned long long s = 0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) s += scan[i].nodes;
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  __shared__ unsigned long long part[kBlock / 64];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t += part[w];
    atomicAdd(stats + 5, t);
  }
}

}  // namespace mdr

// ---------------------------------------------------------------------------
// Launchers used by the host pipeline.
// ---------------------------------------------------------------------------
namespace mdr {

// Models of up to kAdaptRows rows: four traces per block; deeper models (up to
// kMaxRows) take the large-scratch variant, one trace per block.
template <class T, int W>
static void launch_scan_w(const PipeArgs &a, int n, hipStream_t s) {
  hipLaunchKernelGGL((scan_kernel<T, kAdaptRows, W>), dim3((n + W - 1) / W), dim3(64 * W), 0, s, a);
}

template <class T>
static hipError_t launch_scan_t(const PipeArgs &a, hipStream_t s) {
  const int n = a.trace_end - a.trace_begin;
  if (n <= 0) return hipSuccess;
  if (a.max_rows > kAdaptRows) {
    hipLaunchKernelGGL((scan_kernel<T, kMaxRows, 1>), dim3(n), dim3(64), 0, s, a);
  } else if (a.scan_waves == 1) {
    launch_scan_w<T, 1>(a, n, s);
  } else if (a.scan_waves == 2) {
    launch_scan_w<T, 2>(a, n, s);
  } else if (a.scan_waves == 8) {
    if constexpr (sizeof(T) == 4) launch_scan_w<T, 8>(a, n, s);  // 8 FP64 scratches exceed 64 KB of LDS
    else launch_scan_w<T, 4>(a, n, s);
  } else {
    launch_scan_w<T, 4>(a, n, s);
  }
  return hipGetLastError();
}

hipError_t probe_residency(int cus, int wave, int out[2]) {
  int b32 = 0, b64 = 0;
  hipError_t e = wave == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b32, trough_kernel<float, 1>, kBlock, 0)
                           : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b32, trough_kernel<float, 0>, kBlock, 0);
  if (e != hipSuccess) return e;
  e = wave == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b64, trough_kernel<double, 1>, kBlock, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b64, trough_kernel<double, 0>, kBlock, 0);
  if (e != hipSuccess) return e;
  out[0] = b32 * cus * (kBlock / 64);
  out[1] = b64 * cus * (kBlock / 64);
  return hipSuccess;
}

template <class K>
static hipError_t report_of(K kernel, int block, KernelReport *r) {
  hipFuncAttributes at{};
  hipError_t e = hipFuncGetAttributes(&at, reinterpret_cast<const void *>(kernel));
  if (e != hipSuccess) return e;
  r->registers = at.numRegs;
  r->local_bytes = static_cast<int>(at.localSizeBytes);
  r->static_shared = static_cast<int>(at.sharedSizeBytes);
  r->dynamic_shared = 0;
  r->block = block;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(&r->active_per_cu, kernel, block, 0);
}

template <class T>
static hipError_t scan_report(int L, int waves, KernelReport *r) {
  r->table_rows = L > kAdaptRows ? kMaxRows : kAdaptRows;
  if (L > kAdaptRows) return report_of(scan_kernel<T, kMaxRows, 1>, 64, r);
  if (waves == 1) return report_of(scan_kernel<T, kAdaptRows, 1>, 64, r);
  if (waves == 2) return report_of(scan_kernel<T, kAdaptRows, 2>, 128, r);
  if constexpr (sizeof(T) == 4)
    if (waves == 8) return report_of(scan_kernel<T, kAdaptRows, 8>, 512, r);
  return report_of(scan_kernel<T, kAdaptRows, 4>, 256, r);
}

hipError_t kernel_report(int which, int L, bool fp32, int scan_waves, KernelReport *r) {
  r->events_per_trace = kEventCap;
  if (which == 1) {
    r->table_rows = L;
    return report_of(love_phase_kernel, kBlock, r);
  }
  return fp32 ? scan_report<float>(L, scan_waves, r) : scan_report<double>(L, scan_waves, r);
}

hipError_t launch_continuation(const ContArgs &c, int persistent_blocks, hipStream_t s) {
  if (c.nanchors > 0) {
    hipLaunchKernelGGL(slope_kernel, dim3(persistent_blocks), dim3(kBlock), 0, s, c);
    hipLaunchKernelGGL(anchor_flag_kernel, dim3((c.nanchors + 255) / 256), dim3(256), 0, s, c);
  }
  if (c.ntargets > 0) hipLaunchKernelGGL(continuation_kernel, dim3(persistent_blocks), dim3(kBlock), 0, s, c);
  return hipGetLastError();
}

hipError_t launch_gather(const TraceInfo *all, const int *list, int n, TraceInfo *out, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(gather_kernel, dim3((n + 255) / 256), dim3(256), 0, s, all, list, n, out);
  return hipGetLastError();
}

template <int Wave>
static void launch_refine(const PipeArgs &pa, int pb, hipStream_t s) {
  if (pa.refinement == 2) hipLaunchKernelGGL((refine_kernel<2, Wave>), dim3(pb), dim3(kBlock), 0, s, pa);
  else if (pa.refinement == 0) hipLaunchKernelGGL((refine_kernel<0, Wave>), dim3(pb), dim3(kBlock), 0, s, pa);
  else hipLaunchKernelGGL((refine_kernel<1, Wave>), dim3(pb), dim3(kBlock), 0, s, pa);
}

hipError_t launch_pipeline(const ScanArgs &sa, const PipeArgs &pa, bool fp32, bool group, int persistent_blocks,
                           hipStream_t s, hipEvent_t *marks) {
  const int n = sa.trace_end - sa.trace_begin;
  if (n <= 0) return hipSuccess;
  auto mark = [&](int i) {
    if (marks) (void)hipEventRecord(marks[i], s);
  };
  hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, s, sa.counters, 5);
  // Run statistics accumulate over chunks; cleared by a kernel (not a memset
  // node) so that the sequence stays capturable.
  if (pa.stats && pa.trace_begin == 0)
    hipLaunchKernelGGL(reset_stats_kernel, dim3(1), dim3(64), 0, s, pa.stats, kStats);
  mark(0);
  hipError_t err = fp32 ? launch_scan_t<float>(sa, s) : launch_scan_t<double>(sa, s);
  if (err != hipSuccess) return err;
  mark(1);
  const int pb = persistent_blocks;
  if (fp32) {
    if (pa.max_rows <= kAdaptRows)
      hipLaunchKernelGGL((rescan_kernel<kAdaptRows, kWaves>), dim3(pb), dim3(kBlock), 0, s, pa);
    else
      hipLaunchKernelGGL((rescan_kernel<kMaxRows, 1>), dim3(pb), dim3(64), 0, s, pa);
  }
  if (pa.wave == 1 && pa.love_phase)
    hipLaunchKernelGGL(love_range_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, pa);
  mark(2);
  hipLaunchKernelGGL(compact_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, pa);
  if (pa.wave == 1) hipLaunchKernelGGL(recheck_kernel<1>, dim3(pb), dim3(kBlock), 0, s, pa);
  else hipLaunchKernelGGL(recheck_kernel<0>, dim3(pb), dim3(kBlock), 0, s, pa);
  if (fp32) hipLaunchKernelGGL(unc_probe_kernel, dim3(pb), dim3(kBlock), 0, s, pa);
  if (pa.wave == 1) {
    if (fp32) hipLaunchKernelGGL((trough_kernel<float, 1>), dim3(pb), dim3(kBlock), 0, s, pa);
    hipLaunchKernelGGL((trough_kernel<double, 1>), dim3(pb), dim3(kBlock), 0, s, pa);
  } else {
    if (fp32) hipLaunchKernelGGL((trough_kernel<float, 0>), dim3(pb), dim3(kBlock), 0, s, pa);
    hipLaunchKernelGGL((trough_kernel<double, 0>), dim3(pb), dim3(kBlock), 0, s, pa);
  }
  mark(3);
  hipLaunchKernelGGL(assemble_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, pa);
  if (pa.wave == 1) {
    if (pa.love_phase) {
      hipLaunchKernelGGL(love_count_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, pa);
      hipLaunchKernelGGL(love_phase_kernel, dim3(pb), dim3(kBlock), 0, s, pa);
    }
  }
  mark(4);
  if (pa.wave == 1) launch_refine<1>(pa, pb, s);
  else launch_refine<0>(pa, pb, s);
  hipLaunchKernelGGL(finalize_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, pa);
  mark(5);
  if (group && pa.wave == 1) hipLaunchKernelGGL(group_kernel<1>, dim3(pb), dim3(kBlock), 0, s, pa);
  else if (group) hipLaunchKernelGGL(group_kernel<0>, dim3(pb), dim3(kBlock), 0, s, pa);
  mark(6);
  if (pa.stats) {
    hipLaunchKernelGGL(stats_kernel, dim3(1), dim3(64), 0, s, pa.counters, pa.stats, n);
    hipLaunchKernelGGL(nodes_kernel, dim3((n + kBlock - 1) / kBlock < 1024 ? (n + kBlock - 1) / kBlock : 1024), dim3(kBlock), 0, s, pa.scan, n,
                       pa.stats);
  }
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

// Ragged batches resident in HBM. Every trace first becomes a dead trace of
// the sentinel model (index B: empty window, no reported modes, its trace
// output at the trace's own index); then every model whose offsets are
// consistent writes its rows, its ModelInfo and its own traces, one wavefront
// per model. A model whose row, period or cell range is out of bounds, or
// whose cell range does not hold modes x periods cells, keeps the sentinel
// and is counted in *invalid, so no later kernel reads or writes outside the
// caller's arrays, whatever the offsets say.
__device__ ModelInfo sentinel_model() {
  ModelInfo mi{};
  mi.L = 1;
  mi.K = 2;
  mi.empty = 1;
  return mi;
}

__global__ void prep_dead_traces_kernel(long long ntraces, int sentinel, TraceInfo *traces, ModelInfo *models,
                                        int *invalid) {
  const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t == 0) {
    models[sentinel] = sentinel_model();
    *invalid = 0;
  }
  if (t >= ntraces) return;
  TraceInfo ti{};
  ti.omega = 1.;
  ti.model = sentinel;
  ti.f = static_cast<int>(t);
  traces[t] = ti;
}

__global__ __launch_bounds__(kBlock) void prep_ragged_kernel(const double *h, const double *vp, const double *vs,
                                                             const double *rho, const double *bounds,
                                                             const double *periods, RaggedGeometry g, PrepParams p,
                                                             PrepRow *rows, ModelInfo *models, TraceInfo *traces,
                                                             int *invalid) {
  const int lane = threadIdx.x & 63;
  const int m = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (m >= g.B) return;
  const long long r0 = g.row_offsets[m], r1 = g.row_offsets[m + 1];
  const long long p0 = g.period_offsets[m], p1 = g.period_offsets[m + 1];
  const long long o0 = g.cell_offsets[m], o1 = g.cell_offsets[m + 1];
  const long long K = g.modes[m], L = r1 - r0, T = p1 - p0;
  const bool ok = r0 >= 0 && L >= 1 && L <= g.max_rows && r1 <= g.nrows && p0 >= 0 && T >= 0 && p1 <= g.ntraces &&
                  K >= 1 && K <= g.max_modes && o0 >= 0 && o1 <= g.ncells && o1 - o0 == K * T;
  if (!ok) {
    if (lane == 0) {
      models[m] = sentinel_model();
      atomicAdd(invalid, 1);
    }
    return;
  }
  for (long long j = r0 + lane; j < r1; j += 64) rows[j] = make_prep_row(h[j], vp[j], vs[j], rho[j]);
  double slow;
  const ModelInfo mi = make_model_info(h + r0, vp + r0, vs + r0, static_cast<int>(L), bounds[2 * m],
                                       bounds[2 * m + 1], p.wave, static_cast<int>(K), static_cast<int>(T), r0, o0,
                                       p0, &slow);
  if (lane == 0) models[m] = mi;
  for (int f = lane; f < T; f += 64) traces[p0 + f] = make_trace(m, f, periods[p0 + f], slow, p);
}

hipError_t launch_prep_ragged(const double *h, const double *vp, const double *vs, const double *rho,
                              const double *bounds, const double *periods, const RaggedGeometry &g,
                              const PrepParams &p, PrepRow *rows, ModelInfo *models, TraceInfo *traces, int *invalid,
                              hipStream_t s) {
  const long long n = g.ntraces > 0 ? g.ntraces : 1;
  hipLaunchKernelGGL(prep_dead_traces_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s,
                     g.ntraces, g.B, traces, models, invalid);
  const int per_block = kBlock / 64;
  hipLaunchKernelGGL(prep_ragged_kernel, dim3((g.B + per_block - 1) / per_block), dim3(kBlock), 0, s, h, vp, vs,
                     rho, bounds, periods, g, p, rows, models, traces, invalid);
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
