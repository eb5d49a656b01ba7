// Device data layout shared by the host pipeline and the kernels.
//
// HBM layout (one solver call, one chunk of traces):
//   rows     [sum_m L_m]            PrepRow, 64 B per layer row, model-major
//   models   [B]                    ModelInfo
//   traces   [ntrace]               TraceInfo, model-major (trace = model, period)
//   events   [chunk][kEventCap]     Event, scan -> recheck/trough -> assemble
//   scan     [chunk]                TraceScan
//   brackets [chunk][kmax]          Bracket (FP64 endpoints and values)
//   roots    [chunk][kmax]          refined roots
//   outputs  velocity [sum_m modes_m T_m], flags/levels/counts [sum_m T_m],
//            group/group_flags like velocity
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace mdr {

constexpr int kBlock = 256;       // lanes per scan block (4 wave64)
constexpr int kMaxRows = 256;     // layer rows per model (reference limit, __init__.py:152)
constexpr int kEventCap = 64;     // ascending events kept per trace (one bit each in TraceScan's masks)
constexpr int kStats = 9;         // run statistics (stats_kernel)
constexpr int kFloor = 64;        // uniform table points of the adaptive grid (one per lane)
constexpr int kAdaptRows = 64;    // rows of the common scan scratch; deeper models use a kMaxRows scratch
constexpr int kCellMax = 1 << 16; // density nodes per table cell (cap)
constexpr float kTroughMax = 0.05f;  // |F| below which a same-sign trough is probed ...
constexpr float kTroughRel = 0.9f;   // ... or below this share of both neighbours' |F| (a trough at any height)
constexpr int kTroughIters = 28;  // golden-section steps per trough / dip probe
constexpr float kDipDepth = 0.5f; // log2-scale dip (bits below the neighbours' mean) that is probed
constexpr double kGuardTol = 1e-7;  // relative tolerance of the guard roots (modes+1, modes+2)
constexpr int kDipStopIter = 6;   // a dip probe gives up after this many steps ...
constexpr float kDipStopDrop = 1.5f;  // ... unless the log2 scale fell this far below the node's

// Node kinds of a scan event (kBracket, kUncertain, kProbe) and probe kinds
// (kTrough, kDip) of a node searched for a hidden root pair.
enum EventKind : int { kNone = 0, kBracket = 1, kUncertain = 2, kTrough = 3, kDip = 4, kProbe = 5 };

// Trace-flag bits (reference python/modisp/__init__.py:15-32).
constexpr unsigned kFlagUnconverged = 4u, kFlagNearCutoff = 16u, kFlagRefineFailed = 64u, kFlagRejected = 128u,
                   kFlagCloseRoot = 4096u, kFlagLoveFallback = 16384u;
constexpr unsigned kTraceOverflow = 1u << 31;  // internal: event list overflowed
// Group-velocity flags (reference cpp/common/group_velocity.hpp:9-12).
constexpr unsigned kGroupMissing = 1u, kGroupPolish = 2u, kGroupUnstable = 4u, kGroupIll = 8u, kGroupClose = 16u,
                   kGroupNonfinite = 32u;
// search_level codes (reference docs/API.md:203-209): a trace accepted by the
// grid scan reports the code of the reference search its strategy names
// (8/9/10 shape acceptance at high, balanced or auto, fast; 3 high adaptive;
// 6/7 balanced/fast observation grid); Love traces whose roots follow the
// exact phase count report 11.
constexpr int kLevelShapeHigh = 8, kLevelShapeBalanced = 9, kLevelShapeFast = 10, kLevelAdaptive = 3,
              kLevelObservationBalanced = 6, kLevelObservationFast = 7, kLevelLovePhase = 11;
constexpr unsigned kScanLovePhase = 1u << 30;  // TraceScan: Love roots follow the exact phase count
constexpr unsigned kScanPhaseReady = 1u << 29;  // TraceScan: the Love phase count of the window is known (exact, phase_lo)

// One prepared layer row: reciprocals so a trace's constants are products.
struct PrepRow {
  double h;      // thickness (km); 0 for the half-space
  double ivp;    // 1 / vp
  double ivs;    // 1 / vs (0 for water)
  double g2;     // 2 vs^2
  double rho;    // density
  double irho;   // 1 / rho
  double irho2;  // 1 / rho^2
  double imu;    // 1 / (rho vs^2) (0 for water)
};

struct ModelInfo {
  int row0;        // first row in `rows`
  int L;           // rows including the half-space
  int top;         // 1 when row 0 is water (vs = 0)
  int K;           // roots searched: modes + 2
  int modes;       // roots reported
  int T;           // periods of this model
  long long out0;  // velocity/group base offset: cell = out0 + mode*T + f
  long long trace_out0;  // flags/levels/counts base offset: out = trace_out0 + f
  double lo, hi;   // search window
  double vs_half;  // half-space shear speed
  double cutoff_guard;  // FP32 nodes at or above this speed are rechecked in FP64
  double stable_low;    // FP32 nodes below this speed are rechecked in FP64
  double slowness;  // sum of h/vs over the finite layers (FP32 band sizing)
  int empty;       // 0 normal, 1 empty range, 2 no Love guide
  int low_free;    // 1 when no root can lie below stable_low (see make_model)
};

struct TraceInfo {
  double omega;  // angular frequency
  int model;
  int f;         // period index within the model
  float tau;     // FP32 uncertainty band on |F|
  int pad;
};

struct Event {
  double c0, c1, c2;  // nodes j-1, j, j+1 (NaN where absent); a bracket is [c1, c2]
  double v;           // FP64 value at c1 (uncertain nodes)
  double x;           // split point of a hidden root pair (NaN: none)
  double vx;          // F at the split point
  float fl, fc, fr;   // scan values at c0, c1, c2
  float sl, sc, sr;   // summed log2 normalization scales at c0, c1, c2
  int node;           // node index j within the trace
  int kind;           // EventKind of the scan classification
  int probe;          // bits 0-7: kTrough / kDip when the node is probed for a hidden pair;
                      // bit 8 / bits 16-23: probe handed from FP32 to FP64, after that many steps
  int pad;            // best key of a handed-over probe (float bits)
};

// One point of a model's scan table: a window floor point or a layer speed,
// with the frequency-free mode density there (N(c) = omega * d).
struct TablePoint {
  double c;
  float d;
  int kink;  // 1 at a layer speed (the next cell's nodes are spaced quadratically)
};

struct TraceScan {
  int nev;
  unsigned flags;
  int status;   // ModelInfo::empty for empty traces
  int certain;  // certain sign changes seen by the scan
  int count;    // roots found (<= K), -1 unresolved
  int nodes;    // secular evaluations of the trace's scan (lane slots, batch overlaps included)
  uint64_t umask, pmask;  // events (bit e) that are FP32-uncertain / probed for a hidden pair
  unsigned rejected;      // reported modes (bit m) whose |F| failed the f_accept threshold
  int exact;              // Love with kScanPhaseReady: roots in the window by the phase count
  double phase_lo;        // Love with kScanPhaseReady: the unwrapped phase at the window bottom
};

// A root bracket with F at its ends and the summed log2 normalization scales
// there (NaN where unknown). Refinement works on F 2^(ls - max(slo, shi)):
// the unnormalized secular function up to a constant, whose zero is F's but
// which stays smooth where a deep evanescent stack turns the normalized F
// into a +-1 step.
struct Bracket {
  double lo, hi, flo, fhi;
  float slo, shi;
};

struct PipeArgs {
  const TraceInfo *traces;
  const ModelInfo *models;
  const PrepRow *rows;
  Event *events;
  TraceScan *scan;
  Bracket *brackets;
  double *roots;
  uint32_t *unc_queue;
  uint32_t *trough_queue;
  int *counters;  // [0] uncertain nodes, [1] pair probes, [2] overflowed FP32 traces, [3] Love phase searches,
                  // [4] probes handed from FP32 to FP64
  int *unc_count;
  int *trough_count;
  int *overflow_count;
  int *overflow_list;  // chunk-local trace indices to rescan in FP64
  int *phase_count;
  int *handoff_count;   // probes the FP32 probe pass handed to the FP64 pass
  uint32_t *handoff_list;  // their event ids (reuses the recheck queue, free by then)
  int *phase_list;     // chunk-local Love traces whose brackets disagree with the phase count
  int queue_cap;
  int trace_begin, trace_end;
  int nc;
  int max_rows;   // largest row count of any model in the batch (<= kMaxRows)
  int scan_waves; // traces per scan block (1, 2, 4, 8)
  int adaptive;   // 1: density-adaptive grid, 0: uniform nc-node grid
  float rho;      // adaptive grid: nodes per WKB mode
  int kmax;
  int wave;
  int fp32;
  int refinement;  // 0 legacy regula falsi, 1 Brent-Dekker, 2 Chandrupatla
  int level_grid;  // search_level of a trace the grid scan accepts (kLevelShape*, kLevelAdaptive, ...)
  int close_grid;  // 1: close roots of a scan-accepted trace raise kFlagCloseRoot (0 for high qe_adaptive
                   // Rayleigh, whose search never sets it; Love phase traces always test)
  int love_phase;  // 1: Love roots are verified by the exact phase count (reference love_phase=1)
  int probe_waves[2];  // probe-kernel waves resident on the device at once (FP32 pass, FP64 pass);
                      // a probe pass spreads its runs over no fewer waves than this
  unsigned trace_extra;  // flags every searched trace carries (kFlagApproximate for approximate presets)
  double f_accept;  // reported roots with |F| > f_accept are rejected (reference solver.cpp:333); 1 = off
  double tol;
  double close_rtol;
  double group_rtol;  // group velocities with a larger rounding bound are flagged unstable
  double *velocity64;
  float *velocity32;
  uint32_t *flags;
  int32_t *levels;
  int32_t *counts;
  double *group;
  uint32_t *group_flags;
  unsigned long long *stats;  // optional run statistics [kStats] (see stats_kernel)
  double *all_roots;  // optional: every (up to kmax) root per listed trace, [trace list position][kmax]
  const TablePoint *table;  // per-model scan tables [model][table_stride] (adaptive grid)
  const int *ntab;          // points in each model's table
  int table_stride;
};

// ---------------------------------------------------------------------------
// Anchor continuation across periods (reference cpp/common/continuation.hpp,
// cpp/cpu/solver.cpp:605-705): anchors are solved by the full pipeline; every
// other period of the model is predicted from its two anchors (cubic Hermite
// in ln T with slopes dc/dlnT = c (c/U - 1)), bracketed locally, refined in
// FP64 and audited; a target that fails falls back to the full pipeline.
// ---------------------------------------------------------------------------
constexpr unsigned kPredAnchor = 1u, kPredAccepted = 2u, kPredFallback = 4u, kPredBypassed = 8u,
                   kPredAudit = 16u, kPredAuditFailed = 32u, kPredSourceUncertain = 64u, kPredRootCount = 128u,
                   kPredAnchorUnsafe = 256u, kPredWindow = 512u, kPredBracket = 1024u, kPredCorrection = 2048u,
                   kPredReview = 4096u;
constexpr unsigned kFlagApproximate = 2048u, kFlagContinuation = 65536u;
constexpr int kLevelContinuation = 12;
constexpr int kContMax = 16;  // roots per trace the continuation kernel handles (K = modes + 2)

struct TargetPlan {
  int trace;      // index in the full trace array
  int left, right;  // anchor positions (into the anchor list)
  double dl, dr;  // log steps from the target period to the left / right anchor periods
};

struct ContArgs {
  const TraceInfo *traces;   // full trace array
  const ModelInfo *models;
  const PrepRow *rows;
  const int *anchors;        // anchor list: trace indices
  const TargetPlan *targets;
  int ntargets, nanchors, kmax, wave, refinement;
  const double *anchor_roots;  // [anchor][kmax]
  double *anchor_slopes;       // [anchor][kmax]
  double tol, close_rtol;
  double f_accept;             // a continued root with |F| > f_accept falls back (continuation.hpp:201-204)
  double group_rtol;
  double *velocity;
  uint32_t *flags;
  int32_t *levels, *counts;
  double *group;
  uint32_t *group_flags;
  uint32_t *prediction;        // [trace] (cell = model trace_out0 + f)
  int *fallback;               // trace indices to solve with the full pipeline
  int *fallback_count;
  unsigned long long *evaluations;  // optional: FP64 evaluations of the continuation kernel (added)
};
hipError_t launch_continuation(const ContArgs &c, int persistent_blocks, hipStream_t s);
hipError_t launch_review_groups(const PrepRow *rows, int L, int top, int wave, double omega, const double *roots,
                                int known, int modes, double lower, double upper, double close_rtol,
                                double group_rtol, double *out, uint32_t *flags, hipStream_t s);
hipError_t launch_gather(const TraceInfo *all, const int *list, int n, TraceInfo *out, hipStream_t s);
using ScanArgs = PipeArgs;

// marks (optional): 7 events recorded before the scan and after the scan,
// FP64 rescan, recheck/trough, assemble, refine/finalize and group stages.
constexpr int kStageMarks = 7;
hipError_t launch_tables(const ModelInfo *models, const PrepRow *rows, int nmodels, int max_rows, int wave, int stride,
                         TablePoint *table, int *ntab, hipStream_t s);
hipError_t launch_pipeline(const ScanArgs &sa, const PipeArgs &pa, bool fp32, bool group, int persistent_blocks,
                           hipStream_t s, hipEvent_t *marks);
// Waves of trough_kernel<float, wave> (out[0]) and <double, wave> (out[1])
// resident at once on `cus` compute units.
hipError_t probe_residency(int cus, int wave, int out[2]);

}  // namespace mdr

// ---------------------------------------------------------------------------
// Preparation shared by the host (ragged batches) and the device prep kernels.
// ---------------------------------------------------------------------------
#if defined(__HIPCC__)
#define MDP_HD __host__ __device__ inline
#else
#define MDP_HD inline
#endif

namespace mdr {

MDP_HD PrepRow make_prep_row(double h, double vp, double vs, double rho) {
  PrepRow r;
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

struct PrepParams {
  int wave;
  double tau_floor, tau_phase;
};

// Model-level constants (search window, guards, empty-family status) and the
// vertical slowness sum used to size the FP32 uncertainty band.
MDP_HD ModelInfo make_model_info(const double *h, const double *vp, const double *vs, int L, double lo, double hi,
                                 int wave, int modes, int T, long long row0, long long out0, long long trace_out0,
                                 double *slowness) {
  ModelInfo mi;
  mi.row0 = static_cast<int>(row0);
  mi.L = L;
  mi.top = vs[0] == 0. ? 1 : 0;
  mi.K = modes + 2;
  mi.modes = modes;
  mi.T = T;
  mi.out0 = out0;
  mi.trace_out0 = trace_out0;
  mi.lo = lo;
  mi.hi = hi;
  mi.vs_half = vs[L - 1];
  mi.cutoff_guard = vs[L - 1] * (1. - 1e-3);
  double vmax = 0., vmin = vs[L - 1], s = 0.;
  bool guide = false;
  for (int j = 0; j < L; ++j) {
    if (vs[j] > vmax) vmax = vs[j];
    if (j >= mi.top && j < L - 1 && h[j] > 0. && vs[j] < vmin) vmin = vs[j];
    if (j < L - 1) {
      s += vs[j] > 0. ? h[j] / vs[j] : h[j] / vp[j];
      // A Love guide needs a finite solid layer slower than the half-space
      // (reference cpp/common/physics.hpp:23-27).
      if (j >= mi.top && h[j] > 0. && vs[j] < vs[L - 1]) guide = true;
    }
  }
  mi.stable_low = vmax / 16.;
  // In a solid stack every Rayleigh or Love mode is faster than 0.68 of the
  // slowest shear speed (the Rayleigh speed of a layer exceeds 0.68 vs for
  // any vp/vs > sqrt(4/3)); with a margin of 2, no root lies below
  // stable_low, so an FP32 scan only needs FP64 signs for the nodes there.
  // A water top (Scholte modes) keeps the FP64 rescan of the whole trace.
  mi.low_free = (!mi.top && mi.stable_low <= 0.5 * vmin) ? 1 : 0;
  mi.empty = !(hi > lo) ? 1 : ((wave == 1 && !guide) ? 2 : 0);
  mi.slowness = s;
  *slowness = s;
  return mi;
}

MDP_HD TraceInfo make_trace(int model, int f, double period, double slowness, const PrepParams &p) {
  TraceInfo t;
  t.omega = 6.283185307179586476925286766559 / period;
  t.model = model;
  t.f = f;
  const double band = p.tau_phase * (1. + t.omega * slowness);
  t.tau = static_cast<float>(band > p.tau_floor ? band : p.tau_floor);
  t.pad = 0;
  return t;
}

}  // namespace mdr

namespace mdr {
hipError_t launch_prep(const double *h, const double *vp, const double *vs, const double *rho, const double *bounds,
                       const double *periods, int row_periods, int B, int L, int T, int modes, const PrepParams &p,
                       PrepRow *rows, ModelInfo *models, TraceInfo *traces, hipStream_t s);
// Offsets of a ragged batch resident in HBM (device arrays of B + 1 entries)
// and the extents of the caller's arrays they index.
struct RaggedGeometry {
  const int64_t *row_offsets, *period_offsets, *cell_offsets;
  const int32_t *modes;
  int B, max_rows, max_modes;
  long long nrows, ntraces, ncells;
};
// Prepares B models and g.ntraces traces (plus the sentinel model B) on the
// device; models with inconsistent offsets are counted in *invalid.
hipError_t launch_prep_ragged(const double *h, const double *vp, const double *vs, const double *rho,
                              const double *bounds, const double *periods, const RaggedGeometry &g,
                              const PrepParams &p, PrepRow *rows, ModelInfo *models, TraceInfo *traces, int *invalid,
                              hipStream_t s);
}  // namespace mdr

namespace mdr {
// Adjoint launch (adjoint.hip): one 64-lane block per model at a time.
struct AdjointArgs {
  const double *h, *vp, *vs, *rho;  // [B, L] models
  const double *periods;            // [T] or [B, T] (row_periods)
  const double *roots;              // [B, modes, T] phase velocities (NaN: no root)
  const double *weight;             // [B, modes, T] dLoss/dc
  double *gh, *gvp, *gvs, *grho;    // [B, L] sum over roots of weight * dc/dp
  double *tape;                     // [L * 6 * tape_lanes], lane-interleaved
  int B, L, T, modes, row_periods, wave, tape_lanes;
  int group;                        // models a block takes at once (LDS: group * 4 * L doubles)
};
hipError_t launch_adjoint(const AdjointArgs &a, int blocks, hipStream_t s);
// out[i] = float(in[i]) (round to nearest even, NaN kept) for n values.
hipError_t launch_narrow(const double *in, float *out, size_t n, hipStream_t s);
// Compiled resources of the kernel that scans a model of L rows (which = 0)
// or of the Love phase search (which = 1), for the plugin's kernel reports.
struct KernelReport {
  int registers, local_bytes, static_shared, dynamic_shared, block, active_per_cu, table_rows, events_per_trace;
};
hipError_t kernel_report(int which, int L, bool fp32, int scan_waves, KernelReport *r);
hipError_t launch_probe(const TraceInfo *traces, const ModelInfo *models, const PrepRow *rows, int ntraces, int nc,
                        int wave, int fp32, double *out, hipStream_t s);
}  // namespace mdr
