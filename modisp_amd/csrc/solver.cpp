// Host side of the modisp-rocm backend: C++17 over the HIP runtime, exposing
// the C ABI declared in include/modisp_rocm.h.
//
// One solver owns one device stream, its device workspace (sized for its
// capacity and grown on demand) and HIP events for stage timing. A call
// prepares layer rows and trace tables on the device, runs the root pipeline
// over chunks of traces, and (for host entry points) copies results back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/modisp_rocm.h"
#include "pipeline.hpp"

namespace {

thread_local std::string g_error;

int fail(int code, const std::string &msg) {
  g_error = "modisp-rocm: " + msg;
  return code;
}

#define MD_HIP(call)                                                                          \
  do {                                                                                        \
    const hipError_t err_ = (call);                                                           \
    if (err_ != hipSuccess)                                                                   \
      return fail(err_ == hipErrorOutOfMemory ? 3 : 1, std::string(#call) + ": " + hipGetErrorString(err_)); \
  } while (0)

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  int ensure(size_t n) {
    if (n <= bytes) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (n == 0) return 0;
    const hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(e == hipErrorOutOfMemory ? 3 : 1, "device allocation of " + std::to_string(n) + " bytes failed");
    }
    bytes = n;
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T *as() const { return static_cast<T *>(p); }
};

// Pinned host memory that results from host-array calls pass through:
// device-to-host copies into pinned memory run at the link's rate, where
// copies into the caller's pageable arrays are staged by the runtime at a
// fraction of it; the host then moves the bytes on into the caller's arrays
// with several threads (which also share the cost of first touching freshly
// allocated caller pages). Allocated once per solver and grown on demand.
struct HostStage {
  void *p = nullptr;
  size_t bytes = 0;
  bool ensure(size_t n) {
    if (n <= bytes) return true;
    release();
    if (hipHostMalloc(&p, n, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return false;
    }
    bytes = n;
    return true;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// One result array of a host-array call: `bytes` from device `src` to host `dst`.
struct HostCopy {
  void *dst;
  const void *src;
  size_t bytes;
};

// Copies every part on the solver's stream through pinned staging (or, if
// pinned memory cannot be had, straight into the caller's memory) and
// synchronizes the stream. The parts are taken as one byte stream in windows
// of at most kStageBytes: each window is copied to the staging, then spread
// into the caller's arrays by up to eight host threads.
constexpr size_t kStageBytes = size_t(256) << 20;
int copy_to_host(hipStream_t st, HostStage &stage, const std::vector<HostCopy> &parts) {
  size_t payload = 0;
  for (const HostCopy &c : parts) payload += c.bytes;
  if (payload == 0) return hipStreamSynchronize(st) == hipSuccess ? 0 : fail(1, "stream synchronization failed");
  if (!stage.ensure(std::min(payload, kStageBytes))) {
    for (const HostCopy &c : parts)
      if (c.bytes) MD_HIP(hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost, st));
    MD_HIP(hipStreamSynchronize(st));
    return 0;
  }
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  for (size_t w0 = 0; w0 < payload; w0 += kStageBytes) {
    const size_t w1 = std::min(payload, w0 + kStageBytes);
    // Device-to-pinned copies of the window's pieces (stream offset - w0 in the staging).
    size_t base = 0;
    for (const HostCopy &c : parts) {
      const size_t a = std::max(w0, base), b = std::min(w1, base + c.bytes);
      if (a < b)
        MD_HIP(hipMemcpyAsync(static_cast<char *>(stage.p) + (a - w0), static_cast<const char *>(c.src) + (a - base),
                              b - a, hipMemcpyDeviceToHost, st));
      base += c.bytes;
    }
    MD_HIP(hipStreamSynchronize(st));
    auto spread = [&](size_t lo, size_t hi) {  // stream bytes [lo, hi) of this window
      size_t at = 0;
      for (const HostCopy &c : parts) {
        const size_t a = std::max(lo, at), b = std::min(hi, at + c.bytes);
        if (a < b)
          std::memcpy(static_cast<char *>(c.dst) + (a - at), static_cast<const char *>(stage.p) + (a - w0), b - a);
        at += c.bytes;
      }
    };
    const size_t n = w1 - w0;
    const int nt = static_cast<int>(std::min<size_t>(std::min(8u, hw), std::max<size_t>(1, n >> 23)));
    if (nt <= 1) {
      spread(w0, w1);
      continue;
    }
    std::vector<std::thread> pool;
    const size_t step = (n + nt - 1) / nt;
    size_t done = w0 + step;  // the calling thread takes [w0, w0 + step); a thread that cannot start is done here
    for (int t = 1; t < nt; ++t) {
      const size_t a = std::min(w1, w0 + t * step), b = std::min(w1, w0 + (t + 1) * step);
      try {
        pool.emplace_back(spread, a, b);
      } catch (...) {
        spread(a, b);
      }
    }
    spread(w0, std::min(w1, done));
    for (std::thread &th : pool) th.join();
  }
  return 0;
}

constexpr int kStatCount = mdr::kStats;  // see md_rocm_last_stats

// Persistent consumer blocks per CU when the configuration leaves it at 0.
// More blocks than a CU holds resident at once: a block that ends early is
// replaced by a waiting one, which evens out the uneven per-block work of the
// lane-persistent kernels (16 measured fastest on the precision workload's
// recheck, probe and refinement kernels among 4..48).
constexpr int kDefaultConsumerBlocks = 16;

int consumer_blocks_per_cu(const md_rocm_config &c) {
  return c.consumer_blocks > 0 ? c.consumer_blocks : kDefaultConsumerBlocks;
}

}  // namespace

// The fixed device workspace of md_rocm_run_device_static: one array per GPU,
// held until the process ends. In host code `scratch` names only a host-side
// stand-in; the device address of each GPU's copy comes from
// hipGetSymbolAddress, called once per device with that device current
// (scratch_setup, from md_rocm_create), and host code uses that address only.
constexpr size_t kScratchBytes = size_t(MD_ROCM_STATIC_SCRATCH_BYTES);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};
int g_scratch_state[kMaxDevices] = {};  // 0 not set up, 1 address known, -1 lookup failed

void scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_state[device]) return;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  g_scratch_state[device] = e == hipSuccess && ptr ? 1 : -1;
  g_scratch_ptr[device] = e == hipSuccess ? ptr : nullptr;
  if (e != hipSuccess) (void)hipGetLastError();
}

void *scratch_address(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  return g_scratch_state[device] == 1 ? g_scratch_ptr[device] : nullptr;
}
}  // namespace

struct md_rocm_solver {
  md_rocm_config cfg{};
  int capacity = 0, L = 0, nf = 0, nmodes = 0;
  int device = 0;
  int cus = 256;
  hipStream_t stream = nullptr;      // the stream work is issued on
  hipStream_t own_stream = nullptr;  // the solver's own stream
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  DevBuf rows, models, traces, events, scan, brackets, roots, unc_q, tr_q, counters, stats, overflow, phase, tape;
  DevBuf c_anchors, c_targets, c_list, c_roots, c_slopes, c_fallback, c_count, c_pred;
  DevBuf in_h, in_vp, in_vs, in_rho, in_per, in_bounds, in_offsets;
  DevBuf out_v, out_flags, out_levels, out_counts, out_group, out_gflags, out_v32;
  double last_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  std::vector<hipEvent_t> marks;  // kStageMarks per chunk of the last run
  int marked_chunks = 0;
  bool capturing = false;          // recording the device-path sequence into a graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<uintptr_t> graph_key;  // arguments the captured graph was built for
  long long cont_targets = 0, cont_accepted = 0, cont_fallbacks = 0;  // continuation targets of the last run
  long long cont_evaluations = 0;  // FP64 evaluations of the continuation kernel in the last run
  double f_accept = 1.;            // acceptance threshold of the current call (1 = off)
  bool timed = false;              // ev_begin / ev_end were recorded by the last run
  int prep_models = 0;             // models prepared on the device for the current run
  int prep_rows = 0;               // largest row count among them
  DevBuf table, ntab;              // per-model scan tables (adaptive grid)
  DevBuf invalid;                  // models of the last device ragged run with inconsistent offsets
  HostStage host_stage;            // pinned staging of host-array results
  bool invalid_live = false;       // `invalid` belongs to the last run
  int probe_waves[2][2] = {};      // [wave type] resident probe waves (FP32, FP64 pass); 0 if the query failed

  ~md_rocm_solver() {
    for (DevBuf *b : {&rows, &models, &traces, &events, &scan, &brackets, &roots, &unc_q, &tr_q, &counters, &stats,
                      &overflow, &phase, &tape, &c_anchors, &c_targets, &c_list, &c_roots, &c_slopes,
                      &c_fallback, &c_count, &c_pred,
                      &in_h, &in_vp, &in_vs, &in_rho, &in_per, &in_bounds, &in_offsets, &out_v, &out_flags, &out_levels,
                      &out_counts, &out_group, &out_gflags, &out_v32, &table, &ntab, &invalid})
      b->release();
    host_stage.release();
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (hipEvent_t e : marks) (void)hipEventDestroy(e);
    if (ev_begin) (void)hipEventDestroy(ev_begin);
    if (ev_end) (void)hipEventDestroy(ev_end);
    if (own_stream) (void)hipStreamDestroy(own_stream);
  }
};

namespace {

// search_level of a trace the grid scan accepts: the code of the reference
// search its (accuracy, shape) pair selects (docs/API.md:203-209; levels set at
// cpp/cpu/solver.cpp:300 and :307, the adaptive ladder at :258-283).
int grid_level(const md_rocm_config &c) {
  const bool fast = c.accuracy == 2, approximate = c.accuracy != 0;
  if (c.shape) return !approximate ? mdr::kLevelShapeHigh : fast ? mdr::kLevelShapeFast : mdr::kLevelShapeBalanced;
  return !approximate ? mdr::kLevelAdaptive : fast ? mdr::kLevelObservationFast : mdr::kLevelObservationBalanced;
}

// The f_accept of an md_rocm_run call, held by the solver for the call's
// duration (the device entry points run with the test off).
struct AcceptScope {
  md_rocm_solver *s;
  AcceptScope(md_rocm_solver *solver, double accept) : s(solver) { s->f_accept = accept; }
  ~AcceptScope() { s->f_accept = 1.; }
};

uintptr_t accept_bits(const md_rocm_solver *s) {
  uintptr_t b;
  std::memcpy(&b, &s->f_accept, sizeof(b));
  return b;
}

int chunk_size(const md_rocm_solver *s, long long ntraces) {
  long long c = s->cfg.chunk_traces > 0 ? s->cfg.chunk_traces : (1LL << 20);
  if (c > ntraces) c = ntraces;
  if (c < 1) c = 1;
  return static_cast<int>(c);
}

// Device buffers of one run: the prepared rows, models and traces, the
// per-model scan tables, and the per-chunk work buffers of the pipeline.
struct Work {
  mdr::PrepRow *rows = nullptr;
  mdr::ModelInfo *models = nullptr;
  mdr::TraceInfo *traces = nullptr;
  mdr::TablePoint *table = nullptr;
  int *ntab = nullptr;
  mdr::Event *events = nullptr;
  mdr::TraceScan *scan = nullptr;
  mdr::Bracket *brackets = nullptr;
  double *roots = nullptr;
  uint32_t *unc_q = nullptr, *tr_q = nullptr;
  int *counters = nullptr, *overflow = nullptr, *phase = nullptr;
  unsigned long long *stats = nullptr;
  int chunk = 0;           // traces per chunk
  hipEvent_t *marks = nullptr;  // stage events per chunk (nullptr: no stage timing)
};

size_t table_stride(int max_rows) { return mdr::kFloor + 2 * std::max(1, max_rows - 1); }

// The solver's own buffers for a run of ntraces traces (grown on demand).
int solver_work(md_rocm_solver *s, long long ntraces, int kmax, Work &w) {
  const int chunk = chunk_size(s, ntraces);
  const size_t qcap = static_cast<size_t>(chunk) * mdr::kEventCap;
  if (int e = s->events.ensure(qcap * sizeof(mdr::Event))) return e;
  if (int e = s->scan.ensure(static_cast<size_t>(chunk) * sizeof(mdr::TraceScan))) return e;
  if (int e = s->brackets.ensure(static_cast<size_t>(chunk) * kmax * sizeof(mdr::Bracket))) return e;
  if (int e = s->roots.ensure(static_cast<size_t>(chunk) * kmax * sizeof(double))) return e;
  if (int e = s->unc_q.ensure(qcap * sizeof(uint32_t))) return e;
  if (int e = s->tr_q.ensure(qcap * sizeof(uint32_t))) return e;
  if (int e = s->counters.ensure(5 * sizeof(int))) return e;
  if (int e = s->overflow.ensure(static_cast<size_t>(chunk) * sizeof(int))) return e;
  if (int e = s->phase.ensure(static_cast<size_t>(chunk) * sizeof(int))) return e;
  if (int e = s->stats.ensure(kStatCount * sizeof(unsigned long long))) return e;
  if (s->cfg.grid) {
    const size_t stride = table_stride(s->prep_rows), nm = static_cast<size_t>(std::max(1, s->prep_models));
    if (int e = s->table.ensure(nm * stride * sizeof(mdr::TablePoint))) return e;
    if (int e = s->ntab.ensure(nm * sizeof(int))) return e;
  }
  const int nchunks = static_cast<int>((ntraces + chunk - 1) / chunk);
  while (static_cast<int>(s->marks.size()) < nchunks * mdr::kStageMarks) {
    hipEvent_t e;
    MD_HIP(hipEventCreate(&e));
    s->marks.push_back(e);
  }
  w.rows = s->rows.as<mdr::PrepRow>();
  w.models = s->models.as<mdr::ModelInfo>();
  w.traces = s->traces.as<mdr::TraceInfo>();
  w.table = s->table.as<mdr::TablePoint>();
  w.ntab = s->ntab.as<int>();
  w.events = s->events.as<mdr::Event>();
  w.scan = s->scan.as<mdr::TraceScan>();
  w.brackets = s->brackets.as<mdr::Bracket>();
  w.roots = s->roots.as<double>();
  w.unc_q = s->unc_q.as<uint32_t>();
  w.tr_q = s->tr_q.as<uint32_t>();
  w.counters = s->counters.as<int>();
  w.overflow = s->overflow.as<int>();
  w.phase = s->phase.as<int>();
  w.stats = s->stats.as<unsigned long long>();
  w.chunk = chunk;
  w.marks = s->capturing ? nullptr : s->marks.data();
  return 0;
}

// Runs the root pipeline over all traces already prepared on the device.
int run_pipeline(md_rocm_solver *s, const Work &w, long long ntraces, int kmax, int max_rows, int wave, double tol,
                 double *v64, float *v32, uint32_t *flags, int32_t *levels, int32_t *counts, double *group,
                 uint32_t *gflags, const mdr::TraceInfo *list = nullptr, double *all_roots = nullptr) {
  const int chunk = w.chunk;
  const size_t qcap = static_cast<size_t>(chunk) * mdr::kEventCap;
  s->invalid_live = false;
  // Per-model scan tables for the adaptive grid (one wavefront per model).
  const int stride = static_cast<int>(table_stride(s->prep_rows));
  if (s->cfg.grid)
    MD_HIP(mdr::launch_tables(w.models, w.rows, s->prep_models, s->prep_rows, wave, stride, w.table, w.ntab,
                              s->stream));
  mdr::PipeArgs a{};
  a.table = w.table;
  a.ntab = w.ntab;
  a.table_stride = stride;
  a.traces = list ? list : w.traces;
  a.all_roots = all_roots;
  a.models = w.models;
  a.rows = w.rows;
  a.events = w.events;
  a.scan = w.scan;
  a.brackets = w.brackets;
  a.roots = w.roots;
  a.unc_queue = w.unc_q;
  a.trough_queue = w.tr_q;
  a.counters = w.counters;
  a.unc_count = a.counters;
  a.trough_count = a.counters + 1;
  a.overflow_count = a.counters + 2;
  a.overflow_list = w.overflow;
  a.phase_count = a.counters + 3;
  a.handoff_count = a.counters + 4;
  a.handoff_list = w.unc_q;
  a.phase_list = w.phase;
  a.queue_cap = static_cast<int>(qcap);
  a.nc = s->cfg.nc;
  a.max_rows = max_rows;
  a.scan_waves = s->cfg.scan_waves > 0 ? s->cfg.scan_waves : 4;
  a.adaptive = s->cfg.grid;
  a.rho = static_cast<float>(s->cfg.rho);
  a.kmax = kmax;
  a.wave = wave;
  a.fp32 = s->cfg.single_precision ? 1 : 0;
  a.refinement = s->cfg.refinement;
  a.level_grid = grid_level(s->cfg);
  a.close_grid = s->cfg.shape || s->cfg.accuracy != 0 ? 1 : 0;
  a.love_phase = s->cfg.love_phase;
  a.probe_waves[0] = s->probe_waves[wave][0];
  a.probe_waves[1] = s->probe_waves[wave][1];
  a.trace_extra = s->cfg.accuracy != 0 ? mdr::kFlagApproximate : 0u;
  a.f_accept = s->f_accept;
  a.tol = tol;
  a.close_rtol = s->cfg.close_root_rtol;
  a.group_rtol = s->cfg.group_rtol;
  a.velocity64 = v64;
  a.velocity32 = v32;
  a.flags = flags;
  a.levels = levels;
  a.counts = counts;
  a.group = group;
  a.group_flags = gflags;
  a.stats = w.stats;
  const int persistent = s->cus * consumer_blocks_per_cu(s->cfg);
  const int nchunks = static_cast<int>((ntraces + chunk - 1) / chunk);
  s->marked_chunks = w.marks ? nchunks : 0;
  int c = 0;
  for (long long begin = 0; begin < ntraces; begin += chunk, ++c) {
    a.trace_begin = static_cast<int>(begin);
    a.trace_end = static_cast<int>(std::min<long long>(ntraces, begin + chunk));
    const hipError_t err = mdr::launch_pipeline(a, a, a.fp32 != 0, group != nullptr, persistent, s->stream,
                                                w.marks ? w.marks + c * mdr::kStageMarks : nullptr);
    if (err != hipSuccess) return fail(1, std::string("kernel launch: ") + hipGetErrorString(err));
  }
  return 0;
}

// run_pipeline on the solver's own buffers.
int solver_pipeline(md_rocm_solver *s, long long ntraces, int kmax, int max_rows, int wave, double tol, double *v64,
                    uint32_t *flags, int32_t *levels, int32_t *counts, double *group, uint32_t *gflags) {
  Work w;
  if (int e = solver_work(s, ntraces, kmax, w)) return e;
  return run_pipeline(s, w, ntraces, kmax, max_rows, wave, tol, v64, nullptr, flags, levels, counts, group, gflags);
}

// Work buffers carved from the fixed device workspace at `base` (bytes long)
// for B models of L rows and T periods: the prepared inputs and tables first,
// then as many traces per chunk as the rest holds. Returns the bytes needed
// for one trace per chunk when even that does not fit (0 on success).
size_t carve_work(const md_rocm_solver *s, unsigned char *base, size_t bytes, int B, int L, int T, int kmax,
                  Work &w) {
  const auto up = [](size_t n) { return (n + 255) & ~size_t(255); };
  const size_t nb = static_cast<size_t>(B), ntr = nb * T, stride = table_stride(L);
  size_t fixed = up(nb * L * sizeof(mdr::PrepRow)) + up(nb * sizeof(mdr::ModelInfo)) +
                 up(ntr * sizeof(mdr::TraceInfo)) + up(5 * sizeof(int)) + up(kStatCount * sizeof(unsigned long long));
  if (s->cfg.grid) fixed += up(nb * stride * sizeof(mdr::TablePoint)) + up(nb * sizeof(int));
  const size_t per_trace = mdr::kEventCap * (sizeof(mdr::Event) + 2 * sizeof(uint32_t)) + sizeof(mdr::TraceScan) +
                           static_cast<size_t>(kmax) * (sizeof(mdr::Bracket) + sizeof(double)) + 2 * sizeof(int);
  const size_t slack = 9 * 256;  // alignment of the nine per-chunk arrays
  if (fixed + slack + per_trace > bytes) return fixed + slack + per_trace;
  size_t chunk = (bytes - fixed - slack) / per_trace;
  if (s->cfg.chunk_traces > 0) chunk = std::min(chunk, static_cast<size_t>(s->cfg.chunk_traces));
  chunk = std::min(chunk, std::min(ntr, size_t(1) << 20));
  size_t off = 0;
  const auto take = [&](size_t n) {
    unsigned char *p = base + off;
    off += up(n);
    return static_cast<void *>(p);
  };
  w.rows = static_cast<mdr::PrepRow *>(take(nb * L * sizeof(mdr::PrepRow)));
  w.models = static_cast<mdr::ModelInfo *>(take(nb * sizeof(mdr::ModelInfo)));
  w.traces = static_cast<mdr::TraceInfo *>(take(ntr * sizeof(mdr::TraceInfo)));
  w.counters = static_cast<int *>(take(5 * sizeof(int)));
  w.stats = static_cast<unsigned long long *>(take(kStatCount * sizeof(unsigned long long)));
  if (s->cfg.grid) {
    w.table = static_cast<mdr::TablePoint *>(take(nb * stride * sizeof(mdr::TablePoint)));
    w.ntab = static_cast<int *>(take(nb * sizeof(int)));
  }
  const size_t qcap = chunk * mdr::kEventCap;
  w.events = static_cast<mdr::Event *>(take(qcap * sizeof(mdr::Event)));
  w.unc_q = static_cast<uint32_t *>(take(qcap * sizeof(uint32_t)));
  w.tr_q = static_cast<uint32_t *>(take(qcap * sizeof(uint32_t)));
  w.scan = static_cast<mdr::TraceScan *>(take(chunk * sizeof(mdr::TraceScan)));
  w.brackets = static_cast<mdr::Bracket *>(take(chunk * kmax * sizeof(mdr::Bracket)));
  w.roots = static_cast<double *>(take(chunk * kmax * sizeof(double)));
  w.overflow = static_cast<int *>(take(chunk * sizeof(int)));
  w.phase = static_cast<int *>(take(chunk * sizeof(int)));
  w.chunk = static_cast<int>(chunk);
  w.marks = nullptr;
  return off <= bytes ? 0 : off;
}

// Anchor / target plan of one model's periods (reference continuation.hpp:12-20,
// cpu/continuation.hpp:13-24). Models with fewer than 8 periods or periods
// that are not strictly increasing are solved independently (all anchors).
double cont_log_step(double target, double anchor) {
  if (std::fabs(target - anchor) < .5 * anchor) return std::log1p((target - anchor) / anchor);
  return std::log(target) - std::log(anchor);
}

void plan_model(const double *p, int T, long long base, int stride, std::vector<int> &anchors,
                std::vector<mdr::TargetPlan> &targets, bool &bypassed) {
  bool ok = T >= 8;
  for (int f = 1; f < T && ok; ++f) ok = p[f] > p[f - 1];
  bypassed = !ok;
  if (!ok) {
    for (int f = 0; f < T; ++f) anchors.push_back(static_cast<int>(base + f));
    return;
  }
  for (int left = 0;;) {
    const int slot = static_cast<int>(anchors.size());
    anchors.push_back(static_cast<int>(base + left));
    if (left == T - 1) break;
    int right = left + std::min(stride, T - 1 - left);
    while (right > left + 1 && cont_log_step(p[right], p[left]) > .25) --right;
    for (int f = left + 1; f < right; ++f)
      targets.push_back(mdr::TargetPlan{static_cast<int>(base + f), slot, slot + 1, cont_log_step(p[f], p[left]),
                                        cont_log_step(p[f], p[right])});
    left = right;
  }
}

// Anchors through the full pipeline, targets by continuation, failed targets
// through the full pipeline again. Outputs and prediction flags are device
// arrays in the run's layout; traces/models/rows are already on the device.
int run_continuation(md_rocm_solver *s, int kmax, int max_rows, int wave, double tol, double *v64, uint32_t *flags,
                     int32_t *levels, int32_t *counts, double *group, uint32_t *gflags, uint32_t *prediction,
                     const std::vector<int> &anchors, const std::vector<mdr::TargetPlan> &targets) {
  const int na = static_cast<int>(anchors.size()), nt = static_cast<int>(targets.size());
  hipStream_t st = s->stream;
  const size_t list_n = static_cast<size_t>(std::max(na, nt));
  if (int e = s->c_anchors.ensure(std::max<size_t>(1, na) * sizeof(int))) return e;
  if (int e = s->c_targets.ensure(std::max<size_t>(1, nt) * sizeof(mdr::TargetPlan))) return e;
  if (int e = s->c_list.ensure(std::max<size_t>(1, list_n) * sizeof(mdr::TraceInfo))) return e;
  if (int e = s->c_roots.ensure(std::max<size_t>(1, na) * kmax * sizeof(double))) return e;
  if (int e = s->c_slopes.ensure(std::max<size_t>(1, na) * kmax * sizeof(double))) return e;
  if (int e = s->c_fallback.ensure(std::max<size_t>(1, nt) * sizeof(int))) return e;
  if (int e = s->c_count.ensure(2 * sizeof(unsigned long long))) return e;  // fallback count, evaluations
  MD_HIP(hipMemcpyAsync(s->c_anchors.p, anchors.data(), na * sizeof(int), hipMemcpyHostToDevice, st));
  if (nt) MD_HIP(hipMemcpyAsync(s->c_targets.p, targets.data(), nt * sizeof(mdr::TargetPlan), hipMemcpyHostToDevice, st));
  const unsigned long long zero[2] = {0ull, 0ull};
  MD_HIP(hipMemcpyAsync(s->c_count.p, zero, sizeof(zero), hipMemcpyHostToDevice, st));
  MD_HIP(mdr::launch_gather(s->traces.as<mdr::TraceInfo>(), s->c_anchors.as<int>(), na,
                            s->c_list.as<mdr::TraceInfo>(), st));
  Work wa;
  if (int e = solver_work(s, na, kmax, wa)) return e;
  if (int e = run_pipeline(s, wa, na, kmax, max_rows, wave, tol, v64, nullptr, flags, levels, counts, group, gflags,
                           s->c_list.as<mdr::TraceInfo>(), s->c_roots.as<double>()))
    return e;
  mdr::ContArgs c{};
  c.traces = s->traces.as<mdr::TraceInfo>();
  c.models = s->models.as<mdr::ModelInfo>();
  c.rows = s->rows.as<mdr::PrepRow>();
  c.anchors = s->c_anchors.as<int>();
  c.targets = s->c_targets.as<mdr::TargetPlan>();
  c.ntargets = nt;
  c.nanchors = na;
  c.kmax = kmax;
  c.wave = wave;
  c.refinement = s->cfg.refinement;
  c.anchor_roots = s->c_roots.as<double>();
  c.anchor_slopes = s->c_slopes.as<double>();
  c.tol = tol;
  c.close_rtol = s->cfg.close_root_rtol;
  c.f_accept = s->f_accept;
  c.group_rtol = s->cfg.group_rtol;
  c.velocity = v64;
  c.flags = flags;
  c.levels = levels;
  c.counts = counts;
  c.group = group;
  c.group_flags = gflags;
  c.prediction = prediction;
  c.fallback = s->c_fallback.as<int>();
  c.fallback_count = s->c_count.as<int>();
  c.evaluations = s->c_count.as<unsigned long long>() + 1;
  const int pb = s->cus * consumer_blocks_per_cu(s->cfg);
  MD_HIP(mdr::launch_continuation(c, pb, st));
  unsigned long long tally[2] = {0ull, 0ull};
  MD_HIP(hipMemcpyAsync(tally, s->c_count.p, sizeof(tally), hipMemcpyDeviceToHost, st));
  MD_HIP(hipStreamSynchronize(st));
  const int nfb = static_cast<int>(tally[0] & 0xffffffffull);
  s->cont_targets = nt;
  s->cont_accepted = nt - nfb;
  s->cont_fallbacks = nfb;
  s->cont_evaluations = static_cast<long long>(tally[1]);
  if (nfb > 0) {
    MD_HIP(mdr::launch_gather(s->traces.as<mdr::TraceInfo>(), s->c_fallback.as<int>(), nfb,
                              s->c_list.as<mdr::TraceInfo>(), st));
    Work wf;
    if (int e = solver_work(s, nfb, kmax, wf)) return e;
    if (int e = run_pipeline(s, wf, nfb, kmax, max_rows, wave, tol, v64, nullptr, flags, levels, counts, group, gflags,
                             s->c_list.as<mdr::TraceInfo>(), nullptr))
      return e;
  }
  return 0;
}

int check_config(const md_rocm_config &c) {
  if (c.nc < 3 || c.nc > 65536) return fail(2, "nc must be 3..65536");
  if (c.grid != 0 && c.grid != 1) return fail(2, "grid must be 0 (uniform) or 1 (adaptive)");
  if (!(c.rho >= 1. && c.rho <= 1024.)) return fail(2, "rho must be 1..1024 nodes per mode");
  if (!(c.close_root_rtol > 0. && c.close_root_rtol < 1.)) return fail(2, "close_root_rtol must be in (0, 1)");
  if (!(c.group_rtol > 0. && c.group_rtol < 1.)) return fail(2, "group_rtol must be in (0, 1)");
  if (c.continuation && c.anchor_stride < 2) return fail(2, "anchor_stride must be at least 2");
  if (!(c.tau_floor >= 0.) || !(c.tau_phase >= 0.)) return fail(2, "FP32 band parameters must be non-negative");
  if (c.graph != 0 && c.graph != 1) return fail(2, "graph must be 0 or 1");
  if (c.scan_waves != 0 && c.scan_waves != 1 && c.scan_waves != 2 && c.scan_waves != 4 && c.scan_waves != 8)
    return fail(2, "scan_waves must be 0, 1, 2, 4 or 8");
  if (c.consumer_blocks < 0 || c.consumer_blocks > 64) return fail(2, "consumer_blocks must be 0..64");
  if (c.refinement < 0 || c.refinement > 2) return fail(2, "refinement must be 0 (legacy), 1 (brent) or 2 (chandrupatla)");
  if (c.accuracy != 0 && c.accuracy != 1 && c.accuracy != 2 && c.accuracy != 4)
    return fail(2, "accuracy must be 0 (high), 1 (balanced), 2 (fast) or 4 (auto); 3 (uniform) is the legacy "
                   "path: set grid=0");
  if (c.shape != 0 && c.shape != 1) return fail(2, "shape must be 0 or 1");
  if (c.love_phase != 0 && c.love_phase != 1) return fail(2, "invalid refinement or Love strategy");
  if (c.secular != 1)
    return fail(2, "secular must be 1 (fast): the released 25-entry Dunkin evaluator (secular=0) is not built on "
                   "this backend");
  if (!(c.sampling_rtol == -1. || (c.sampling_rtol > 0. && c.sampling_rtol <= 1.)))
    return fail(2, "sampling_rtol must be finite and between 0 and 1 (or -1 for automatic sampling)");
  if (!(c.period_threshold >= 0.) || !std::isfinite(c.period_threshold))
    return fail(2, "period_threshold must be finite and non-negative");
  return 0;
}

}  // namespace

extern "C" {

void md_rocm_default_config(md_rocm_config *cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->accuracy = 0;
  cfg->close_root_rtol = .005;
  cfg->single_precision = 1;
  cfg->group_velocity = 0;
  cfg->group_rtol = .001;
  cfg->continuation = 0;
  cfg->anchor_stride = 4;
  cfg->nc = 512;
  cfg->tau_floor = 1e-4;
  cfg->tau_phase = 5e-5;
  cfg->chunk_traces = 0;
  cfg->grid = 1;
  cfg->rho = 8.;
  cfg->refinement = 1;
  cfg->shape = 1;
  cfg->love_phase = 1;
  cfg->secular = 1;
  cfg->sampling_rtol = 1e-3;
  cfg->period_threshold = 0.;
}

const char *md_rocm_last_error(void) { return g_error.c_str(); }

int md_rocm_available(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return 0;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) return 0;
  return std::strncmp(p.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

int md_rocm_set_device(int device) {
  MD_HIP(hipSetDevice(device));
  return 0;
}

int md_rocm_memory_info(size_t *free_bytes, size_t *total_bytes) {
  MD_HIP(hipMemGetInfo(free_bytes, total_bytes));
  return 0;
}

int md_rocm_device_name(int device, char *buf, int len) {
  hipDeviceProp_t p;
  MD_HIP(hipGetDeviceProperties(&p, device));
  if (buf && len > 0) std::snprintf(buf, static_cast<size_t>(len), "%s (%s)", p.name, p.gcnArchName);
  return 0;
}

int md_rocm_review_groups(const double *packed, size_t size, int layers, int modes, int wave, int secular,
                          double period, double lower, double upper, double group_rtol, double close_rtol,
                          double *group, uint32_t *flags, int *evaluations) {
  if (!packed || !group || !flags || layers < 1 || layers > mdr::kMaxRows || modes < 1 || modes > 256 ||
      size < static_cast<size_t>(4) * layers || size > static_cast<size_t>(4) * layers + 256 || !(period > 0.) ||
      (wave != 0 && wave != 1))
    return fail(2, "invalid review trace");
  if (secular != 1) return fail(2, "secular must be 1 (fast): the only evaluator built on this backend");
  const int known = static_cast<int>(size - static_cast<size_t>(4) * layers);
  const double *h = packed, *vp = packed + layers, *vs = packed + 2 * layers, *rho = packed + 3 * layers;
  std::vector<mdr::PrepRow> rows(layers);
  for (int j = 0; j < layers; ++j) rows[j] = mdr::make_prep_row(h[j], vp[j], vs[j], rho[j]);
  const size_t bytes = layers * sizeof(mdr::PrepRow) + (known + 2 * modes) * sizeof(double) + modes * 4 + 256;
  DevBuf buf;
  if (int e = buf.ensure(bytes)) return e;
  auto *base = static_cast<unsigned char *>(buf.p);
  auto *d_rows = reinterpret_cast<mdr::PrepRow *>(base);
  auto *d_roots = reinterpret_cast<double *>(base + layers * sizeof(mdr::PrepRow));
  auto *d_group = d_roots + std::max(known, 1);
  auto *d_flags = reinterpret_cast<uint32_t *>(d_group + modes);
  MD_HIP(hipMemcpy(d_rows, rows.data(), layers * sizeof(mdr::PrepRow), hipMemcpyHostToDevice));
  if (known) MD_HIP(hipMemcpy(d_roots, packed + 4 * layers, known * sizeof(double), hipMemcpyHostToDevice));
  const int top = vs[0] == 0. ? 1 : 0;
  constexpr double two_pi = 6.283185307179586;
  MD_HIP(mdr::launch_review_groups(d_rows, layers, top, wave, two_pi / period, d_roots, known, modes, lower, upper,
                                   close_rtol, group_rtol, d_group, d_flags, nullptr));
  MD_HIP(hipMemcpy(group, d_group, modes * sizeof(double), hipMemcpyDeviceToHost));
  MD_HIP(hipMemcpy(flags, d_flags, modes * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (evaluations) {  // one dual-number evaluation per root that reached the derivative
    int n = 0;
    for (int m = 0; m < modes; ++m) n += m < known && flags[m] != mdr::kGroupMissing && flags[m] != mdr::kGroupClose &&
                                         flags[m] != mdr::kGroupPolish;
    *evaluations = n;
  }
  return 0;
}

static int fill_report(int which, int L, int single_precision, const md_rocm_config *cfg, md_rocm_kernel_report *out) {
  if (L < 1 || L > mdr::kMaxRows) return fail(2, "L must be 1..256");
  if (!out) return fail(2, "null report");
  mdr::KernelReport r{};
  const int waves = cfg && cfg->scan_waves > 0 ? cfg->scan_waves : 4;
  MD_HIP(mdr::kernel_report(which, L, single_precision != 0, waves, &r));
  out->registers_per_thread = r.registers;
  out->local_bytes_per_thread = r.local_bytes;
  out->static_shared_bytes = r.static_shared;
  out->dynamic_shared_bytes = r.dynamic_shared;
  out->search_block_threads = r.block;
  out->max_active_blocks_per_cu = r.active_per_cu;
  out->configured_blocks_per_cu = cfg ? consumer_blocks_per_cu(*cfg) : kDefaultConsumerBlocks;
  out->table_capacity_per_model = r.table_rows;
  out->cache_points_per_trace = r.events_per_trace;
  return 0;
}

int md_rocm_qe_kernel_report(int L, int accuracy, int shape, int single_precision, const md_rocm_config *cfg,
                             md_rocm_kernel_report *out) {
  (void)accuracy;
  (void)shape;
  return fill_report(0, L, single_precision, cfg, out);
}

int md_rocm_love_phase_kernel_report(int L, int single_precision, md_rocm_kernel_report *out) {
  return fill_report(1, L, single_precision, nullptr, out);
}

md_rocm_solver *md_rocm_create(int capacity, int L, int nf, int nmodes, const md_rocm_config *cfg) {
  md_rocm_config c;
  md_rocm_default_config(&c);
  if (cfg) c = *cfg;
  if (capacity < 1 || L < 1 || L > mdr::kMaxRows || nf < 1 || nmodes < 1 || nmodes > 256) {
    fail(2, "invalid solver geometry (capacity >= 1, 1 <= L <= 256, nf >= 1, 1 <= nmodes <= 256)");
    return nullptr;
  }
  if (check_config(c)) return nullptr;
  auto *s = new md_rocm_solver();
  s->cfg = c;
  s->capacity = capacity;
  s->L = L;
  s->nf = nf;
  s->nmodes = nmodes;
  if (hipGetDevice(&s->device) != hipSuccess) {
    delete s;
    fail(1, "no HIP device");
    return nullptr;
  }
  scratch_setup(s->device);
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, s->device) == hipSuccess) s->cus = p.multiProcessorCount;
  // Resident probe waves per wave type (an occupancy query, made here so that
  // no run, captured or not, has to make it).
  for (int w = 0; w < 2; ++w)
    if (mdr::probe_residency(s->cus, w, s->probe_waves[w]) != hipSuccess)
      s->probe_waves[w][0] = s->probe_waves[w][1] = 0;
  if (hipStreamCreateWithFlags(&s->own_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&s->ev_begin) != hipSuccess || hipEventCreate(&s->ev_end) != hipSuccess) {
    delete s;
    fail(1, "stream/event creation failed");
    return nullptr;
  }
  s->stream = s->own_stream;
  return s;
}

void md_rocm_free(md_rocm_solver *s) { delete s; }

int md_rocm_synchronize(md_rocm_solver *s) {
  if (!s) return fail(2, "null solver");
  MD_HIP(hipStreamSynchronize(s->stream));
  float ms = 0.f;
  for (double &x : s->last_ms) x = 0.;
  // Only events the last run recorded are read: a query of an event never
  // recorded fails and would leave that error as the thread's last HIP error.
  if (s->timed && hipEventElapsedTime(&ms, s->ev_begin, s->ev_end) == hipSuccess) s->last_ms[0] = ms;
  for (int c = 0; c < s->marked_chunks; ++c) {
    hipEvent_t *m = s->marks.data() + c * mdr::kStageMarks;
    for (int k = 0; k + 1 < mdr::kStageMarks; ++k)
      if (hipEventElapsedTime(&ms, m[k], m[k + 1]) == hipSuccess) s->last_ms[1 + k] += ms;
  }
  (void)hipGetLastError();
  return 0;
}

int md_rocm_last_timing(md_rocm_solver *s, double *ms, int n) {
  if (!s || !ms) return 0;
  const int k = std::min(n, 7);
  for (int i = 0; i < k; ++i) ms[i] = s->last_ms[i];
  return k;
}

int md_rocm_last_stats(md_rocm_solver *s, uint64_t *stats, int n) {
  if (!s || !stats || n < 1) {
    fail(2, "null solver or stats");
    return -1;
  }
  const int k = std::min(n, kStatCount + 4);
  unsigned long long st[kStatCount + 4] = {};
  int invalid = 0;
  if (s->stats.p && hipMemcpyAsync(st, s->stats.p, kStatCount * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                   s->stream) != hipSuccess) {
    fail(1, "stats copy failed");
    return -1;
  }
  if (s->invalid_live && hipMemcpyAsync(&invalid, s->invalid.p, sizeof(int), hipMemcpyDeviceToHost, s->stream) !=
                             hipSuccess) {
    fail(1, "stats copy failed");
    return -1;
  }
  if (hipStreamSynchronize(s->stream) != hipSuccess) {
    fail(1, "stats copy failed");
    return -1;
  }
  st[kStatCount] = static_cast<unsigned long long>(invalid);
  st[kStatCount + 1] = static_cast<unsigned long long>(s->cont_targets);
  st[kStatCount + 2] = static_cast<unsigned long long>(s->cont_accepted);
  st[kStatCount + 3] = static_cast<unsigned long long>(s->cont_evaluations);
  for (int i = 0; i < k; ++i) stats[i] = st[i];
  return k;
}

int md_rocm_set_stream(md_rocm_solver *s, void *stream, int own) {
  if (!s) return fail(2, "null solver");
  s->stream = own ? s->own_stream : static_cast<hipStream_t>(stream);
  return 0;
}

int md_rocm_adjoint(md_rocm_solver *s, const double *thickness, const double *vp, const double *vs,
                    const double *rho, int B, const double *periods, int row_periods, int wave,
                    const double *roots, const double *weight, double *grad_thickness, double *grad_vp,
                    double *grad_vs, double *grad_rho) {
  if (!s) return fail(2, "null solver");
  if (B < 1 || B > s->capacity) return fail(2, "B must be 1..capacity");
  if (wave != 0 && wave != 1) return fail(2, "wave must be 0 (Rayleigh) or 1 (Love)");
  if (!thickness || !vp || !vs || !rho || !periods || !roots || !weight || !grad_thickness || !grad_vp ||
      !grad_vs || !grad_rho)
    return fail(2, "null device pointer");
  // A block takes `group` models at once so that their roots fill its 64
  // lanes; the group shrinks as the batch falls below the 8 blocks per CU
  // that keep the device full, and its LDS stays within 16 KiB.
  const int group = std::max(1, std::min({8, B / (s->cus * 8), 512 / s->L}));
  const int blocks = std::max(1, std::min((B + group - 1) / group, s->cus * 8));
  const int lanes = blocks * 64;
  if (int e = s->tape.ensure(static_cast<size_t>(lanes) * s->L * 6 * sizeof(double))) return e;
  mdr::AdjointArgs a{};
  a.h = thickness; a.vp = vp; a.vs = vs; a.rho = rho;
  a.periods = periods; a.roots = roots; a.weight = weight;
  a.gh = grad_thickness; a.gvp = grad_vp; a.gvs = grad_vs; a.grho = grad_rho;
  a.tape = s->tape.as<double>();
  a.B = B; a.L = s->L; a.T = s->nf; a.modes = s->nmodes; a.row_periods = row_periods ? 1 : 0; a.wave = wave;
  a.tape_lanes = lanes;
  a.group = group;
  const hipError_t err = mdr::launch_adjoint(a, blocks, s->stream);
  if (err != hipSuccess) return fail(1, std::string("adjoint launch: ") + hipGetErrorString(err));
  return 0;
}

extern "C++" {
namespace {
// Prep + pipeline of a device batch on s->stream (eager, or being captured).
int device_sequence(md_rocm_solver *s, const double *thickness, const double *vp, const double *vs,
                    const double *rho, int B, const double *periods, int row_periods, const double *bounds,
                    double c_tol_rel, int wave, double *velocity, uint32_t *flags, int32_t *levels,
                    int32_t *counts, double *group, uint32_t *group_flags) {
  const int L = s->L, T = s->nf;
  const long long ntraces = static_cast<long long>(B) * T;
  const mdr::PrepParams pp{wave, s->cfg.tau_floor, s->cfg.tau_phase};
  MD_HIP(mdr::launch_prep(thickness, vp, vs, rho, bounds, periods, row_periods, B, L, T, s->nmodes, pp,
                          s->rows.as<mdr::PrepRow>(), s->models.as<mdr::ModelInfo>(), s->traces.as<mdr::TraceInfo>(),
                          s->stream));
  s->prep_models = B;
  s->prep_rows = L;
  Work w;
  if (int e = solver_work(s, ntraces, s->nmodes + 2, w)) return e;
  return run_pipeline(s, w, ntraces, s->nmodes + 2, s->L, wave, c_tol_rel, velocity, nullptr, flags, levels, counts,
                      s->cfg.group_velocity ? group : nullptr, group_flags);
}

// A device-resident launch sequence, replayed from the solver's hipGraph when
// `key` equals the arguments it was captured for. Otherwise the sequence runs
// eagerly and, with cfg.graph, is then recorded once more (every buffer is
// allocated by the eager run) on the solver's own stream for later calls.
template <class Sequence>
int graph_run(md_rocm_solver *s, const std::vector<uintptr_t> &key, Sequence &&sequence) {
  s->timed = true;
  if (s->cfg.graph && s->graph_exec && key == s->graph_key) {
    MD_HIP(hipEventRecord(s->ev_begin, s->stream));
    MD_HIP(hipGraphLaunch(s->graph_exec, s->stream));
    MD_HIP(hipEventRecord(s->ev_end, s->stream));
    s->marked_chunks = 0;
    return 0;
  }
  MD_HIP(hipEventRecord(s->ev_begin, s->stream));
  if (int e = sequence()) return e;
  MD_HIP(hipEventRecord(s->ev_end, s->stream));
  if (!s->cfg.graph) return 0;
  if (s->graph_exec) { (void)hipGraphExecDestroy(s->graph_exec); s->graph_exec = nullptr; }
  if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
  s->graph_key.clear();
  hipStream_t issue = s->stream;
  s->stream = s->own_stream;
  s->capturing = true;
  const hipError_t err = hipStreamBeginCapture(s->own_stream, hipStreamCaptureModeRelaxed);
  const int e = err == hipSuccess ? sequence() : 1;
  hipGraph_t g = nullptr;
  const hipError_t end = err == hipSuccess ? hipStreamEndCapture(s->own_stream, &g) : err;
  s->capturing = false;
  s->stream = issue;
  if (e) {
    if (g) (void)hipGraphDestroy(g);
    return e;
  }
  if (end != hipSuccess) return fail(1, std::string("graph capture: ") + hipGetErrorString(end));
  s->graph = g;
  MD_HIP(hipGraphInstantiate(&s->graph_exec, s->graph, nullptr, nullptr, 0));
  s->graph_key = key;
  return 0;
}

uintptr_t double_bits(double x) {
  uintptr_t b;
  std::memcpy(&b, &x, sizeof(b));
  return b;
}

uintptr_t addr(const void *p) { return reinterpret_cast<uintptr_t>(p); }

// Times the launch-shape candidates (traces per scan block x consumer blocks
// per CU) with `run` (an eager device run) and keeps the fastest.
template <class Run>
int autotune_tiles(md_rocm_solver *s, Run &&run, int *scan_waves, int *consumer_blocks) {
  const md_rocm_config saved = s->cfg;
  s->cfg.graph = 0;
  const int waves[] = {1, 2, 4, 8};
  const int blocks[] = {8, 16, 24};
  double best = 1e300;
  int bw = 4, bb = kDefaultConsumerBlocks;
  for (int w : waves) {
    if (w == 8 && !s->cfg.single_precision) continue;
    for (int b : blocks) {
      s->cfg.scan_waves = w;
      s->cfg.consumer_blocks = b;
      double ms = 0.;
      for (int rep = 0; rep < 2; ++rep) {  // the first pass warms up, the second is timed
        if (int e = run()) {
          s->cfg = saved;
          return e;
        }
        MD_HIP(hipStreamSynchronize(s->stream));
        float t = 0.f;
        MD_HIP(hipEventElapsedTime(&t, s->ev_begin, s->ev_end));
        ms = t;
      }
      if (ms < best) { best = ms; bw = w; bb = b; }
    }
  }
  s->cfg = saved;
  s->cfg.scan_waves = bw;
  s->cfg.consumer_blocks = bb;
  if (scan_waves) *scan_waves = bw;
  if (consumer_blocks) *consumer_blocks = bb;
  return 0;
}
}  // namespace
}  // extern "C++"

int md_rocm_run_device(md_rocm_solver *s, const double *thickness, const double *vp, const double *vs,
                       const double *rho, int B, const double *periods, int row_periods, const double *bounds,
                       double c_tol_rel, int wave, double *velocity, uint32_t *flags, int32_t *levels,
                       int32_t *counts, double *group, uint32_t *group_flags) {
  if (!s) return fail(2, "null solver");
  if (B < 1 || B > s->capacity) return fail(2, "B must be 1..capacity");
  if (wave != 0 && wave != 1) return fail(2, "wave must be 0 (Rayleigh) or 1 (Love)");
  if (!(c_tol_rel > 0.) || !std::isfinite(c_tol_rel)) return fail(2, "c_tol_rel must be positive");
  if (!thickness || !vp || !vs || !rho || !periods || !bounds || !velocity || !flags || !levels || !counts)
    return fail(2, "null input or output pointer");
  if (s->cfg.group_velocity && (!group || !group_flags)) return fail(2, "group outputs required");
  const int L = s->L, T = s->nf;
  const long long ntraces = static_cast<long long>(B) * T;
  if (int e = s->rows.ensure(static_cast<size_t>(B) * L * sizeof(mdr::PrepRow))) return e;
  if (int e = s->models.ensure(static_cast<size_t>(B) * sizeof(mdr::ModelInfo))) return e;
  if (int e = s->traces.ensure(static_cast<size_t>(ntraces) * sizeof(mdr::TraceInfo))) return e;
  const std::vector<uintptr_t> key = {
      0, addr(thickness), addr(vp), addr(vs), addr(rho), static_cast<uintptr_t>(B), addr(periods),
      static_cast<uintptr_t>(row_periods), addr(bounds), double_bits(c_tol_rel), static_cast<uintptr_t>(wave),
      addr(velocity), addr(flags), addr(levels), addr(counts), addr(group), addr(group_flags),
      static_cast<uintptr_t>(s->cfg.scan_waves), static_cast<uintptr_t>(s->cfg.consumer_blocks), accept_bits(s)};
  return graph_run(s, key, [&] {
    return device_sequence(s, thickness, vp, vs, rho, B, periods, row_periods, bounds, c_tol_rel, wave, velocity,
                           flags, levels, counts, group, group_flags);
  });
}

int md_rocm_run_device_static(md_rocm_solver *s, const double *thickness, const double *vp, const double *vs,
                              const double *rho, int B, const double *periods, int row_periods,
                              const double *bounds, double c_tol_rel, int wave, double *velocity, uint32_t *flags,
                              int32_t *levels, int32_t *counts, double *group, uint32_t *group_flags) {
  if (!s) return fail(2, "null solver");
  if (B < 1 || B > s->capacity) return fail(2, "B must be 1..capacity");
  if (wave != 0 && wave != 1) return fail(2, "wave must be 0 (Rayleigh) or 1 (Love)");
  if (!(c_tol_rel > 0.) || !std::isfinite(c_tol_rel)) return fail(2, "c_tol_rel must be positive");
  if (!thickness || !vp || !vs || !rho || !periods || !bounds || !velocity || !flags || !levels || !counts)
    return fail(2, "null input or output pointer");
  if (s->cfg.group_velocity && (!group || !group_flags)) return fail(2, "group outputs required");
  auto *base = static_cast<unsigned char *>(scratch_address(s->device));
  if (!base) return fail(1, "the static workspace's device address is unavailable on this device");
  const int L = s->L, T = s->nf, kmax = s->nmodes + 2;
  Work w;
  if (const size_t need = carve_work(s, base, kScratchBytes, B, L, T, kmax, w))
    return fail(2, "the batch needs " + std::to_string(need) + " bytes of static workspace; the limit is " +
                       std::to_string(kScratchBytes) + " (MD_ROCM_STATIC_SCRATCH_BYTES): split the batch");
  const mdr::PrepParams pp{wave, s->cfg.tau_floor, s->cfg.tau_phase};
  MD_HIP(mdr::launch_prep(thickness, vp, vs, rho, bounds, periods, row_periods, B, L, T, s->nmodes, pp, w.rows,
                          w.models, w.traces, s->stream));
  s->prep_models = B;
  s->prep_rows = L;
  s->timed = false;
  return run_pipeline(s, w, static_cast<long long>(B) * T, kmax, L, wave, c_tol_rel, velocity, nullptr, flags,
                      levels, counts, s->cfg.group_velocity ? group : nullptr, group_flags);
}

int md_rocm_autotune(md_rocm_solver *s, const double *thickness, const double *vp, const double *vs,
                     const double *rho, int B, const double *periods, int row_periods, const double *bounds,
                     double c_tol_rel, int wave, double *velocity, uint32_t *flags, int32_t *levels,
                     int32_t *counts, double *group, uint32_t *group_flags, int *scan_waves,
                     int *consumer_blocks) {
  if (!s) return fail(2, "null solver");
  return autotune_tiles(s, [&] {
    return md_rocm_run_device(s, thickness, vp, vs, rho, B, periods, row_periods, bounds, c_tol_rel, wave, velocity,
                              flags, levels, counts, group, group_flags);
  }, scan_waves, consumer_blocks);
}

int md_rocm_run(md_rocm_solver *s, const double *thickness, const double *vp, const double *vs, const double *rho,
                int B, const double *periods, int row_periods, const double *bounds, double c_tol_rel,
                double f_accept, int wave, int automatic, void *velocity, int velocity_is_f32, uint32_t *flags,
                int32_t *levels, int32_t *counts, double *group, uint32_t *group_flags, uint32_t *prediction_flags,
                uint64_t *stats, int nstats) {
  (void)automatic;
  if (!s) return fail(2, "null solver");
  if (B < 1 || B > s->capacity) return fail(2, "B must be 1..capacity");
  if (!thickness || !vp || !vs || !rho || !periods || !bounds || !velocity || !flags || !levels || !counts)
    return fail(2, "null input or output pointer");
  if (!(f_accept >= 0.) || !std::isfinite(f_accept)) return fail(2, "f_accept must be non-negative");
  // Reported roots with |F| > f_accept (the max-normalised secular function,
  // as the reference's) become NaN with TRACE_F_REJECTED; 1 disables the test.
  AcceptScope accept_scope(s, f_accept);
  const int L = s->L, T = s->nf, M = s->nmodes;
  const size_t nrow = static_cast<size_t>(B) * L, ncell = static_cast<size_t>(B) * M * T,
               ntr = static_cast<size_t>(B) * T;
  const size_t nper = row_periods ? ntr : static_cast<size_t>(T);
  for (DevBuf *b : {&s->in_h, &s->in_vp, &s->in_vs, &s->in_rho})
    if (int e = b->ensure(nrow * sizeof(double))) return e;
  if (int e = s->in_per.ensure(nper * sizeof(double))) return e;
  if (int e = s->in_bounds.ensure(static_cast<size_t>(B) * 2 * sizeof(double))) return e;
  if (int e = s->out_v.ensure(ncell * sizeof(double))) return e;
  if (int e = s->out_flags.ensure(ntr * sizeof(uint32_t))) return e;
  if (int e = s->out_levels.ensure(ntr * sizeof(int32_t))) return e;
  if (int e = s->out_counts.ensure(ntr * sizeof(int32_t))) return e;
  const bool want_group = s->cfg.group_velocity != 0;
  if (want_group) {
    if (!group || !group_flags) return fail(2, "group outputs required");
    if (int e = s->out_group.ensure(ncell * sizeof(double))) return e;
    if (int e = s->out_gflags.ensure(ncell * sizeof(uint32_t))) return e;
  }
  hipStream_t st = s->stream;
  MD_HIP(hipMemcpyAsync(s->in_h.p, thickness, nrow * sizeof(double), hipMemcpyHostToDevice, st));
  MD_HIP(hipMemcpyAsync(s->in_vp.p, vp, nrow * sizeof(double), hipMemcpyHostToDevice, st));
  MD_HIP(hipMemcpyAsync(s->in_vs.p, vs, nrow * sizeof(double), hipMemcpyHostToDevice, st));
  MD_HIP(hipMemcpyAsync(s->in_rho.p, rho, nrow * sizeof(double), hipMemcpyHostToDevice, st));
  MD_HIP(hipMemcpyAsync(s->in_per.p, periods, nper * sizeof(double), hipMemcpyHostToDevice, st));
  MD_HIP(hipMemcpyAsync(s->in_bounds.p, bounds, static_cast<size_t>(B) * 2 * sizeof(double), hipMemcpyHostToDevice,
                        st));
  std::vector<char> bypassed;
  if (s->cfg.continuation) {
    if (int e = s->rows.ensure(nrow * sizeof(mdr::PrepRow))) return e;
    if (int e = s->models.ensure(static_cast<size_t>(B) * sizeof(mdr::ModelInfo))) return e;
    if (int e = s->traces.ensure(ntr * sizeof(mdr::TraceInfo))) return e;
    if (int e = s->c_pred.ensure(ntr * sizeof(uint32_t))) return e;
    s->timed = true;
    MD_HIP(hipEventRecord(s->ev_begin, st));
    const mdr::PrepParams pp{wave, s->cfg.tau_floor, s->cfg.tau_phase};
    MD_HIP(mdr::launch_prep(s->in_h.as<double>(), s->in_vp.as<double>(), s->in_vs.as<double>(), s->in_rho.as<double>(),
                            s->in_bounds.as<double>(), s->in_per.as<double>(), row_periods, B, L, T, M, pp,
                            s->rows.as<mdr::PrepRow>(), s->models.as<mdr::ModelInfo>(),
                            s->traces.as<mdr::TraceInfo>(), st));
    s->prep_models = B;
    s->prep_rows = L;
    std::vector<int> anchors;
    std::vector<mdr::TargetPlan> targets;
    bypassed.assign(B, 0);
    for (int m = 0; m < B; ++m) {
      bool by = false;
      plan_model(row_periods ? periods + static_cast<size_t>(m) * T : periods, T, static_cast<long long>(m) * T,
                 s->cfg.anchor_stride, anchors, targets, by);
      bypassed[m] = by;
    }
    if (int e = run_continuation(s, M + 2, L, wave, c_tol_rel, s->out_v.as<double>(), s->out_flags.as<uint32_t>(),
                                 s->out_levels.as<int32_t>(), s->out_counts.as<int32_t>(),
                                 want_group ? s->out_group.as<double>() : nullptr,
                                 want_group ? s->out_gflags.as<uint32_t>() : nullptr, s->c_pred.as<uint32_t>(),
                                 anchors, targets))
      return e;
    MD_HIP(hipEventRecord(s->ev_end, st));
  } else if (int e = md_rocm_run_device(s, s->in_h.as<double>(), s->in_vp.as<double>(), s->in_vs.as<double>(),
                                        s->in_rho.as<double>(), B, s->in_per.as<double>(), row_periods,
                                        s->in_bounds.as<double>(), c_tol_rel, wave, s->out_v.as<double>(),
                                        s->out_flags.as<uint32_t>(), s->out_levels.as<int32_t>(),
                                        s->out_counts.as<int32_t>(), want_group ? s->out_group.as<double>() : nullptr,
                                        want_group ? s->out_gflags.as<uint32_t>() : nullptr)) {
    return e;
  }
  std::vector<HostCopy> parts;
  if (velocity_is_f32) {  // narrowed on the device: half the bytes cross to the host
    if (int e = s->out_v32.ensure(ncell * sizeof(float))) return e;
    MD_HIP(mdr::launch_narrow(s->out_v.as<double>(), s->out_v32.as<float>(), ncell, st));
    parts.push_back({velocity, s->out_v32.p, ncell * sizeof(float)});
  } else {
    parts.push_back({velocity, s->out_v.p, ncell * sizeof(double)});
  }
  parts.push_back({flags, s->out_flags.p, ntr * sizeof(uint32_t)});
  parts.push_back({levels, s->out_levels.p, ntr * sizeof(int32_t)});
  parts.push_back({counts, s->out_counts.p, ntr * sizeof(int32_t)});
  if (want_group) {
    parts.push_back({group, s->out_group.p, ncell * sizeof(double)});
    parts.push_back({group_flags, s->out_gflags.p, ncell * sizeof(uint32_t)});
  }
  if (int e = copy_to_host(st, s->host_stage, parts)) return e;
  if (int e = md_rocm_synchronize(s)) return e;
  if (prediction_flags) {
    if (s->cfg.continuation) {
      MD_HIP(hipMemcpy(prediction_flags, s->c_pred.p, ntr * sizeof(uint32_t), hipMemcpyDeviceToHost));
      for (int m = 0; m < B; ++m)
        if (bypassed[m]) std::fill_n(prediction_flags + static_cast<size_t>(m) * T, T, mdr::kPredBypassed);
    } else {
      std::fill_n(prediction_flags, ntr, mdr::kPredBypassed);  // independent periods
    }
  }
  if (stats && nstats > 0) {
    unsigned long long st4[kStatCount] = {};
    MD_HIP(hipMemcpy(st4, s->stats.p, sizeof(st4), hipMemcpyDeviceToHost));
    for (int i = 0; i < std::min(nstats, kStatCount); ++i) stats[i] = st4[i];
  }
  return 0;
}

int md_rocm_run_ragged(md_rocm_solver *s, const double *thickness, const double *vp, const double *vs,
                       const double *rho, const int64_t *row_offsets, int B, const double *periods,
                       const int64_t *period_offsets, const int32_t *modes, const double *bounds,
                       double c_tol_rel, int wave, double *velocity, uint32_t *flags, int32_t *levels,
                       int32_t *counts, double *group, uint32_t *group_flags, uint32_t *prediction_flags) {
  if (!s) return fail(2, "null solver");
  if (B < 1) return fail(2, "B must be positive");
  if (!thickness || !vp || !vs || !rho || !row_offsets || !periods || !period_offsets || !modes || !bounds ||
      !velocity || !flags || !levels || !counts)
    return fail(2, "null input or output pointer");
  const bool want_group = s->cfg.group_velocity != 0;
  if (want_group && (!group || !group_flags)) return fail(2, "group outputs required");
  if (row_offsets[0] != 0 || period_offsets[0] != 0) return fail(2, "offsets must start at 0");
  std::vector<int64_t> cells(static_cast<size_t>(B) + 1, 0);
  int kmax = 0, max_rows = 1, max_modes = 1;
  for (int m = 0; m < B; ++m) {
    const long long L = row_offsets[m + 1] - row_offsets[m], T = period_offsets[m + 1] - period_offsets[m];
    if (L < 1 || L > mdr::kMaxRows || T < 0 || modes[m] < 1 || modes[m] > 256)
      return fail(2, "invalid ragged geometry at model " + std::to_string(m));
    cells[m + 1] = cells[m] + static_cast<long long>(modes[m]) * T;
    kmax = std::max(kmax, modes[m] + 2);
    max_modes = std::max(max_modes, static_cast<int>(modes[m]));
    max_rows = std::max(max_rows, static_cast<int>(L));
  }
  const long long nrows = row_offsets[B], ntr = period_offsets[B];
  if (nrows > INT32_MAX || ntr > INT32_MAX) return fail(2, "rows and traces must each be below 2^31");
  const size_t ncell = static_cast<size_t>(cells[B]);
  // The packed arrays go to the device as they are; the models and traces are
  // prepared there by the kernels md_rocm_run_ragged_device uses.
  const size_t nb = static_cast<size_t>(B), r8 = static_cast<size_t>(nrows) * sizeof(double);
  for (DevBuf *d : {&s->in_h, &s->in_vp, &s->in_vs, &s->in_rho})
    if (int e = d->ensure(r8)) return e;
  if (int e = s->in_per.ensure(std::max<size_t>(1, static_cast<size_t>(ntr)) * sizeof(double))) return e;
  if (int e = s->in_bounds.ensure(2 * nb * sizeof(double))) return e;
  if (int e = s->in_offsets.ensure(3 * (nb + 1) * sizeof(int64_t) + nb * sizeof(int32_t))) return e;
  if (int e = s->rows.ensure(static_cast<size_t>(nrows) * sizeof(mdr::PrepRow))) return e;
  if (int e = s->models.ensure((nb + 1) * sizeof(mdr::ModelInfo))) return e;
  if (int e = s->traces.ensure(std::max<size_t>(1, static_cast<size_t>(ntr)) * sizeof(mdr::TraceInfo))) return e;
  if (int e = s->invalid.ensure(sizeof(int))) return e;
  if (int e = s->out_v.ensure(ncell * sizeof(double))) return e;
  if (int e = s->out_flags.ensure(static_cast<size_t>(ntr) * sizeof(uint32_t))) return e;
  if (int e = s->out_levels.ensure(static_cast<size_t>(ntr) * sizeof(int32_t))) return e;
  if (int e = s->out_counts.ensure(static_cast<size_t>(ntr) * sizeof(int32_t))) return e;
  if (want_group) {
    if (int e = s->out_group.ensure(ncell * sizeof(double))) return e;
    if (int e = s->out_gflags.ensure(ncell * sizeof(uint32_t))) return e;
  }
  hipStream_t st = s->stream;
  s->timed = true;
  MD_HIP(hipEventRecord(s->ev_begin, st));
  const std::pair<DevBuf *, const double *> cols[] = {
      {&s->in_h, thickness}, {&s->in_vp, vp}, {&s->in_vs, vs}, {&s->in_rho, rho}};
  for (const auto &c : cols) MD_HIP(hipMemcpyAsync(c.first->p, c.second, r8, hipMemcpyHostToDevice, st));
  if (ntr > 0)
    MD_HIP(hipMemcpyAsync(s->in_per.p, periods, static_cast<size_t>(ntr) * sizeof(double), hipMemcpyHostToDevice,
                          st));
  MD_HIP(hipMemcpyAsync(s->in_bounds.p, bounds, 2 * nb * sizeof(double), hipMemcpyHostToDevice, st));
  auto *offs = s->in_offsets.as<int64_t>();
  MD_HIP(hipMemcpyAsync(offs, row_offsets, (nb + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
  MD_HIP(hipMemcpyAsync(offs + nb + 1, period_offsets, (nb + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
  MD_HIP(hipMemcpyAsync(offs + 2 * (nb + 1), cells.data(), (nb + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
  auto *dmodes = reinterpret_cast<int32_t *>(offs + 3 * (nb + 1));
  MD_HIP(hipMemcpyAsync(dmodes, modes, nb * sizeof(int32_t), hipMemcpyHostToDevice, st));
  const mdr::RaggedGeometry g{offs, offs + nb + 1, offs + 2 * (nb + 1), dmodes, B, max_rows, max_modes,
                              nrows, ntr, static_cast<long long>(ncell)};
  const mdr::PrepParams pp{wave, s->cfg.tau_floor, s->cfg.tau_phase};
  MD_HIP(mdr::launch_prep_ragged(s->in_h.as<double>(), s->in_vp.as<double>(), s->in_vs.as<double>(),
                                 s->in_rho.as<double>(), s->in_bounds.as<double>(), s->in_per.as<double>(), g, pp,
                                 s->rows.as<mdr::PrepRow>(), s->models.as<mdr::ModelInfo>(),
                                 s->traces.as<mdr::TraceInfo>(), s->invalid.as<int>(), st));
  s->prep_models = B + 1;  // the sentinel model B (no valid model leaves traces to it here)
  s->prep_rows = max_rows;
  std::vector<char> bypassed;
  if (s->cfg.continuation) {
    if (int e = s->c_pred.ensure(static_cast<size_t>(ntr) * sizeof(uint32_t))) return e;
    std::vector<int> anchors;
    std::vector<mdr::TargetPlan> targets;
    bypassed.assign(B, 0);
    for (int m = 0; m < B; ++m) {
      bool by = false;
      const long long p0 = period_offsets[m];
      plan_model(periods + p0, static_cast<int>(period_offsets[m + 1] - p0), p0, s->cfg.anchor_stride, anchors,
                 targets, by);
      bypassed[m] = by;
    }
    if (int e = run_continuation(s, kmax, max_rows, wave, c_tol_rel, s->out_v.as<double>(),
                                 s->out_flags.as<uint32_t>(), s->out_levels.as<int32_t>(), s->out_counts.as<int32_t>(),
                                 want_group ? s->out_group.as<double>() : nullptr,
                                 want_group ? s->out_gflags.as<uint32_t>() : nullptr, s->c_pred.as<uint32_t>(),
                                 anchors, targets))
      return e;
  } else if (int e = solver_pipeline(s, ntr, kmax, max_rows, wave, c_tol_rel, s->out_v.as<double>(),
                                     s->out_flags.as<uint32_t>(), s->out_levels.as<int32_t>(),
                                  s->out_counts.as<int32_t>(), want_group ? s->out_group.as<double>() : nullptr,
                                  want_group ? s->out_gflags.as<uint32_t>() : nullptr)) {
    return e;
  }
  MD_HIP(hipEventRecord(s->ev_end, st));
  std::vector<HostCopy> parts = {{velocity, s->out_v.p, ncell * sizeof(double)},
                                 {flags, s->out_flags.p, static_cast<size_t>(ntr) * sizeof(uint32_t)},
                                 {levels, s->out_levels.p, static_cast<size_t>(ntr) * sizeof(int32_t)},
                                 {counts, s->out_counts.p, static_cast<size_t>(ntr) * sizeof(int32_t)}};
  if (want_group) {
    parts.push_back({group, s->out_group.p, ncell * sizeof(double)});
    parts.push_back({group_flags, s->out_gflags.p, ncell * sizeof(uint32_t)});
  }
  if (prediction_flags && s->cfg.continuation)
    parts.push_back({prediction_flags, s->c_pred.p, static_cast<size_t>(ntr) * sizeof(uint32_t)});
  if (int e = copy_to_host(st, s->host_stage, parts)) return e;
  if (int e = md_rocm_synchronize(s)) return e;
  if (prediction_flags) {
    for (int m = 0; m < B; ++m)
      if (!s->cfg.continuation || bypassed[m])
        std::fill(prediction_flags + period_offsets[m], prediction_flags + period_offsets[m + 1], mdr::kPredBypassed);
  }
  return 0;
}

}  // extern "C"

int md_rocm_run_ragged_device(md_rocm_solver *s, const double *thickness, const double *vp, const double *vs,
                              const double *rho, const int64_t *row_offsets, long long nrows, int B,
                              const double *periods, const int64_t *period_offsets, long long ntraces,
                              const int32_t *modes, const int64_t *cell_offsets, long long ncells,
                              const double *bounds, int max_rows, int max_modes, double c_tol_rel, int wave,
                              double *velocity, uint32_t *flags, int32_t *levels, int32_t *counts, double *group,
                              uint32_t *group_flags) {
  if (!s) return fail(2, "null solver");
  if (B < 1) return fail(2, "B must be positive");
  if (nrows < 1 || nrows > INT32_MAX) return fail(2, "nrows must be 1..2^31-1");
  if (ntraces < 0 || ntraces > INT32_MAX) return fail(2, "ntraces must be 0..2^31-1");
  if (ncells < 0) return fail(2, "ncells must be non-negative");
  if (max_rows < 1 || max_rows > mdr::kMaxRows)
    return fail(2, "max_rows must be 1.." + std::to_string(mdr::kMaxRows));
  if (max_modes < 1 || max_modes > 256) return fail(2, "max_modes must be 1..256");
  if (wave != 0 && wave != 1) return fail(2, "wave must be 0 (Rayleigh) or 1 (Love)");
  if (!(c_tol_rel > 0.) || !std::isfinite(c_tol_rel)) return fail(2, "c_tol_rel must be positive");
  if (!thickness || !vp || !vs || !rho || !row_offsets || !periods || !period_offsets || !modes || !cell_offsets ||
      !bounds || !velocity || !flags || !levels || !counts)
    return fail(2, "null input or output pointer");
  const bool want_group = s->cfg.group_velocity != 0;
  if (want_group && (!group || !group_flags)) return fail(2, "group outputs required");
  if (s->cfg.continuation)
    return fail(2, "anchor continuation plans its anchors from host periods: use md_rocm_run_ragged");
  if (int e = s->rows.ensure(static_cast<size_t>(nrows) * sizeof(mdr::PrepRow))) return e;
  if (int e = s->models.ensure(static_cast<size_t>(B + 1) * sizeof(mdr::ModelInfo))) return e;
  if (int e = s->traces.ensure(static_cast<size_t>(std::max(ntraces, 1LL)) * sizeof(mdr::TraceInfo))) return e;
  if (int e = s->invalid.ensure(sizeof(int))) return e;
  const mdr::RaggedGeometry g{row_offsets, period_offsets, cell_offsets, modes, B, max_rows, max_modes,
                              nrows, ntraces, ncells};
  const int kmax = max_modes + 2;
  const std::vector<uintptr_t> key = {
      1, addr(thickness), addr(vp), addr(vs), addr(rho), addr(row_offsets), static_cast<uintptr_t>(nrows),
      static_cast<uintptr_t>(B), addr(periods), addr(period_offsets), static_cast<uintptr_t>(ntraces), addr(modes),
      addr(cell_offsets), static_cast<uintptr_t>(ncells), addr(bounds), static_cast<uintptr_t>(max_rows),
      static_cast<uintptr_t>(max_modes), double_bits(c_tol_rel), static_cast<uintptr_t>(wave), addr(velocity),
      addr(flags), addr(levels), addr(counts), addr(group), addr(group_flags),
      static_cast<uintptr_t>(s->cfg.scan_waves), static_cast<uintptr_t>(s->cfg.consumer_blocks), accept_bits(s)};
  const int e = graph_run(s, key, [&]() -> int {
    const mdr::PrepParams pp{wave, s->cfg.tau_floor, s->cfg.tau_phase};
    MD_HIP(mdr::launch_prep_ragged(thickness, vp, vs, rho, bounds, periods, g, pp, s->rows.as<mdr::PrepRow>(),
                                   s->models.as<mdr::ModelInfo>(), s->traces.as<mdr::TraceInfo>(),
                                   s->invalid.as<int>(), s->stream));
    s->prep_models = B + 1;  // the sentinel model B holds traces no valid model claims
    s->prep_rows = max_rows;
    Work w;
    if (int e = solver_work(s, ntraces, kmax, w)) return e;
    return run_pipeline(s, w, ntraces, kmax, max_rows, wave, c_tol_rel, velocity, nullptr, flags, levels, counts,
                        want_group ? group : nullptr, group_flags);
  });
  s->invalid_live = e == 0;
  return e;
}

int md_rocm_autotune_ragged_device(md_rocm_solver *s, const double *thickness, const double *vp, const double *vs,
                                   const double *rho, const int64_t *row_offsets, long long nrows, int B,
                                   const double *periods, const int64_t *period_offsets, long long ntraces,
                                   const int32_t *modes, const int64_t *cell_offsets, long long ncells,
                                   const double *bounds, int max_rows, int max_modes, double c_tol_rel, int wave,
                                   double *velocity, uint32_t *flags, int32_t *levels, int32_t *counts,
                                   double *group, uint32_t *group_flags, int *scan_waves, int *consumer_blocks) {
  if (!s) return fail(2, "null solver");
  return autotune_tiles(s, [&] {
    return md_rocm_run_ragged_device(s, thickness, vp, vs, rho, row_offsets, nrows, B, periods, period_offsets,
                                     ntraces, modes, cell_offsets, ncells, bounds, max_rows, max_modes, c_tol_rel,
                                     wave, velocity, flags, levels, counts, group, group_flags);
  }, scan_waves, consumer_blocks);
}

extern "C" int md_rocm_probe_grid(const double *thickness, const double *vp, const double *vs, const double *rho,
                                  int B, int L, const double *periods, int nf, const double *bounds, int nc, int wave,
                                  int fp32, double *out) {
  if (!thickness || !vp || !vs || !rho || !periods || !bounds || !out) return fail(2, "null pointer");
  if (B < 1 || L < 1 || L > mdr::kMaxRows || nf < 1 || nc < 2 || (wave != 0 && wave != 1))
    return fail(2, "invalid probe geometry");
  const size_t nrow = static_cast<size_t>(B) * L, ntr = static_cast<size_t>(B) * nf;
  DevBuf h, a, b, r, per, bnd, rows, models, traces, res;
  for (DevBuf *x : {&h, &a, &b, &r})
    if (int e = x->ensure(nrow * sizeof(double))) return e;
  if (int e = per.ensure(nf * sizeof(double))) return e;
  if (int e = bnd.ensure(static_cast<size_t>(B) * 2 * sizeof(double))) return e;
  if (int e = rows.ensure(nrow * sizeof(mdr::PrepRow))) return e;
  if (int e = models.ensure(static_cast<size_t>(B) * sizeof(mdr::ModelInfo))) return e;
  if (int e = traces.ensure(ntr * sizeof(mdr::TraceInfo))) return e;
  if (int e = res.ensure(ntr * nc * sizeof(double))) return e;
  auto cleanup = [&] {
    for (DevBuf *x : {&h, &a, &b, &r, &per, &bnd, &rows, &models, &traces, &res}) x->release();
  };
  hipStream_t st = nullptr;
  int code = 0;
  do {
    if (hipMemcpy(h.p, thickness, nrow * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(a.p, vp, nrow * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b.p, vs, nrow * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(r.p, rho, nrow * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(per.p, periods, nf * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(bnd.p, bounds, static_cast<size_t>(B) * 16, hipMemcpyHostToDevice) != hipSuccess) {
      code = fail(1, "probe upload failed");
      break;
    }
    const mdr::PrepParams pp{wave, 0., 0.};
    if (mdr::launch_prep(h.as<double>(), a.as<double>(), b.as<double>(), r.as<double>(), bnd.as<double>(),
                         per.as<double>(), 0, B, L, nf, 1, pp, rows.as<mdr::PrepRow>(), models.as<mdr::ModelInfo>(),
                         traces.as<mdr::TraceInfo>(), st) != hipSuccess ||
        mdr::launch_probe(traces.as<mdr::TraceInfo>(), models.as<mdr::ModelInfo>(), rows.as<mdr::PrepRow>(),
                          static_cast<int>(ntr), nc, wave, fp32, res.as<double>(), st) != hipSuccess) {
      code = fail(1, "probe launch failed");
      break;
    }
    if (hipMemcpy(out, res.p, ntr * nc * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      code = fail(1, "probe download failed");
  } while (false);
  cleanup();
  return code;
}

// Continuity screen of completed phases in caller grid order (semantics of
// the reference's cpp/cpu/continuity_screen.hpp:8-67, used by
// python/modisp/__init__.py:446-455): per (model, period) a byte of
// ContinuityFlag bits -- 8 NO_BILATERAL (no mode has finite values at both
// neighbouring periods), 16 ENVELOPE (a mode leaves [0.95 min, 1.05 max] of its
// neighbours, or vanishes between them), 32 CLOSE_FP32 (FP32 phases of
// adjacent modes within 16 ulp), 64 PREDICTION (continuation asked for
// review). Neighbours are the adjacent periods in ascending order, used only
// when their log step is at most 0.25.
//
// The neighbour table (left, right, usable) is built once from the period
// order; each model is then screened row by row over its contiguous [K, T]
// block, accumulating per-period masks: "some mode has both neighbours"
// (clears NO_BILATERAL), "some mode leaves its envelope", "some adjacent
// modes alias in FP32".
struct Neighbours {
  std::vector<int> left, right;
  std::vector<uint8_t> usable;
  Neighbours(const double *p, int T, bool bilateral) : left(T), right(T), usable(T, 0) {
    std::vector<int> rank(T);
    for (int i = 0; i < T; ++i) rank[i] = i;
    if (bilateral && T > 2 && !std::is_sorted(p, p + T))
      std::stable_sort(rank.begin(), rank.end(), [&](int i, int j) { return p[i] < p[j]; });
    for (int j = 0; j < T; ++j) {
      const int f = rank[j];
      left[f] = rank[j > 0 ? j - 1 : j];
      right[f] = rank[j + 1 < T ? j + 1 : j];
      usable[f] = bilateral && j > 0 && j + 1 < T && cont_log_step(p[right[f]], p[left[f]]) <= .25;
    }
  }
};

template <class T>
static bool leaves_envelope(T x, T a, T b) {
  return !std::isfinite(x) || x < T(.95) * std::min(a, b) || x > T(1.05) * std::max(a, b);
}

template <class T>
static bool fp32_alias(T x, T y) {
  if constexpr (sizeof(T) != 4) {
    (void)x;
    (void)y;
    return false;
  } else {
    if (!std::isfinite(x) || !std::isfinite(y)) return false;
    uint32_t a, c;
    std::memcpy(&a, &x, 4);
    std::memcpy(&c, &y, 4);
    return (a > c ? a - c : c - a) <= 16u;
  }
}

template <class T>
static int screen(const double *p, int Tn, const T *v, int B, int K, const uint32_t *pred, int bilateral,
                  uint8_t *out, int *suspect) {
  const Neighbours nb(p, Tn, bilateral != 0);
  const bool aliases = bilateral && sizeof(T) == 4 && K > 1;
  std::vector<uint8_t> paired(Tn), jump(Tn), alias(Tn);
  int any = 0;
  for (int b = 0; b < B; ++b) {
    std::fill(paired.begin(), paired.end(), 0);
    std::fill(jump.begin(), jump.end(), 0);
    std::fill(alias.begin(), alias.end(), 0);
    const T *block = v + static_cast<size_t>(b) * K * Tn;
    for (int k = 0; k < K; ++k) {
      const T *row = block + static_cast<size_t>(k) * Tn;
      for (int f = 0; f < Tn; ++f) {
        if (!nb.usable[f]) continue;
        const T a = row[nb.left[f]], c = row[nb.right[f]];
        if (!std::isfinite(a) || !std::isfinite(c)) continue;
        paired[f] = 1;
        jump[f] |= leaves_envelope(row[f], a, c);
      }
      if (aliases && k > 0)
        for (int f = 0; f < Tn; ++f) alias[f] |= fp32_alias(row[f], row[f - Tn]);
    }
    for (int f = 0; f < Tn; ++f) {
      uint8_t flag = 0;
      if (bilateral && !paired[f]) flag |= 8;
      if (jump[f]) flag |= 16;
      if (alias[f]) flag |= 32;
      if (pred && (pred[static_cast<size_t>(b) * Tn + f] & mdr::kPredReview)) flag |= 64;
      out[static_cast<size_t>(b) * Tn + f] = flag;
      any |= (flag & (16 | 32 | 64)) != 0;
    }
  }
  *suspect = any;
  return 0;
}

extern "C" int md_rocm_continuity_screen(const double *periods, int T, const void *phase, int phase_is_f32, int B,
                                         int K, const uint32_t *prediction, int bilateral, uint8_t *flags,
                                         int *suspect) {
  if (!periods || !phase || !flags || !suspect || T < 1 || B < 0 || K < 1)
    return fail(2, "continuity screening requires periods[T] and phases[B,K,T]");
  return phase_is_f32 ? screen(periods, T, static_cast<const float *>(phase), B, K, prediction, bilateral, flags, suspect)
                      : screen(periods, T, static_cast<const double *>(phase), B, K, prediction, bilateral, flags,
                               suspect);
}
