/* modisp-rocm: C ABI of the MI355X (gfx950) multimode dispersion backend.
 *
 * This is the drop-in boundary of the `modisp-rocm` plugin. Each entry point
 * replaces one binding of the reference's CUDA plugin (YuanYusung/modisp
 * 0.5.0), cited per declaration. All pointers are plain host or device
 * pointers; no framework types cross this boundary.
 *
 * Conventions (as the reference's legacy C ABI, cpp/cuda/coarse_scan.cu:31-47):
 * functions return 0 on success and a non-zero code on failure, with the
 * message available from md_rocm_last_error() (thread-local here):
 *   1 = device/runtime error (reference: RuntimeError "QE CUDA: ...", qe.cu:823-825)
 *   2 = invalid argument      (reference: ValueError, qe.cu:978-995)
 *   3 = allocation failure    (reference: MemoryError, review_groups.cu:15-17)
 * Numerical failures never fail a call; they become flags, NaN and counts.
 */
#ifndef MODISP_ROCM_H
#define MODISP_ROCM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MD_ROCM_ABI_VERSION 9 /* reference python/modisp/_loader.py:4 */
#define MD_ROCM_VERSION "0.5.0"

typedef struct md_rocm_solver md_rocm_solver;

/* Solver configuration; the fields mirror the reference QeSolver constructor
 * (cpp/cuda/qe.cu:1166): accuracy, close_root_rtol, single_precision,
 * group_velocity, group_rtol, continuation, anchor_stride; plus this
 * backend's grid size and FP32 uncertainty band. */
typedef struct md_rocm_config {
  int accuracy;            /* reference preset code: 0 high, 1 balanced, 2 fast, 4 auto (3, the legacy
                              uniform path, is grid = 0 here). Every preset runs this backend's exhaustive
                              grid scan with hidden-pair probes; the preset sets the reported contract:
                              search_level (with `shape`) and TRACE_APPROXIMATE_SEARCH (2048) on every
                              searched trace of balanced, fast and auto, as the reference reports them */
  double close_root_rtol;  /* close-root threshold, default 0.005 */
  int single_precision;    /* 1: FP32 grid scan, FP64 recheck/refine; 0: FP64 grid scan */
  int group_velocity;      /* 1: analytic FP64 group velocity */
  double group_rtol;       /* group velocities whose rounding-error bound exceeds 0.25 group_rtol
                              max(c, |U|) are flagged DERIVATIVE_UNSTABLE (group_velocity.hpp:87-92) */
  int continuation;        /* 1: anchor continuation across periods on the host-array entry points
                              (md_rocm_run, md_rocm_run_ragged): anchors every <= anchor_stride periods
                              (log step <= 0.25) by the full pipeline, the rest predicted, refined and
                              audited, failures re-solved by the full pipeline; models with fewer than
                              8 strictly increasing periods are solved independently */
  int anchor_stride;       /* anchors every <= stride periods, default 4 */
  int nc;                  /* uniform grid: samples per trace, 3..65536 (default 512) */
  double tau_floor;        /* FP32 |F| uncertainty band floor (default 1e-4) */
  double tau_phase;        /* band growth per radian of vertical S phase (default 5e-5) */
  int chunk_traces;        /* traces per device chunk (0 = automatic) */
  int grid;                /* 1: density-adaptive nodes (default), 0: uniform nc-node grid */
  double rho;              /* adaptive grid: nodes per WKB mode (default 8) */
  int graph;               /* 1: md_rocm_run_device and md_rocm_run_ragged_device capture their
                              launch sequence in a hipGraph and replay it for repeated calls
                              with identical arguments */
  int scan_waves;          /* traces (wavefronts) per scan block: 1, 2, 4 or 8 (FP64: <= 4);
                              0 = 4, or what md_rocm_autotune picked */
  int consumer_blocks;     /* persistent consumer blocks per compute unit, 1..64 (0 = 16) */
  int refinement;          /* bracket refinement (reference QeSolver refinement, refinement.hpp:11-65):
                              0 legacy regula falsi, 1 Brent-Dekker (default), 2 Chandrupatla */
  int shape;               /* reference `shape` (search_strategy shape_adaptive, the default 1, or
                              qe_adaptive, 0): selects the search_level codes (8/9/10 against 3/6/7,
                              docs/API.md:203-209) and, as the reference's searches do, whether close
                              roots of a Rayleigh trace raise TRACE_CLOSE_ROOT_RISK (4096): the shape and
                              observation searches test them (cpp/cpu/solver.cpp:196,145), the high
                              adaptive ladder does not */
  int love_phase;          /* 1 (default): Love roots verified by the exact phase count and re-bracketed
                              by phase bisection where the grid disagrees (love_phase.hpp:76-91), level
                              11; 0: the grid scan alone (reference love_strategy="scan") */
  int secular;             /* 1 (fast, the contracted Dunkin form, dunkin.hpp:459-566): the only form
                              built here; 0 (the released 25-entry form) is refused */
  double sampling_rtol;    /* validated as the reference (0 < x <= 1, or -1 for auto sampling) and kept
                              for reports: it loosens the reference's shape sampling, which this
                              backend's fixed-density scan does not use */
  double period_threshold; /* validated (finite, >= 0) and kept for reports; the reference also keeps it
                              as metadata only (cpp/common/accuracy.hpp:47-53) */
} md_rocm_config;

/* Defaults for every field. */
void md_rocm_default_config(md_rocm_config *cfg);

/* Create a solver for up to `capacity` models of `L` rows, `nf` periods and
 * `nmodes` reported modes. Replaces QeSolver.__init__ (cpp/cuda/qe.cu:966-1038).
 * Returns NULL on failure (see md_rocm_last_error). */
md_rocm_solver *md_rocm_create(int capacity, int L, int nf, int nmodes, const md_rocm_config *cfg);

/* Release every resource; NULL is accepted. Replaces QeSolver.close (qe.cu:1039-1059). */
void md_rocm_free(md_rocm_solver *s);

/* Solve B <= capacity models from HOST arrays. Replaces QeSolver.run
 * (cpp/cuda/qe.cu:1060-1118). f_accept < 1 rejects reported roots whose
 * max-normalised |F| exceeds it: NaN velocity, TRACE_F_REJECTED (128), group
 * flag MISSING_ROOT (reference cpp/cpu/solver.cpp:333); a continued target
 * that fails it falls back to the full search (continuation.hpp:201-204).
 *   thickness, vp, vs, rho : [B, L] row-major float64
 *   periods                : [nf] shared, or [B, nf] when row_periods != 0
 *   bounds                 : [B, 2] search window (lo, hi)
 *   wave                   : 0 Rayleigh, 1 Love
 * Outputs (host, caller-allocated):
 *   velocity [B, nmodes, nf] float64 (or float32 when velocity_is_f32),
 *   flags [B, nf] uint32, levels [B, nf] int32, counts [B, nf] int32,
 *   group / group_flags [B, nmodes, nf] (may be NULL when group velocity is off),
 *   prediction_flags [B, nf] (may be NULL), stats[nstats] uint64 (may be NULL). */
int md_rocm_run(md_rocm_solver *s, const double *thickness, const double *vp, const double *vs,
                const double *rho, int B, const double *periods, int row_periods, const double *bounds,
                double c_tol_rel, double f_accept, int wave, int automatic, void *velocity,
                int velocity_is_f32, uint32_t *flags, int32_t *levels, int32_t *counts, double *group,
                uint32_t *group_flags, uint32_t *prediction_flags, uint64_t *stats, int nstats);

/* As md_rocm_run, but every input and output pointer is DEVICE memory on the
 * solver's device and the call is asynchronous on the solver's stream until
 * md_rocm_synchronize. No counterpart in the reference (whose run() copies
 * through pinned staging, qe.cu:1091); it exists so device-resident callers
 * (the torch.autograd binding, benchmarks) avoid host round trips. */
int md_rocm_run_device(md_rocm_solver *s, const double *thickness, const double *vp, const double *vs,
                       const double *rho, int B, const double *periods, int row_periods, const double *bounds,
                       double c_tol_rel, int wave, double *velocity, uint32_t *flags, int32_t *levels,
                       int32_t *counts, double *group, uint32_t *group_flags);

/* Bytes of the fixed device workspace of md_rocm_run_device_static, per GPU. */
#define MD_ROCM_STATIC_SCRATCH_BYTES (128u << 20)

/* As md_rocm_run_device (same arguments, results bit for bit the same), but
 * it allocates nothing: no hipMalloc, hipFree or allocating library call.
 * Every work buffer (prepared rows, models, traces, scan tables, events,
 * brackets, queues, counters) lives in one fixed device array of
 * MD_ROCM_STATIC_SCRATCH_BYTES per GPU, set up by md_rocm_create and held
 * until the process ends; a batch is processed in as many trace chunks as
 * that array needs (about 3.4 KB of chunk buffers per trace; per model its
 * prepared rows, scan table and 24 B per period stay resident, 4.7 KB for 13
 * rows and 100 periods). A batch whose prepared inputs alone do not fit
 * (above about 28,000 such models) is rejected with code 2 before any launch.
 * Every call on one GPU shares that GPU's array and leaves its contents
 * behind (counters are cleared by a kernel at the start of each chunk), so
 * calls that overlap on one GPU, from any solver or thread, race: serialise
 * them on one stream. Asynchronous on the solver's stream; no stage timing.
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation). */
int md_rocm_run_device_static(md_rocm_solver *s, const double *thickness, const double *vp, const double *vs,
                              const double *rho, int B, const double *periods, int row_periods,
                              const double *bounds, double c_tol_rel, int wave, double *velocity, uint32_t *flags,
                              int32_t *levels, int32_t *counts, double *group, uint32_t *group_flags);

/* Adjoint of md_rocm_run_device's roots (all pointers DEVICE memory, [B, L]
 * models, [nf] or [B, nf] periods, [B, nmodes, nf] roots and weights):
 * grad_p[b, j] = sum over modes and periods of weight * dc/dp_j for
 * p in (thickness, vp, vs, rho), by the implicit rule dc/dp = -(dF/dp)/(dF/dc)
 * at each root with the propagator recomputed in FP64 (forward sweep in
 * wavenumber duals, reverse sweep with the transposed layer matrices).
 * Roots that are NaN or have zero weight contribute nothing; the water
 * layer's vs entry gets no gradient. No reference counterpart (the reference
 * has no gradients); this is the entry point its torch.autograd binding calls. */
int md_rocm_adjoint(md_rocm_solver *s, const double *thickness, const double *vp, const double *vs,
                    const double *rho, int B, const double *periods, int row_periods, int wave,
                    const double *roots, const double *weight, double *grad_thickness, double *grad_vp,
                    double *grad_vs, double *grad_rho);

/* Tile autotuner: runs md_rocm_run_device's pipeline on the given DEVICE
 * batch (arguments as md_rocm_run_device; outputs are overwritten) for each
 * candidate (traces per scan block x persistent consumer blocks per CU),
 * keeps the fastest in the solver's configuration and returns it through
 * scan_waves / consumer_blocks (either may be NULL). No reference
 * counterpart (the reference fixes its launch shapes, cpp/cuda/qe.cu:962-964). */
int md_rocm_autotune(md_rocm_solver *s, const double *thickness, const double *vp, const double *vs,
                     const double *rho, int B, const double *periods, int row_periods, const double *bounds,
                     double c_tol_rel, int wave, double *velocity, uint32_t *flags, int32_t *levels,
                     int32_t *counts, double *group, uint32_t *group_flags, int *scan_waves,
                     int *consumer_blocks);

/* Run this solver's work on `stream` (a hipStream_t, e.g. the caller's
 * framework stream; NULL is the default stream) instead of its own, or, with
 * own != 0, return to the solver's own stream. The solver does not own or
 * destroy a stream set here. No reference counterpart (the reference keeps
 * one private stream per solver, cpp/cuda/qe.cu:998). */
int md_rocm_set_stream(md_rocm_solver *s, void *stream, int own);

/* Ragged batch from HOST arrays: model m has rows [row_offsets[m], row_offsets[m+1])
 * of the packed thickness/vp/vs/rho arrays, periods [period_offsets[m],
 * period_offsets[m+1]) and modes[m] reported modes. Outputs are packed per
 * model: velocity/group cell = out_offsets[m] + mode * T_m + f with
 * out_offsets[m] = sum_{i<m} modes[i] T_i, flags/levels/counts index
 * period_offsets[m] + f. The reference requires rectangular batches (padding,
 * python/modisp/models.py:224-270); this entry point is its ragged equivalent. */
int md_rocm_run_ragged(md_rocm_solver *s, const double *thickness, const double *vp, const double *vs,
                       const double *rho, const int64_t *row_offsets, int B, const double *periods,
                       const int64_t *period_offsets, const int32_t *modes, const double *bounds,
                       double c_tol_rel, int wave, double *velocity, uint32_t *flags, int32_t *levels,
                       int32_t *counts, double *group, uint32_t *group_flags, uint32_t *prediction_flags);

/* Ragged batch resident in DEVICE memory (HBM), the device counterpart of
 * md_rocm_run_ragged: every pointer is a device address. row_offsets,
 * period_offsets and cell_offsets have B + 1 entries (model m owns rows
 * [row_offsets[m], row_offsets[m+1]) of the packed thickness/vp/vs/rho arrays
 * of nrows rows, traces [period_offsets[m], period_offsets[m+1]) of the packed
 * periods and of flags/levels/counts (ntraces entries), and velocity/group cells
 * [cell_offsets[m], cell_offsets[m+1]) with cell = cell_offsets[m] + mode * T_m + f
 * (ncells entries)); modes[m] reported modes; bounds [B, 2]. max_rows and
 * max_modes bound every model's rows and modes (they size the kernels). The
 * preparation runs on the device; a model whose offsets are out of range or
 * inconsistent (cell range != modes x periods) is solved as an empty model
 * and counted (md_rocm_last_stats value 9), and no kernel touches memory
 * outside the given extents. With cfg.graph the whole sequence (preparation,
 * scan, rechecks, refinement, group velocity) is captured once and replayed
 * for calls with identical arguments. Asynchronous on the solver's stream.
 * Anchor continuation (cfg.continuation) plans from host periods and is
 * refused here. No reference counterpart (the reference requires rectangular
 * batches, python/modisp/models.py:224-270, and keeps them on the host). */
int md_rocm_run_ragged_device(md_rocm_solver *s, const double *thickness, const double *vp, const double *vs,
                              const double *rho, const int64_t *row_offsets, long long nrows, int B,
                              const double *periods, const int64_t *period_offsets, long long ntraces,
                              const int32_t *modes, const int64_t *cell_offsets, long long ncells,
                              const double *bounds, int max_rows, int max_modes, double c_tol_rel, int wave,
                              double *velocity, uint32_t *flags, int32_t *levels, int32_t *counts, double *group,
                              uint32_t *group_flags);

/* md_rocm_autotune for a device ragged batch (arguments as
 * md_rocm_run_ragged_device; outputs are overwritten). No reference
 * counterpart. */
int md_rocm_autotune_ragged_device(md_rocm_solver *s, const double *thickness, const double *vp, const double *vs,
                                   const double *rho, const int64_t *row_offsets, long long nrows, int B,
                                   const double *periods, const int64_t *period_offsets, long long ntraces,
                                   const int32_t *modes, const int64_t *cell_offsets, long long ncells,
                                   const double *bounds, int max_rows, int max_modes, double c_tol_rel, int wave,
                                   double *velocity, uint32_t *flags, int32_t *levels, int32_t *counts,
                                   double *group, uint32_t *group_flags, int *scan_waves, int *consumer_blocks);

/* Wait for the solver's stream; returns the first asynchronous error. */
int md_rocm_synchronize(md_rocm_solver *s);

/* Device time of the last run in milliseconds, from HIP events recorded on the
 * solver's stream (summed over trace chunks): 0 whole call, 1 grid scan,
 * 2 FP64 rescan of overflowed FP32 traces, 3 FP64 recheck + trough probes,
 * 4 bracket assembly, 5 refinement + output, 6 group velocity.
 * Returns the number of values written (at most 7). */
int md_rocm_last_timing(md_rocm_solver *s, double *ms, int n);

/* Work counters of the last run (the stats array md_rocm_run returns, also
 * for md_rocm_run_device): 0 traces, 1 FP64 rechecks of FP32-uncertain nodes,
 * 2 hidden-pair probes, 3 FP64 rescans of overflowed FP32 traces, 4 Love phase
 * searches, 5 secular evaluations of the grid scans, 6 FP64 evaluations of
 * the bracket refinement, 7 FP32 and 8 FP64 evaluations of the hidden-pair
 * probes, 9 models of the last md_rocm_run_ragged_device call with
 * inconsistent offsets (0 after other entry points), and for the last run
 * with anchor continuation: 10 continuation targets, 11 targets accepted
 * without the full pipeline, 12 FP64 evaluations (secular or Love phase) of
 * the continuation kernel. Synchronizes the solver's stream; returns the
 * number of values written (at most 13) or -1 on error.
 * No reference counterpart: the reference returns these only from run()
 * (cpp/cuda/qe.cu:815-821). */
int md_rocm_last_stats(md_rocm_solver *s, uint64_t *stats, int n);

/* Automatic search window [B, 2] (reference _cpu.bounds, cpp/cpu/solver.cpp:781-790):
 * lower = 1/8 of the slowest Rayleigh/Scholte/Love reference speed unless
 * auto_low == 0 (then `low`), upper = min(high or Vs_halfspace, nextafter(Vs_halfspace, 0)). */
int md_rocm_bounds(const double *thickness, const double *vp, const double *vs, const double *rho, int B, int L,
                   int wave, int auto_low, double low, int auto_high, double high, double *out);

/* Secular values on the scan grid of each (model, period), computed on the
 * device with this backend's FP32 (fp32 = 1) or FP64 evaluator: host inputs
 * [B, L], periods [nf], bounds [B, 2]; host output [B, nf, nc] float64 with
 * node i at lo + i (hi - lo) / (nc - 1). Diagnostic counterpart of the
 * reference's probe_grid / md_dunkin_grid[_f32] (cpp/cpu/solver.cpp:811-816,
 * cpp/cpu/dunkin_ref.cpp:41-83). */
int md_rocm_probe_grid(const double *thickness, const double *vp, const double *vs, const double *rho, int B, int L,
                       const double *periods, int nf, const double *bounds, int nc, int wave, int fp32, double *out);

/* Compiled resources of a search kernel (the fields of the reference's
 * qe_kernel_report / love_phase_kernel_report dictionaries). "Per CU" here is
 * the reference's "per SM". */
typedef struct md_rocm_kernel_report {
  int registers_per_thread;     /* VGPRs */
  int local_bytes_per_thread;   /* scratch (register spills) */
  int static_shared_bytes;      /* LDS */
  int dynamic_shared_bytes;
  int search_block_threads;
  int max_active_blocks_per_cu; /* occupancy at that block size */
  int configured_blocks_per_cu; /* persistent consumer blocks per CU (0 in the config = 8) */
  int table_capacity_per_model; /* layer rows held in LDS by the scan */
  int cache_points_per_trace;   /* scan events kept per trace */
} md_rocm_kernel_report;

/* The kernel that scans a model of L rows (1 <= L <= 256) with the FP32
 * (single_precision = 1) or FP64 evaluator, at the given config's scan shape
 * (cfg may be NULL). accuracy and shape select among the reference's CUDA
 * search kernels; this backend has one scan kernel for all of them.
 * Reference: qe_kernel_report (cpp/cuda/qe.cu:1143-1164). */
int md_rocm_qe_kernel_report(int L, int accuracy, int shape, int single_precision, const md_rocm_config *cfg,
                             md_rocm_kernel_report *out);

/* The Love phase search kernel. Reference: love_phase_kernel_report
 * (cpp/cuda/qe.cu:1131-1139). */
int md_rocm_love_phase_kernel_report(int L, int single_precision, md_rocm_kernel_report *out);

/* Continuity screen of completed phases, host-side (reference
 * continuity_screen, cpp/cpu/continuity_screen.hpp:8-67): periods [T] in the
 * caller's order, phase [B, K, T] (float32 if phase_is_f32, else float64),
 * prediction flags [B, T] or NULL, bilateral (compare with neighbouring
 * periods). Writes flags [B, T] (8 no bilateral neighbours, 16 envelope jump,
 * 32 FP32 phases of adjacent modes within 16 ulp, 64 continuation review) and
 * *suspect = 1 if any trace has 16, 32 or 64. */
int md_rocm_continuity_screen(const double *periods, int T, const void *phase, int phase_is_f32, int B, int K,
                              const uint32_t *prediction, int bilateral, uint8_t *flags, int *suspect);

/* Per-cell result masks of a solve, host-side (reference DispersionBatch,
 * python/modisp/__init__.py:100-123, computes them with numpy): velocity
 * [B, modes, T] (float32 if velocity_is_f32, else float64), trace flags and
 * root counts [B, T], group velocity and its flags [B, modes, T] (both NULL
 * when group_mask is NULL). Writes status [B, modes, T] (0 valid, 1 not found,
 * 2 unresolved, 3 rejected), mode_mask (finite), close_root_mask (a gap to a
 * neighbouring mode at most close_root_rtol times the upper velocity, in the
 * velocities' precision), close_root_risk [B, T] and group_mask (finite group
 * velocity with zero flags), each 0 or 1 per byte. Models are spread over
 * host threads. Returns 0, or 2 on a null pointer or bad size. */
int md_rocm_result_masks(const void *velocity, int velocity_is_f32, const uint32_t *flags, const int32_t *counts,
                         const double *group, const uint32_t *group_flags, int B, int modes, int T,
                         double close_root_rtol, uint8_t *status, uint8_t *mode_mask, uint8_t *close_root_mask,
                         uint8_t *close_root_risk, uint8_t *group_mask);

/* Group velocities of one confirmed trace for the continuity review.
 * Replaces the plugin function _review_groups (reference
 * cpp/common/review_groups.hpp:39-63, cpp/cuda/review_groups.cu:6-35):
 * packed = one model's h, vp, vs, rho ([layers] each) followed by its ordered
 * roots (guards included, at most 256), so size = 4 layers + roots. Writes
 * group[modes] (NaN where flagged) and flags[modes] with the group-velocity
 * flag contract (modes beyond the roots: MISSING_ROOT; neighbour midpoints as
 * windows, lower / upper at the ends) and the number of derivative
 * evaluations. secular must be 1. Returns 2 on invalid geometry
 * ("invalid review trace", as the reference), 3 on allocation failure. */
int md_rocm_review_groups(const double *packed, size_t size, int layers, int modes, int wave, int secular,
                          double period, double lower, double upper, double group_rtol, double close_rtol,
                          double *group, uint32_t *flags, int *evaluations);

/* Device helpers (reference cpp/cuda/qe.cu:1140-1142, bindings.cpp:354-357). */
int md_rocm_available(int device); /* 1 if the device exists and is gfx950-capable */
int md_rocm_set_device(int device);
int md_rocm_memory_info(size_t *free_bytes, size_t *total_bytes);
int md_rocm_device_name(int device, char *buf, int len);

/* Thread-local message of the last failed call. */
const char *md_rocm_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MODISP_ROCM_H */
