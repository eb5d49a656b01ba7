// Per-cell result masks of a solve, on the host (modisp-rocm backend, C++17).
//
// The reference's DispersionBatch builds these with numpy over the
// [B, modes, T] result (python/modisp/__init__.py:100-123): the finite mask,
// the close-root mask (neighbouring modes whose gap is at most
// close_root_rtol times the upper one), the per-trace close-root risk, and
// the status code of every cell (valid; rejected where the trace's F
// acceptance failed below its root count; unresolved where the trace did not
// converge or kept an unresolved bracket; else not found). Here each model is
// one pass over its cells and the models are spread over host threads; every
// output is what the numpy expressions give, with the gap and its threshold
// in the velocities' own precision (float32 arithmetic for float32 results,
// as numpy computes them).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/modisp_rocm.h"

namespace {

constexpr uint32_t kUnresolvedBracket = 1u, kSearchUnconverged = 4u, kFRejected = 128u, kCloseRootRisk = 4096u;
constexpr uint8_t kValid = 0, kNotFound = 1, kUnresolved = 2, kRejected = 3;

template <class V>
void model_masks(int b, const V *v, const uint32_t *flags, const int32_t *counts, const double *group,
                 const uint32_t *gflags, int M, int T, V rtol, uint8_t *status, uint8_t *mode_mask, uint8_t *close,
                 uint8_t *risk, uint8_t *group_mask) {
  const size_t c0 = static_cast<size_t>(b) * M * T, t0 = static_cast<size_t>(b) * T;
  const V *vb = v + c0;
  uint8_t *cb = close + c0;
  std::fill_n(cb, static_cast<size_t>(M) * T, uint8_t(0));
  for (int m = 1; m < M; ++m) {
    const V *lo = vb + static_cast<size_t>(m - 1) * T, *hi = vb + static_cast<size_t>(m) * T;
    uint8_t *clo = cb + static_cast<size_t>(m - 1) * T, *chi = cb + static_cast<size_t>(m) * T;
    for (int t = 0; t < T; ++t) {
      const V gap = hi[t] - lo[t];
      const uint8_t c = std::isfinite(gap) && gap <= rtol * hi[t];
      clo[t] |= c;
      chi[t] |= c;
    }
  }
  for (int t = 0; t < T; ++t) {
    const uint32_t f = flags[t0 + t];
    uint8_t r = (f & kCloseRootRisk) != 0;
    for (int m = 0; m < M; ++m) r |= cb[static_cast<size_t>(m) * T + t];
    risk[t0 + t] = r;
  }
  for (int m = 0; m < M; ++m) {
    const size_t o = c0 + static_cast<size_t>(m) * T;
    for (int t = 0; t < T; ++t) {
      const uint32_t f = flags[t0 + t];
      const V x = v[o + t];
      const bool fin = std::isfinite(x);
      mode_mask[o + t] = fin;
      uint8_t s = kValid;
      if (!fin) {
        if (std::isnan(x) && (f & kFRejected) && m < counts[t0 + t]) s = kRejected;
        else s = (f & (kSearchUnconverged | kUnresolvedBracket)) ? kUnresolved : kNotFound;
      }
      status[o + t] = s;
      if (group_mask) group_mask[o + t] = std::isfinite(group[o + t]) && gflags[o + t] == 0;
    }
  }
}

template <class V>
void all_masks(const V *v, const uint32_t *flags, const int32_t *counts, const double *group, const uint32_t *gflags,
               int B, int M, int T, double rtol, uint8_t *status, uint8_t *mode_mask, uint8_t *close, uint8_t *risk,
               uint8_t *group_mask) {
  const V r = static_cast<V>(rtol);
  const long long cells = static_cast<long long>(B) * M * T;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int nthreads = static_cast<int>(std::min<long long>({16, hw, std::max(1LL, cells / 262144), B}));
  auto work = [&](int b0, int b1) {
    for (int b = b0; b < b1; ++b)
      model_masks<V>(b, v, flags, counts, group, gflags, M, T, r, status, mode_mask, close, risk, group_mask);
  };
  if (nthreads <= 1) {
    work(0, B);
    return;
  }
  std::vector<std::thread> pool;
  for (int i = 0; i < nthreads; ++i) {
    const int b0 = static_cast<int>(static_cast<long long>(B) * i / nthreads);
    const int b1 = static_cast<int>(static_cast<long long>(B) * (i + 1) / nthreads);
    try {
      pool.emplace_back(work, b0, b1);
    } catch (...) {  // a thread that cannot start: its models are done here
      work(b0, b1);
    }
  }
  for (std::thread &th : pool) th.join();
}

}  // namespace

extern "C" int md_rocm_result_masks(const void *velocity, int velocity_is_f32, const uint32_t *flags,
                                    const int32_t *counts, const double *group, const uint32_t *group_flags, int B,
                                    int modes, int T, double close_root_rtol, uint8_t *status, uint8_t *mode_mask,
                                    uint8_t *close_root_mask, uint8_t *close_root_risk, uint8_t *group_mask) {
  if (!velocity || !flags || !counts || !status || !mode_mask || !close_root_mask || !close_root_risk) return 2;
  if (B < 0 || modes < 1 || T < 0) return 2;
  if (group_mask && (!group || !group_flags)) return 2;
  if (velocity_is_f32)
    all_masks(static_cast<const float *>(velocity), flags, counts, group, group_flags, B, modes, T, close_root_rtol,
              status, mode_mask, close_root_mask, close_root_risk, group_mask);
  else
    all_masks(static_cast<const double *>(velocity), flags, counts, group, group_flags, B, modes, T, close_root_rtol,
              status, mode_mask, close_root_mask, close_root_risk, group_mask);
  return 0;
}
