// gfx950 adjoint of the phase-velocity roots: dc/d(h, vp, vs, rho) of every
// layer by the implicit-function rule at a converged root,
//
//     dc/dp = -(dF/dp) / (dF/dc) = (dF/dp) c / (k dF/dk),   k = w / c,
//
// (no reference counterpart; the implicit formula is the one the reference
// uses for its group velocity, cpp/common/group_velocity.hpp:2-4). F is the
// normalised secular function of the propagator: every positive factor it
// carries -- the per-layer max-abs rescale and the e^-(p+q) attenuation of
// the compound matrix -- multiplies F, which vanishes at the root, so it
// cancels from the ratio and the factors can be held constant.
//
// Per root (one lane), in FP64:
//   1. forward sweep, half-space to surface, in one-direction dual numbers of
//      the wavenumber k: the states e_{j+1} entering each layer and the
//      scales s_j go to a per-lane tape in device memory, and dF/dk comes out
//      at the top;
//   2. reverse sweep, surface to half-space: the adjoint row vector
//      l_j = dF/de_j is carried down with the transposed layer matrix,
//      l_{j+1} = l_j C_j / s_j, and at each layer
//      dF/dp_j = l_j (dC_j/dp_j) e_{j+1} / s_j is the derivative of the
//      bilinear form l_j C_j e_{j+1}, taken in reverse mode by
//      rayleigh_layer_adjoint (one pass over the layer's wave factors and
//      their closed-form h and wavenumber derivatives, no dual matrix);
//      Love's 2x2 layer matrix is small enough for dual numbers seeded in
//      two parameters at a time;
//   3. the half-space vector and the water top contribute their own terms
//      (dual numbers; once per root).
// The propagator is recomputed here from the model, not kept from the root
// search. A block (one wavefront) takes `group` consecutive models at once
// and spreads their roots over its lanes; the weighted sums over each
// model's roots, sum_roots w dc/dp, accumulate in LDS and are written once.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>
#include "pipeline.hpp"
#include "secular.hpp"
#include "adjoint_root.hpp"

namespace mdr {

__device__ inline PrepRow adj_row(const AdjointArgs &a, int m, int j) {
  const size_t i = static_cast<size_t>(m) * a.L + j;
  return make_prep_row(a.h[i], a.vp[i], a.vs[i], a.rho[i]);
}

// One wavefront per block. Block b takes models [m0, m0 + group) for
// m0 = b * group, b * group + gridDim.x * group, ...; the group's
// modes * T cells are spread over the lanes in model-major order.
template <int Wave>
__global__ __launch_bounds__(64) void adjoint_kernel(AdjointArgs a) {
  extern __shared__ double g[];  // [group][4][L]
  const int lane = threadIdx.x;
  const int L = a.L, R = a.modes * a.T, per_model = 4 * L;
  double *tape = a.tape + blockIdx.x * 64 + lane;
  const int stride = a.tape_lanes;
  constexpr double two_pi = 6.283185307179586;
  for (int m0 = blockIdx.x * a.group; m0 < a.B; m0 += gridDim.x * a.group) {
    const int nm = min(a.group, a.B - m0);
    for (int i = lane; i < nm * per_model; i += 64) g[i] = 0.;
    __syncthreads();
    for (int q = lane; q < nm * R; q += 64) {
      const int mi = q / R, cell = q - mi * R;
      const int m = m0 + mi, mode = cell / a.T, f = cell - mode * a.T;
      const size_t o = static_cast<size_t>(m) * R + cell;
      const double c = a.roots[o], wgt = a.weight[o];
      if (!(c == c) || !(wgt == wgt) || wgt == 0.) continue;
      const double period = a.row_periods ? a.periods[static_cast<size_t>(m) * a.T + f] : a.periods[f];
      const int top = a.vs[static_cast<size_t>(m) * L] == 0. ? 1 : 0;
      double *gm = g + mi * per_model;
      auto row = [&](int j) { return adj_row(a, m, j); };
      auto add = [&](int p, int j, double v) { atomicAdd(&gm[p * L + j], v); };
      auto tp = [&](int j, int i) -> double & { return tape[(static_cast<size_t>(j) * 6 + i) * stride]; };
      if (Wave == 0)
        rayleigh_root_adjoint(row, add, tp, L, top, two_pi / period, c, wgt);
      else
        love_root_adjoint(row, add, tp, L, top, two_pi / period, c, wgt);
    }
    __syncthreads();
    for (int i = lane; i < nm * per_model; i += 64) {
      const int mi = i / per_model, pj = i - mi * per_model, p = pj / L, j = pj - p * L;
      const size_t o = static_cast<size_t>(m0 + mi) * L + j;
      (p == 0 ? a.gh : p == 1 ? a.gvp : p == 2 ? a.gvs : a.grho)[o] = g[i];
    }
    __syncthreads();
  }
}

// FP64 -> FP32 velocities for callers that take float32 results: one
// grid-stride pass on the device, so half the bytes cross to the host.
__global__ __launch_bounds__(256) void narrow_kernel(const double *in, float *out, size_t n) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = static_cast<float>(in[i]);
}

hipError_t launch_narrow(const double *in, float *out, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(narrow_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, in, out, n);
  return hipGetLastError();
}

hipError_t launch_adjoint(const AdjointArgs &a, int blocks, hipStream_t s) {
  const size_t lds = static_cast<size_t>(a.group) * 4 * a.L * sizeof(double);
  if (a.wave == 0)
    hipLaunchKernelGGL(adjoint_kernel<0>, dim3(blocks), dim3(64), lds, s, a);
  else
    hipLaunchKernelGGL(adjoint_kernel<1>, dim3(blocks), dim3(64), lds, s, a);
  return hipGetLastError();
}

}  // namespace mdr
