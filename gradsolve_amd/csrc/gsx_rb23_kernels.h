// gsx_rb23_kernels.h -- the Rosenbrock23 kernels (one trajectory per lane) as templates over
// the field, shared like gsx_tsit5_kernels.h by the library and by fields compiled at run time.
#pragma once

#include "gsx_core.h"
#include "gsx_rb23.h"
#include "gsx_rb23_f32.h"

namespace GSX_NS {

template <class F>
__global__ __launch_bounds__(kFwdBlock) void rb23_fwd_kernel(FwdArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  if (j >= a.n) return;
  rb23_fwd_lane<F>(a, j);
}

template <class F>
__global__ __launch_bounds__(kFwdBlock) void rb23_replay_kernel(ReplayArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  if (j >= a.n) return;
  rb23_replay_lane<F>(a, j);
}

template <class F, int CMAX>
__global__ __launch_bounds__(kAdjBlock) void rb23_adj_kernel(AdjArgs a) {
  __shared__ double seg[CMAX * ckpt_rows<F>() * kAdjBlock];
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kAdjBlock + threadIdx.x;
  const int32_t mtop = wave_max(j < a.n ? segment_count(a, j) : 0);
  if (j >= a.n) return;
  rb23_adj_lane<F>(a, j, seg + threadIdx.x, kAdjBlock, mtop);
}

// Checkpoint interval 2: the pair adjoint (rb23_adj_pair_segment; no LDS), walked in step
// across the wavefront from its largest segment count down, as tsit5_adj_pair_kernel.
template <class F, bool kSaves>
__global__ __launch_bounds__(kFwdBlock) void rb23_adj_pair_kernel(AdjArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  const bool valid = j < a.n;
  const int32_t na = valid ? a.n_acc[j] : 0;
  const int32_t mmax = wave_max((na + 1) / 2);
  if (!valid) return;
  rb23_adj_pair_lane<F, kSaves>(a, j, na, mmax);
}

// Dense output on a recorded mesh and the VJPs of its re-steps, one trajectory per lane.
template <class F>
__global__ __launch_bounds__(kFwdBlock) void rb23_saveat_kernel(SaveArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  if (j >= a.n) return;
  rb23_saveat_lane<F>(a, j);
}

template <class F>
__global__ __launch_bounds__(kFwdBlock) void rb23_saveat_vjp_kernel(SaveVjpArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  if (j >= a.n) return;
  rb23_saveat_vjp_lane<F>(a, j);
}

// The FP32 variant (gsx_rb23_f32.h): FP32 stages on a compensated state, FP64 LU and solves.
template <class F>
__global__ __launch_bounds__(kFwdBlock) void rb23_fwd_f32_kernel(FwdArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  if (j >= a.n) return;
  rb23_fwd_lane_f32<F>(a, j);
}

template <class F>
__global__ __launch_bounds__(kFwdBlock) void rb23_replay_f32_kernel(ReplayArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  if (j >= a.n) return;
  rb23_replay_lane_f32<F>(a, j);
}

template <class F, int CMAX>
__global__ __launch_bounds__(kAdjBlock) void rb23_adj_f32_kernel(AdjArgs a) {
  __shared__ double seg[CMAX * F::D * kAdjBlock];
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kAdjBlock + threadIdx.x;
  const int32_t mtop = wave_max(j < a.n ? segment_count(a, j) : 0);
  if (j >= a.n) return;
  rb23_adj_lane_f32<F>(a, j, seg + threadIdx.x, kAdjBlock, mtop);
}

}  // namespace GSX_NS
