// gsx_rodas5p_kernels.h -- the Rodas5P kernels (one trajectory per lane) as templates over the
// field, shared like gsx_rb23_kernels.h by the library and by fields compiled at run time.
#pragma once

#include "gsx_core.h"
#include "gsx_rodas5p.h"

namespace GSX_NS {

template <class F>
__global__ __launch_bounds__(kFwdBlock) void r5_fwd_kernel(FwdArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  if (j >= a.n) return;
  r5_fwd_lane<F>(a, j);
}

template <class F>
__global__ __launch_bounds__(kFwdBlock) void r5_replay_kernel(ReplayArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  if (j >= a.n) return;
  r5_replay_lane<F>(a, j);
}

template <class F, int CMAX>
__global__ __launch_bounds__(kAdjBlock) void r5_adj_kernel(AdjArgs a) {
  __shared__ double seg[CMAX * ckpt_rows<F>() * kAdjBlock];
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kAdjBlock + threadIdx.x;
  const int32_t mtop = wave_max(j < a.n ? segment_count(a, j) : 0);
  if (j >= a.n) return;
  r5_adj_lane<F>(a, j, seg + threadIdx.x, kAdjBlock, mtop);
}

template <class F>
__global__ __launch_bounds__(kFwdBlock) void r5_saveat_kernel(SaveArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  if (j >= a.n) return;
  r5_saveat_lane<F>(a, j);
}

template <class F>
__global__ __launch_bounds__(kFwdBlock) void r5_saveat_vjp_kernel(SaveVjpArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  if (j >= a.n) return;
  r5_saveat_vjp_lane<F>(a, j);
}

template <class F>
__global__ __launch_bounds__(kFwdBlock) void r5_saveat_dense_kernel(SaveArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  if (j >= a.n) return;
  r5_saveat_dense_lane<F>(a, j);
}

}  // namespace GSX_NS
