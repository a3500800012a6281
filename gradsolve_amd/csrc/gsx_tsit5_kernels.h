// gsx_tsit5_kernels.h -- the Tsit5 kernels (one trajectory per lane) as templates over the
// field. The library instantiates them for its built-in fields (gsx_tsit5.hip); fields
// registered at run time instantiate the same templates through hiprtc (gsx_runtime.hip).
#pragma once

#include "gsx_core.h"
#include "gsx_f32.h"

namespace GSX_NS {

// kRecCk: the launch records both the tape and the checkpoints (see tsit5_fwd_lane).
template <class F, bool kRecCk>
__global__ __launch_bounds__(kFwdBlock) void tsit5_fwd_kernel(FwdArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  if (j >= a.n) return;
  tsit5_fwd_lane<F, kRecCk>(a, j);
}

// The segment buffer holds, for each lane, the states at the start of the (at most CMAX)
// steps of the segment being reversed: seg[(s*DCK + d)*kAdjBlock + lane]. Consecutive lanes
// touch consecutive doubles, so every ds_read/ds_write_b64 of a wavefront is conflict-free.
template <class F, int CMAX>
__global__ __launch_bounds__(kAdjBlock) void tsit5_adj_kernel(AdjArgs a) {
  __shared__ double seg[CMAX * ckpt_rows<F>() * kAdjBlock];
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kAdjBlock + threadIdx.x;
  const int32_t mtop = wave_max(j < a.n ? segment_count(a, j) : 0);
  if (j >= a.n) return;
  tsit5_adj_lane<F>(a, j, seg + threadIdx.x, kAdjBlock, mtop);
}

template <class F>
__global__ __launch_bounds__(kFwdBlock) void tsit5_fwd_f32_kernel(FwdArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  if (j >= a.n) return;
  tsit5_fwd_lane_f32<F>(a, j);
}

template <class F>
__global__ __launch_bounds__(kFwdBlock) void tsit5_replay_f32_kernel(ReplayArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  if (j >= a.n) return;
  tsit5_replay_lane_f32<F>(a, j);
}

template <class F, int CMAX>
__global__ __launch_bounds__(kAdjBlock) void tsit5_adj_f32_kernel(AdjArgs a) {
  __shared__ double seg[CMAX * F::D * kAdjBlock];
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kAdjBlock + threadIdx.x;
  const int32_t mtop = wave_max(j < a.n ? segment_count(a, j) : 0);
  if (j >= a.n) return;
  tsit5_adj_lane_f32<F>(a, j, seg + threadIdx.x, kAdjBlock, mtop);
}

template <class F>
__global__ __launch_bounds__(kFwdBlock) void tsit5_replay_kernel(ReplayArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  if (j >= a.n) return;
  tsit5_replay_lane<F>(a, j);
}

// Checkpoint interval 2: the pair adjoint (tsit5_adj_pair_segment; no LDS; kSaves = false when
// no save cotangents or incoming pbar are given, and only that variant takes the VGPR cap below:
// with the save injection the cap would spill 24 VGPRs). The segment loop
// runs in step across the wavefront, from its largest segment count down, so the active lanes
// read the same checkpoint and tape rows; a lane idles above its own count. Fields with at
// most three states and one parameter are held to 128 VGPRs (four waves per SIMD; Lorenz:
// 136 -> 128 VGPRs with two spilled, measured 1.3% faster at N = 2^20); other fields keep
// the compiler's choice (Lorenz with three parameters would spill 112 bytes per lane).
template <class F>
constexpr int pair_waves_per_eu() { return (F::D <= 3 && F::P == 1) ? 4 : 1; }

template <class F, bool kSaves>
__global__ __launch_bounds__(kFwdBlock)
__attribute__((amdgpu_waves_per_eu(kSaves ? 1 : pair_waves_per_eu<F>())))
void tsit5_adj_pair_kernel(AdjArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  const bool valid = j < a.n;
  const int32_t na = valid ? a.n_acc[j] : 0;
  const int32_t mmax = wave_max((na + 1) / 2);
  if (!valid) return;
  tsit5_adj_pair_lane<F, kSaves>(a, j, na, mmax);
}

// Dense output on a recorded mesh (tsit5_saveat_lane) and the VJPs of its re-steps
// (tsit5_saveat_vjp_lane), one trajectory per lane.
template <class F>
__global__ __launch_bounds__(kFwdBlock) void tsit5_saveat_kernel(SaveArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  if (j >= a.n) return;
  tsit5_saveat_lane<F>(a, j);
}

template <class F>
__global__ __launch_bounds__(kFwdBlock) void tsit5_saveat_vjp_kernel(SaveVjpArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  if (j >= a.n) return;
  tsit5_saveat_vjp_lane<F>(a, j);
}

}  // namespace GSX_NS
