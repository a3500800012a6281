// gsx_tsit5.hip -- Tsit5 forward+record, replay and adjoint kernels for gfx950 (one
// trajectory per lane; the per-lane bodies live in gsx_core.h).
//
// This translation unit is compiled twice: as namespace gsx with the default fused
// multiply-add contraction, and with -DGSX_NS=gsx_strict -ffp-contract=off.
#include <hip/hip_runtime.h>

#include "gsx_internal.h"
#include "gsx_f32.h"

namespace GSX_NS {

using ::gsx::check_launch;
using ::gsx::kMaxCkptInterval;
using ::gsx::set_error;

namespace {

constexpr int kFwdBlock = 256;
constexpr int kAdjBlock = 64;  // one wavefront per workgroup: LDS is allocated per wave

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
  if (j >= a.n) return;
  tsit5_adj_lane<F>(a, j, seg + threadIdx.x, kAdjBlock);
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
  if (j >= a.n) return;
  tsit5_adj_lane_f32<F>(a, j, seg + threadIdx.x, kAdjBlock);
}

template <class F>
__global__ __launch_bounds__(kFwdBlock) void tsit5_replay_kernel(ReplayArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  if (j >= a.n) return;
  tsit5_replay_lane<F>(a, j);
}

template <class F>
int launch_replay(const ReplayArgs& a, hipStream_t s) {
  if (a.n == 0) return GSX_OK;
  const unsigned grid = static_cast<unsigned>((a.n + kFwdBlock - 1) / kFwdBlock);
  tsit5_replay_kernel<F><<<grid, kFwdBlock, 0, s>>>(a);
  return check_launch("tsit5_replay_kernel");
}

template <class F>
int launch_fwd(const FwdArgs& a, hipStream_t s) {
  if (a.n == 0) return GSX_OK;
  const unsigned grid = static_cast<unsigned>((a.n + kFwdBlock - 1) / kFwdBlock);
  if (a.tape != nullptr && a.ckpt != nullptr)
    tsit5_fwd_kernel<F, true><<<grid, kFwdBlock, 0, s>>>(a);
  else
    tsit5_fwd_kernel<F, false><<<grid, kFwdBlock, 0, s>>>(a);
  return check_launch("tsit5_fwd_kernel");
}

// Checkpoint interval 2: the pair adjoint (tsit5_adj_pair_segment; no LDS). The segment loop
// runs in step across the wavefront, from its largest segment count down, so the active lanes
// read the same checkpoint and tape rows; a lane idles above its own count. Fields with at
// most three states and one parameter are held to 128 VGPRs (four waves per SIMD; Lorenz:
// 136 -> 128 VGPRs with two spilled, measured 1.3% faster at N = 2^20); other fields keep
// the compiler's choice (Lorenz with three parameters would spill 112 bytes per lane).
template <class F>
constexpr int pair_waves_per_eu() { return (F::D <= 3 && F::P == 1) ? 4 : 1; }

template <class F>
__global__ __launch_bounds__(kFwdBlock) __attribute__((amdgpu_waves_per_eu(pair_waves_per_eu<F>())))
void tsit5_adj_pair_kernel(AdjArgs a) {
  constexpr int D = F::D, P = F::P;
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  const bool valid = j < a.n;
  const int32_t na = valid ? a.n_acc[j] : 0;
  const int32_t nseg = (na + 1) / 2;
  int32_t mmax = nseg;
  GSX_UNROLL for (int off = 32; off >= 1; off >>= 1) {
    const int32_t o = __shfl_xor(mmax, off);
    mmax = o > mmax ? o : mmax;
  }
  if (!valid) return;
  const int64_t n = a.n;
  double p[P], pb[P], lam[D];
  GSX_UNROLL for (int q = 0; q < P; ++q) {
    p[q] = a.p[q * n + j];
    pb[q] = 0.0;
  }
  GSX_UNROLL for (int d = 0; d < D; ++d) lam[d] = a.ybar[d * n + j];
  for (int32_t m = mmax - 1; m >= 0; --m) {
    if (m < nseg) tsit5_adj_pair_segment<F>(a, j, m, na, p, lam, pb);
  }
  GSX_UNROLL for (int d = 0; d < D; ++d) a.y0bar[d * n + j] = lam[d];
  GSX_UNROLL for (int q = 0; q < P; ++q) a.pbar[q * n + j] = pb[q];
}

template <class F, int CMAX>
int launch_adj_c(const AdjArgs& a, hipStream_t s) {
  const unsigned grid = static_cast<unsigned>((a.n + kAdjBlock - 1) / kAdjBlock);
  tsit5_adj_kernel<F, CMAX><<<grid, kAdjBlock, 0, s>>>(a);
  return check_launch("tsit5_adj_kernel");
}

template <class F>
int launch_adj(const AdjArgs& a, hipStream_t s) {
  if (a.n == 0) return GSX_OK;
  const int C = a.ckpt_interval;
  if (C <= 1) return launch_adj_c<F, 1>(a, s);
  if (C == 2) {
    const unsigned grid = static_cast<unsigned>((a.n + kFwdBlock - 1) / kFwdBlock);
    tsit5_adj_pair_kernel<F><<<grid, kFwdBlock, 0, s>>>(a);
    return check_launch("tsit5_adj_pair_kernel");
  }
  if (C <= 4) return launch_adj_c<F, 4>(a, s);
  if (C <= 8) return launch_adj_c<F, 8>(a, s);
  if (C <= 16) return launch_adj_c<F, 16>(a, s);
  if (C <= 32) return launch_adj_c<F, 32>(a, s);
  return set_error(GSX_E_INVALID, "ckpt_interval %d exceeds the adjoint's maximum of %d", C,
                   kMaxCkptInterval);
}

template <class F>
int launch_fwd_f32(const FwdArgs& a, hipStream_t s) {
  if (a.n == 0) return GSX_OK;
  const unsigned grid = static_cast<unsigned>((a.n + kFwdBlock - 1) / kFwdBlock);
  tsit5_fwd_f32_kernel<F><<<grid, kFwdBlock, 0, s>>>(a);
  return check_launch("tsit5_fwd_f32_kernel");
}

template <class F>
int launch_replay_f32(const ReplayArgs& a, hipStream_t s) {
  if (a.n == 0) return GSX_OK;
  const unsigned grid = static_cast<unsigned>((a.n + kFwdBlock - 1) / kFwdBlock);
  tsit5_replay_f32_kernel<F><<<grid, kFwdBlock, 0, s>>>(a);
  return check_launch("tsit5_replay_f32_kernel");
}

template <class F, int CMAX>
int launch_adj_f32_c(const AdjArgs& a, hipStream_t s) {
  const unsigned grid = static_cast<unsigned>((a.n + kAdjBlock - 1) / kAdjBlock);
  tsit5_adj_f32_kernel<F, CMAX><<<grid, kAdjBlock, 0, s>>>(a);
  return check_launch("tsit5_adj_f32_kernel");
}

template <class F>
int launch_adj_f32(const AdjArgs& a, hipStream_t s) {
  if (a.n == 0) return GSX_OK;
  const int C = a.ckpt_interval;
  if (C <= 1) return launch_adj_f32_c<F, 1>(a, s);
  if (C <= 4) return launch_adj_f32_c<F, 4>(a, s);
  if (C <= 8) return launch_adj_f32_c<F, 8>(a, s);
  if (C <= 16) return launch_adj_f32_c<F, 16>(a, s);
  if (C <= 32) return launch_adj_f32_c<F, 32>(a, s);
  return set_error(GSX_E_INVALID, "ckpt_interval %d exceeds the adjoint's maximum of %d", C,
                   kMaxCkptInterval);
}

}  // namespace

int tsit5_fwd_f32_dispatch(int field, int D, const FwdArgs& a, hipStream_t s) {
  switch (field) {
    case GSX_FIELD_LORENZ: return launch_fwd_f32<FLorenz>(a, s);
    case GSX_FIELD_LORENZ3: return launch_fwd_f32<FLorenz3>(a, s);
    default: break;
  }
  return set_error(GSX_E_UNSUPPORTED, "tsit5 FP32 forward: field %d with D=%d is not built",
                   field, D);
}

int tsit5_replay_f32_dispatch(int field, int D, const ReplayArgs& a, hipStream_t s) {
  switch (field) {
    case GSX_FIELD_LORENZ: return launch_replay_f32<FLorenz>(a, s);
    case GSX_FIELD_LORENZ3: return launch_replay_f32<FLorenz3>(a, s);
    default: break;
  }
  return set_error(GSX_E_UNSUPPORTED, "tsit5 FP32 replay: field %d with D=%d is not built",
                   field, D);
}

int tsit5_adj_f32_dispatch(int field, int D, const AdjArgs& a, hipStream_t s) {
  switch (field) {
    case GSX_FIELD_LORENZ: return launch_adj_f32<FLorenz>(a, s);
    case GSX_FIELD_LORENZ3: return launch_adj_f32<FLorenz3>(a, s);
    default: break;
  }
  return set_error(GSX_E_UNSUPPORTED, "tsit5 FP32 adjoint: field %d with D=%d is not built",
                   field, D);
}

int tsit5_fwd_dispatch(int field, int D, const FwdArgs& a, hipStream_t s) {
  switch (field) {
    case GSX_FIELD_LORENZ: return launch_fwd<FLorenz>(a, s);
    case GSX_FIELD_LORENZ3: return launch_fwd<FLorenz3>(a, s);
    case GSX_FIELD_VDP: return launch_fwd<FVdp>(a, s);
    case GSX_FIELD_LINEAR_DECAY:
      if (D == 1) return launch_fwd<FLinearDecay<1>>(a, s);
      if (D == 2) return launch_fwd<FLinearDecay<2>>(a, s);
      if (D == 3) return launch_fwd<FLinearDecay<3>>(a, s);
      break;
    default: break;
  }
  return set_error(GSX_E_UNSUPPORTED, "tsit5 forward: field %d with D=%d is not built", field,
                   D);
}

int tsit5_replay_dispatch(int field, int D, const ReplayArgs& a, hipStream_t s) {
  switch (field) {
    case GSX_FIELD_LORENZ: return launch_replay<FLorenz>(a, s);
    case GSX_FIELD_LORENZ3: return launch_replay<FLorenz3>(a, s);
    case GSX_FIELD_VDP: return launch_replay<FVdp>(a, s);
    case GSX_FIELD_LINEAR_DECAY:
      if (D == 1) return launch_replay<FLinearDecay<1>>(a, s);
      if (D == 2) return launch_replay<FLinearDecay<2>>(a, s);
      if (D == 3) return launch_replay<FLinearDecay<3>>(a, s);
      break;
    default: break;
  }
  return set_error(GSX_E_UNSUPPORTED, "tsit5 replay: field %d with D=%d is not built", field,
                   D);
}

int tsit5_adj_dispatch(int field, int D, const AdjArgs& a, hipStream_t s) {
  switch (field) {
    case GSX_FIELD_LORENZ: return launch_adj<FLorenz>(a, s);
    case GSX_FIELD_LORENZ3: return launch_adj<FLorenz3>(a, s);
    case GSX_FIELD_VDP: return launch_adj<FVdp>(a, s);
    case GSX_FIELD_LINEAR_DECAY:
      if (D == 1) return launch_adj<FLinearDecay<1>>(a, s);
      if (D == 2) return launch_adj<FLinearDecay<2>>(a, s);
      if (D == 3) return launch_adj<FLinearDecay<3>>(a, s);
      break;
    default: break;
  }
  return set_error(GSX_E_UNSUPPORTED, "tsit5 adjoint: field %d with D=%d is not built", field,
                   D);
}

}  // namespace GSX_NS
