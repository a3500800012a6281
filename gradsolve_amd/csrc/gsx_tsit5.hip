// gsx_tsit5.hip -- Tsit5 forward+record, replay and adjoint kernels for gfx950 (one
// trajectory per lane; the per-lane bodies live in gsx_core.h).
//
// This translation unit is compiled twice: as namespace gsx with the default fused
// multiply-add contraction, and with -DGSX_NS=gsx_strict -ffp-contract=off.
#include <hip/hip_runtime.h>

#include "gsx_internal.h"
#include "gsx_tsit5_kernels.h"

namespace GSX_NS {

using ::gsx::check_launch;
using ::gsx::kMaxCkptInterval;
using ::gsx::set_error;

namespace {

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
    if (a.nsave > 0 || a.pbar_in)
      tsit5_adj_pair_kernel<F, true><<<grid, kFwdBlock, 0, s>>>(a);
    else
      tsit5_adj_pair_kernel<F, false><<<grid, kFwdBlock, 0, s>>>(a);
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

template <class F>
int launch_saveat(const SaveArgs& a, hipStream_t s) {
  if (a.n == 0) return GSX_OK;
  const unsigned grid = static_cast<unsigned>((a.n + kFwdBlock - 1) / kFwdBlock);
  tsit5_saveat_kernel<F><<<grid, kFwdBlock, 0, s>>>(a);
  return check_launch("tsit5_saveat_kernel");
}

template <class F>
int launch_saveat_vjp(const SaveVjpArgs& a, hipStream_t s) {
  if (a.n == 0) return GSX_OK;
  const unsigned grid = static_cast<unsigned>((a.n + kFwdBlock - 1) / kFwdBlock);
  tsit5_saveat_vjp_kernel<F><<<grid, kFwdBlock, 0, s>>>(a);
  return check_launch("tsit5_saveat_vjp_kernel");
}

#if defined(GSX_STRICT_ARITH)
constexpr ::gsx::rt::Family kRtFamily = ::gsx::rt::kTsit5Strict;
#else
constexpr ::gsx::rt::Family kRtFamily = ::gsx::rt::kTsit5;
#endif

}  // namespace

int tsit5_fwd_f32_dispatch(int field, int D, const FwdArgs& a, hipStream_t s) {
  switch (field) {
    case GSX_FIELD_LORENZ: return launch_fwd_f32<FLorenz>(a, s);
    case GSX_FIELD_LORENZ3: return launch_fwd_f32<FLorenz3>(a, s);
    default:
      if (::gsx::rt::is_runtime(field)) return ::gsx::rt::tsit5_fwd(field, ::gsx::rt::kTsit5F32, a, s);
      break;
  }
  return set_error(GSX_E_UNSUPPORTED, "tsit5 FP32 forward: field %d with D=%d is not built",
                   field, D);
}

int tsit5_replay_f32_dispatch(int field, int D, const ReplayArgs& a, hipStream_t s) {
  switch (field) {
    case GSX_FIELD_LORENZ: return launch_replay_f32<FLorenz>(a, s);
    case GSX_FIELD_LORENZ3: return launch_replay_f32<FLorenz3>(a, s);
    default:
      if (::gsx::rt::is_runtime(field)) return ::gsx::rt::tsit5_replay(field, ::gsx::rt::kTsit5F32, a, s);
      break;
  }
  return set_error(GSX_E_UNSUPPORTED, "tsit5 FP32 replay: field %d with D=%d is not built",
                   field, D);
}

int tsit5_adj_f32_dispatch(int field, int D, const AdjArgs& a, hipStream_t s) {
  switch (field) {
    case GSX_FIELD_LORENZ: return launch_adj_f32<FLorenz>(a, s);
    case GSX_FIELD_LORENZ3: return launch_adj_f32<FLorenz3>(a, s);
    default:
      if (::gsx::rt::is_runtime(field)) return ::gsx::rt::tsit5_adj(field, ::gsx::rt::kTsit5F32, a, s);
      break;
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
    default:
      if (::gsx::rt::is_runtime(field)) return ::gsx::rt::tsit5_fwd(field, kRtFamily, a, s);
      break;
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
    default:
      if (::gsx::rt::is_runtime(field)) return ::gsx::rt::tsit5_replay(field, kRtFamily, a, s);
      break;
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
    default:
      if (::gsx::rt::is_runtime(field)) return ::gsx::rt::tsit5_adj(field, kRtFamily, a, s);
      break;
  }
  return set_error(GSX_E_UNSUPPORTED, "tsit5 adjoint: field %d with D=%d is not built", field,
                   D);
}

int tsit5_saveat_dispatch(int field, int D, const SaveArgs& a, hipStream_t s) {
  switch (field) {
    case GSX_FIELD_LORENZ: return launch_saveat<FLorenz>(a, s);
    case GSX_FIELD_LORENZ3: return launch_saveat<FLorenz3>(a, s);
    case GSX_FIELD_VDP: return launch_saveat<FVdp>(a, s);
    case GSX_FIELD_LINEAR_DECAY:
      if (D == 1) return launch_saveat<FLinearDecay<1>>(a, s);
      if (D == 2) return launch_saveat<FLinearDecay<2>>(a, s);
      if (D == 3) return launch_saveat<FLinearDecay<3>>(a, s);
      break;
    default:
      if (::gsx::rt::is_runtime(field)) return ::gsx::rt::tsit5_saveat(field, kRtFamily, a, s);
      break;
  }
  return set_error(GSX_E_UNSUPPORTED, "tsit5 saveat: field %d with D=%d is not built", field,
                   D);
}

int tsit5_saveat_vjp_dispatch(int field, int D, const SaveVjpArgs& a, hipStream_t s) {
  switch (field) {
    case GSX_FIELD_LORENZ: return launch_saveat_vjp<FLorenz>(a, s);
    case GSX_FIELD_LORENZ3: return launch_saveat_vjp<FLorenz3>(a, s);
    case GSX_FIELD_VDP: return launch_saveat_vjp<FVdp>(a, s);
    case GSX_FIELD_LINEAR_DECAY:
      if (D == 1) return launch_saveat_vjp<FLinearDecay<1>>(a, s);
      if (D == 2) return launch_saveat_vjp<FLinearDecay<2>>(a, s);
      if (D == 3) return launch_saveat_vjp<FLinearDecay<3>>(a, s);
      break;
    default:
      if (::gsx::rt::is_runtime(field))
        return ::gsx::rt::tsit5_saveat_vjp(field, kRtFamily, a, s);
      break;
  }
  return set_error(GSX_E_UNSUPPORTED, "tsit5 saveat adjoint: field %d with D=%d is not built",
                   field, D);
}

}  // namespace GSX_NS
