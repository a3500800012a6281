// gsx_rb23.hip -- Rosenbrock23 forward+record, replay and adjoint kernels for gfx950 (one
// trajectory per lane; per-lane bodies in gsx_rb23.h). Compiled twice like gsx_tsit5.hip:
// namespace gsx (contraction on) and -DGSX_NS=gsx_strict -ffp-contract=off.
#include <hip/hip_runtime.h>

#include "gsx_internal.h"
#include "gsx_rb23_kernels.h"

namespace GSX_NS {

using ::gsx::check_launch;
using ::gsx::kMaxCkptInterval;
using ::gsx::set_error;

namespace {

template <class F>
int launch_fwd(const FwdArgs& a, hipStream_t s) {
  if (a.n == 0) return GSX_OK;
  const unsigned grid = static_cast<unsigned>((a.n + kFwdBlock - 1) / kFwdBlock);
  rb23_fwd_kernel<F><<<grid, kFwdBlock, 0, s>>>(a);
  return check_launch("rb23_fwd_kernel");
}

template <class F>
int launch_replay(const ReplayArgs& a, hipStream_t s) {
  if (a.n == 0) return GSX_OK;
  const unsigned grid = static_cast<unsigned>((a.n + kFwdBlock - 1) / kFwdBlock);
  rb23_replay_kernel<F><<<grid, kFwdBlock, 0, s>>>(a);
  return check_launch("rb23_replay_kernel");
}

template <class F, int CMAX>
int launch_adj_c(const AdjArgs& a, hipStream_t s) {
  const unsigned grid = static_cast<unsigned>((a.n + kAdjBlock - 1) / kAdjBlock);
  rb23_adj_kernel<F, CMAX><<<grid, kAdjBlock, 0, s>>>(a);
  return check_launch("rb23_adj_kernel");
}

template <class F>
int launch_adj(const AdjArgs& a, hipStream_t s) {
  if (a.n == 0) return GSX_OK;
  const int C = a.ckpt_interval;
  if (C <= 1) return launch_adj_c<F, 1>(a, s);
  if (C == 2) {
    const unsigned grid = static_cast<unsigned>((a.n + kFwdBlock - 1) / kFwdBlock);
    if (a.nsave > 0 || a.pbar_in)
      rb23_adj_pair_kernel<F, true><<<grid, kFwdBlock, 0, s>>>(a);
    else
      rb23_adj_pair_kernel<F, false><<<grid, kFwdBlock, 0, s>>>(a);
    return check_launch("rb23_adj_pair_kernel");
  }
  if (C <= 4) return launch_adj_c<F, 4>(a, s);
  if (C <= 8) return launch_adj_c<F, 8>(a, s);
  if (C <= 16) return launch_adj_c<F, 16>(a, s);
  if (C <= 32) return launch_adj_c<F, 32>(a, s);
  return set_error(GSX_E_INVALID, "ckpt_interval %d exceeds the adjoint's maximum of %d", C,
                   kMaxCkptInterval);
}

template <class F>
int launch_saveat(const SaveArgs& a, hipStream_t s) {
  const unsigned grid = static_cast<unsigned>((a.n + kFwdBlock - 1) / kFwdBlock);
  rb23_saveat_kernel<F><<<grid, kFwdBlock, 0, s>>>(a);
  return check_launch("rb23_saveat_kernel");
}

template <class F>
int launch_saveat_vjp(const SaveVjpArgs& a, hipStream_t s) {
  const unsigned grid = static_cast<unsigned>((a.n + kFwdBlock - 1) / kFwdBlock);
  rb23_saveat_vjp_kernel<F><<<grid, kFwdBlock, 0, s>>>(a);
  return check_launch("rb23_saveat_vjp_kernel");
}

template <class F>
int launch_fwd_f32(const FwdArgs& a, hipStream_t s) {
  if (a.n == 0) return GSX_OK;
  const unsigned grid = static_cast<unsigned>((a.n + kFwdBlock - 1) / kFwdBlock);
  rb23_fwd_f32_kernel<F><<<grid, kFwdBlock, 0, s>>>(a);
  return check_launch("rb23_fwd_f32_kernel");
}

template <class F>
int launch_replay_f32(const ReplayArgs& a, hipStream_t s) {
  if (a.n == 0) return GSX_OK;
  const unsigned grid = static_cast<unsigned>((a.n + kFwdBlock - 1) / kFwdBlock);
  rb23_replay_f32_kernel<F><<<grid, kFwdBlock, 0, s>>>(a);
  return check_launch("rb23_replay_f32_kernel");
}

template <class F, int CMAX>
int launch_adj_f32_c(const AdjArgs& a, hipStream_t s) {
  const unsigned grid = static_cast<unsigned>((a.n + kAdjBlock - 1) / kAdjBlock);
  rb23_adj_f32_kernel<F, CMAX><<<grid, kAdjBlock, 0, s>>>(a);
  return check_launch("rb23_adj_f32_kernel");
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

#if defined(GSX_STRICT_ARITH)
constexpr ::gsx::rt::Family kRtFamily = ::gsx::rt::kRb23Strict;
#else
constexpr ::gsx::rt::Family kRtFamily = ::gsx::rt::kRb23;
#endif

}  // namespace

#if !defined(GSX_STRICT_ARITH)
// The FP32 variant: autonomous stiff fields (built in, or compiled at run time).
int rb23_fwd_f32_dispatch(int field, int D, const FwdArgs& a, hipStream_t s) {
  switch (field) {
    case GSX_FIELD_ROBERTSON: return launch_fwd_f32<FRobertson>(a, s);
    case GSX_FIELD_LINEAR_DECAY:
      if (D == 1) return launch_fwd_f32<FLinearDecay<1>>(a, s);
      if (D == 2) return launch_fwd_f32<FLinearDecay<2>>(a, s);
      if (D == 3) return launch_fwd_f32<FLinearDecay<3>>(a, s);
      break;
    default:
      if (::gsx::rt::is_runtime(field))
        return ::gsx::rt::rb23_fwd(field, ::gsx::rt::kRb23F32, a, s);
      break;
  }
  return set_error(GSX_E_UNSUPPORTED, "rosenbrock23 FP32 forward: field %d with D=%d is not "
                   "built (the FP32 variant serves autonomous stiff fields)", field, D);
}

int rb23_replay_f32_dispatch(int field, int D, const ReplayArgs& a, hipStream_t s) {
  switch (field) {
    case GSX_FIELD_ROBERTSON: return launch_replay_f32<FRobertson>(a, s);
    case GSX_FIELD_LINEAR_DECAY:
      if (D == 1) return launch_replay_f32<FLinearDecay<1>>(a, s);
      if (D == 2) return launch_replay_f32<FLinearDecay<2>>(a, s);
      if (D == 3) return launch_replay_f32<FLinearDecay<3>>(a, s);
      break;
    default:
      if (::gsx::rt::is_runtime(field))
        return ::gsx::rt::rb23_replay(field, ::gsx::rt::kRb23F32, a, s);
      break;
  }
  return set_error(GSX_E_UNSUPPORTED, "rosenbrock23 FP32 replay: field %d with D=%d is not "
                   "built (the FP32 variant serves autonomous stiff fields)", field, D);
}

int rb23_adj_f32_dispatch(int field, int D, const AdjArgs& a, hipStream_t s) {
  switch (field) {
    case GSX_FIELD_ROBERTSON: return launch_adj_f32<FRobertson>(a, s);
    case GSX_FIELD_LINEAR_DECAY:
      if (D == 1) return launch_adj_f32<FLinearDecay<1>>(a, s);
      if (D == 2) return launch_adj_f32<FLinearDecay<2>>(a, s);
      if (D == 3) return launch_adj_f32<FLinearDecay<3>>(a, s);
      break;
    default:
      if (::gsx::rt::is_runtime(field))
        return ::gsx::rt::rb23_adj(field, ::gsx::rt::kRb23F32, a, s);
      break;
  }
  return set_error(GSX_E_UNSUPPORTED, "rosenbrock23 FP32 adjoint: field %d with D=%d is not "
                   "built (the FP32 variant serves autonomous stiff fields)", field, D);
}
#endif

int rb23_fwd_dispatch(int field, int D, const FwdArgs& a, hipStream_t s) {
  switch (field) {
    case GSX_FIELD_ROBERTSON: return launch_fwd<FRobertson>(a, s);
    case GSX_FIELD_FORCED_STIFF: return launch_fwd<FForcedStiff>(a, s);
    case GSX_FIELD_LINEAR_DECAY:
      if (D == 1) return launch_fwd<FLinearDecay<1>>(a, s);
      if (D == 2) return launch_fwd<FLinearDecay<2>>(a, s);
      if (D == 3) return launch_fwd<FLinearDecay<3>>(a, s);
      break;
    default:
      if (::gsx::rt::is_runtime(field)) return ::gsx::rt::rb23_fwd(field, kRtFamily, a, s);
      break;
  }
  return set_error(GSX_E_UNSUPPORTED, "rosenbrock23 forward: field %d with D=%d has no "
                   "analytic Jacobian in this build", field, D);
}

int rb23_replay_dispatch(int field, int D, const ReplayArgs& a, hipStream_t s) {
  switch (field) {
    case GSX_FIELD_ROBERTSON: return launch_replay<FRobertson>(a, s);
    case GSX_FIELD_FORCED_STIFF: return launch_replay<FForcedStiff>(a, s);
    case GSX_FIELD_LINEAR_DECAY:
      if (D == 1) return launch_replay<FLinearDecay<1>>(a, s);
      if (D == 2) return launch_replay<FLinearDecay<2>>(a, s);
      if (D == 3) return launch_replay<FLinearDecay<3>>(a, s);
      break;
    default:
      if (::gsx::rt::is_runtime(field)) return ::gsx::rt::rb23_replay(field, kRtFamily, a, s);
      break;
  }
  return set_error(GSX_E_UNSUPPORTED, "rosenbrock23 replay: field %d with D=%d has no "
                   "analytic Jacobian in this build", field, D);
}

int rb23_adj_dispatch(int field, int D, const AdjArgs& a, hipStream_t s) {
  switch (field) {
    case GSX_FIELD_ROBERTSON: return launch_adj<FRobertson>(a, s);
    case GSX_FIELD_FORCED_STIFF: return launch_adj<FForcedStiff>(a, s);
    case GSX_FIELD_LINEAR_DECAY:
      if (D == 1) return launch_adj<FLinearDecay<1>>(a, s);
      if (D == 2) return launch_adj<FLinearDecay<2>>(a, s);
      if (D == 3) return launch_adj<FLinearDecay<3>>(a, s);
      break;
    default:
      if (::gsx::rt::is_runtime(field)) return ::gsx::rt::rb23_adj(field, kRtFamily, a, s);
      break;
  }
  return set_error(GSX_E_UNSUPPORTED, "rosenbrock23 adjoint: field %d with D=%d has no "
                   "analytic Jacobian in this build", field, D);
}

int rb23_saveat_dispatch(int field, int D, const SaveArgs& a, hipStream_t s) {
  if (a.n == 0) return GSX_OK;
  switch (field) {
    case GSX_FIELD_ROBERTSON: return launch_saveat<FRobertson>(a, s);
    case GSX_FIELD_FORCED_STIFF: return launch_saveat<FForcedStiff>(a, s);
    case GSX_FIELD_LINEAR_DECAY:
      if (D == 1) return launch_saveat<FLinearDecay<1>>(a, s);
      if (D == 2) return launch_saveat<FLinearDecay<2>>(a, s);
      if (D == 3) return launch_saveat<FLinearDecay<3>>(a, s);
      break;
    default:
      if (::gsx::rt::is_runtime(field)) return ::gsx::rt::rb23_saveat(field, kRtFamily, a, s);
      break;
  }
  return set_error(GSX_E_UNSUPPORTED, "rosenbrock23 saveat: field %d with D=%d has no "
                   "analytic Jacobian in this build", field, D);
}

int rb23_saveat_vjp_dispatch(int field, int D, const SaveVjpArgs& a, hipStream_t s) {
  if (a.n == 0) return GSX_OK;
  switch (field) {
    case GSX_FIELD_ROBERTSON: return launch_saveat_vjp<FRobertson>(a, s);
    case GSX_FIELD_FORCED_STIFF: return launch_saveat_vjp<FForcedStiff>(a, s);
    case GSX_FIELD_LINEAR_DECAY:
      if (D == 1) return launch_saveat_vjp<FLinearDecay<1>>(a, s);
      if (D == 2) return launch_saveat_vjp<FLinearDecay<2>>(a, s);
      if (D == 3) return launch_saveat_vjp<FLinearDecay<3>>(a, s);
      break;
    default:
      if (::gsx::rt::is_runtime(field))
        return ::gsx::rt::rb23_saveat_vjp(field, kRtFamily, a, s);
      break;
  }
  return set_error(GSX_E_UNSUPPORTED, "rosenbrock23 saveat adjoint: field %d with D=%d has no "
                   "analytic Jacobian in this build", field, D);
}

}  // namespace GSX_NS
