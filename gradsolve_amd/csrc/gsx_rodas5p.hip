// gsx_rodas5p.hip -- Rodas5P forward+record, replay, adjoint and dense-output kernels for gfx950
// (one trajectory per lane; per-lane bodies in gsx_rodas5p.h). Compiled twice like
// gsx_rb23.hip: namespace gsx (contraction on) and -DGSX_NS=gsx_strict -ffp-contract=off.
#include <hip/hip_runtime.h>

#include "gsx_internal.h"
#include "gsx_rodas5p_kernels.h"

namespace GSX_NS {

using ::gsx::check_launch;
using ::gsx::kMaxCkptInterval;
using ::gsx::set_error;

namespace {

unsigned grid_of(int64_t n, int block) { return static_cast<unsigned>((n + block - 1) / block); }

template <class F>
int launch_fwd(const FwdArgs& a, hipStream_t s) {
  if (a.n == 0) return GSX_OK;
  r5_fwd_kernel<F><<<grid_of(a.n, kFwdBlock), kFwdBlock, 0, s>>>(a);
  return check_launch("r5_fwd_kernel");
}

template <class F>
int launch_replay(const ReplayArgs& a, hipStream_t s) {
  if (a.n == 0) return GSX_OK;
  r5_replay_kernel<F><<<grid_of(a.n, kFwdBlock), kFwdBlock, 0, s>>>(a);
  return check_launch("r5_replay_kernel");
}

template <class F, int CMAX>
int launch_adj_c(const AdjArgs& a, hipStream_t s) {
  r5_adj_kernel<F, CMAX><<<grid_of(a.n, kAdjBlock), kAdjBlock, 0, s>>>(a);
  return check_launch("r5_adj_kernel");
}

template <class F>
int launch_adj(const AdjArgs& a, hipStream_t s) {
  if (a.n == 0) return GSX_OK;
  const int C = a.ckpt_interval;
  if (C <= 1) return launch_adj_c<F, 1>(a, s);
  if (C <= 4) return launch_adj_c<F, 4>(a, s);
  if (C <= 8) return launch_adj_c<F, 8>(a, s);
  if (C <= 16) return launch_adj_c<F, 16>(a, s);
  if (C <= 32) return launch_adj_c<F, 32>(a, s);
  return set_error(GSX_E_INVALID, "ckpt_interval %d exceeds the adjoint's maximum of %d", C,
                   kMaxCkptInterval);
}

template <class F>
int launch_saveat(const SaveArgs& a, bool dense, hipStream_t s) {
  if (a.n == 0) return GSX_OK;
  if (dense) {
    r5_saveat_dense_kernel<F><<<grid_of(a.n, kFwdBlock), kFwdBlock, 0, s>>>(a);
    return check_launch("r5_saveat_dense_kernel");
  }
  r5_saveat_kernel<F><<<grid_of(a.n, kFwdBlock), kFwdBlock, 0, s>>>(a);
  return check_launch("r5_saveat_kernel");
}

template <class F>
int launch_saveat_vjp(const SaveVjpArgs& a, hipStream_t s) {
  if (a.n == 0) return GSX_OK;
  r5_saveat_vjp_kernel<F><<<grid_of(a.n, kFwdBlock), kFwdBlock, 0, s>>>(a);
  return check_launch("r5_saveat_vjp_kernel");
}

#if defined(GSX_STRICT_ARITH)
constexpr ::gsx::rt::Family kRtOwn = ::gsx::rt::kRodas5pStrict;
#else
constexpr ::gsx::rt::Family kRtOwn = ::gsx::rt::kRodas5p;
#endif

// The built-in fields with an analytic Jacobian; run-time fields go to the hiprtc kernels.
#define GSX_R5_FIELDS(CALL, ...)                                                  \
  switch (field) {                                                                \
    case GSX_FIELD_ROBERTSON: return CALL<FRobertson>(__VA_ARGS__);               \
    case GSX_FIELD_FORCED_STIFF: return CALL<FForcedStiff>(__VA_ARGS__);          \
    case GSX_FIELD_LINEAR_DECAY:                                                  \
      if (D == 1) return CALL<FLinearDecay<1>>(__VA_ARGS__);                      \
      if (D == 2) return CALL<FLinearDecay<2>>(__VA_ARGS__);                      \
      if (D == 3) return CALL<FLinearDecay<3>>(__VA_ARGS__);                      \
      break;                                                                      \
    default: break;                                                               \
  }

int unsupported(const char* what, int field, int D) {
  return set_error(GSX_E_UNSUPPORTED, "rodas5p %s: field %d with D=%d has no analytic Jacobian "
                   "in this build", what, field, D);
}

}  // namespace

int r5_fwd_dispatch(int field, int D, const FwdArgs& a, hipStream_t s) {
  GSX_R5_FIELDS(launch_fwd, a, s)
  if (::gsx::rt::is_runtime(field)) return ::gsx::rt::r5_launch(field, kRtOwn, "fwd", a.n, a, s);
  return unsupported("forward", field, D);
}

int r5_replay_dispatch(int field, int D, const ReplayArgs& a, hipStream_t s) {
  GSX_R5_FIELDS(launch_replay, a, s)
  if (::gsx::rt::is_runtime(field))
    return ::gsx::rt::r5_launch(field, kRtOwn, "replay", a.n, a, s);
  return unsupported("replay", field, D);
}

int r5_adj_dispatch(int field, int D, const AdjArgs& a, hipStream_t s) {
  GSX_R5_FIELDS(launch_adj, a, s)
  if (::gsx::rt::is_runtime(field)) return ::gsx::rt::r5_adj(field, kRtOwn, a, s);
  return unsupported("adjoint", field, D);
}

int r5_saveat_dispatch(int field, int D, const SaveArgs& a, bool dense, hipStream_t s) {
  GSX_R5_FIELDS(launch_saveat, a, dense, s)
  if (::gsx::rt::is_runtime(field))
    return ::gsx::rt::r5_launch(field, kRtOwn, dense ? "saveat_dense" : "saveat", a.n, a, s);
  return unsupported("saveat", field, D);
}

int r5_saveat_vjp_dispatch(int field, int D, const SaveVjpArgs& a, hipStream_t s) {
  GSX_R5_FIELDS(launch_saveat_vjp, a, s)
  if (::gsx::rt::is_runtime(field))
    return ::gsx::rt::r5_launch(field, kRtOwn, "saveat_vjp", a.n, a, s);
  return unsupported("saveat adjoint", field, D);
}

#undef GSX_R5_FIELDS

}  // namespace GSX_NS
