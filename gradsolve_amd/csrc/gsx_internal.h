// gsx_internal.h -- declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/gradsolve_amd.h"
#include "gsx_core.h"
#include "gsx_persist.h"

namespace gsx {

constexpr int kMaxCkptInterval = 32;

// Records a printf-style message as the calling thread's last error; returns `code`.
int set_error(int code, const char* fmt, ...);
// Checks hipGetLastError after a launch; returns GSX_OK or GSX_E_HIP.
int check_launch(const char* what);

struct FieldInfo {
  int D, P;
  bool autonomous;
  bool stiff;
};
// Static facts about a built-in field at dimension D; false if (field, D) is not built.
bool field_info(int field, int D, FieldInfo* out);

int tsit5_fwd_dispatch(int field, int D, const FwdArgs& a, hipStream_t s);
int tsit5_replay_dispatch(int field, int D, const ReplayArgs& a, hipStream_t s);
int tsit5_adj_dispatch(int field, int D, const AdjArgs& a, hipStream_t s);
int tsit5_saveat_dispatch(int field, int D, const SaveArgs& a, hipStream_t s);
int tsit5_saveat_vjp_dispatch(int field, int D, const SaveVjpArgs& a, hipStream_t s);
int rb23_fwd_dispatch(int field, int D, const FwdArgs& a, hipStream_t s);
int rb23_replay_dispatch(int field, int D, const ReplayArgs& a, hipStream_t s);
int rb23_adj_dispatch(int field, int D, const AdjArgs& a, hipStream_t s);
int rb23_saveat_dispatch(int field, int D, const SaveArgs& a, hipStream_t s);
int rb23_saveat_vjp_dispatch(int field, int D, const SaveVjpArgs& a, hipStream_t s);
int r5_fwd_dispatch(int field, int D, const FwdArgs& a, hipStream_t s);
int r5_replay_dispatch(int field, int D, const ReplayArgs& a, hipStream_t s);
int r5_adj_dispatch(int field, int D, const AdjArgs& a, hipStream_t s);
int r5_saveat_dispatch(int field, int D, const SaveArgs& a, bool dense, hipStream_t s);
int r5_saveat_vjp_dispatch(int field, int D, const SaveVjpArgs& a, hipStream_t s);
int tsit5_fwd_f32_dispatch(int field, int D, const FwdArgs& a, hipStream_t s);
int tsit5_replay_f32_dispatch(int field, int D, const ReplayArgs& a, hipStream_t s);
int tsit5_adj_f32_dispatch(int field, int D, const AdjArgs& a, hipStream_t s);
int rb23_fwd_f32_dispatch(int field, int D, const FwdArgs& a, hipStream_t s);
int rb23_replay_f32_dispatch(int field, int D, const ReplayArgs& a, hipStream_t s);
int rb23_adj_f32_dispatch(int field, int D, const AdjArgs& a, hipStream_t s);
int persist_fwd_dispatch(int field, int D, bool fp32, const PersistFwdArgs& pa, hipStream_t s);
int persist_adj_dispatch(int field, int D, bool fp32, const PersistAdjArgs& pa, hipStream_t s);
int kura_fwd_launch(const KuraFwdArgs& a, bool fp32, hipStream_t s);
int kura_replay_launch(const KuraReplayArgs& a, bool fp32, hipStream_t s);
int kura_saveat_launch(const KuraSaveatArgs& a, bool fp32, hipStream_t s);
int kura_adj_launch(const KuraAdjArgs& a, bool fp32, hipStream_t s);
int plain_dispatch(int field, bool adjoint, const PlainArgs& a, hipStream_t s);

// Fields registered at run time (gsx_runtime.hip): ids from GSX_FIELD_RUNTIME_BASE up,
// kernels compiled with hiprtc per family and launched with the built-in argument blocks.
namespace rt {
enum Family { kTsit5, kTsit5Strict, kTsit5F32, kRb23, kRb23Strict, kRodas5p, kRodas5pStrict,
              kRb23F32, kTsit5Wide, kNumFamilies };
bool is_runtime(int field);
bool info(int field, FieldInfo* out);
int tsit5_fwd(int field, Family fam, const FwdArgs& a, hipStream_t s);
int tsit5_replay(int field, Family fam, const ReplayArgs& a, hipStream_t s);
int tsit5_adj(int field, Family fam, const AdjArgs& a, hipStream_t s);
int tsit5_saveat(int field, Family fam, const SaveArgs& a, hipStream_t s);
int tsit5_saveat_vjp(int field, Family fam, const SaveVjpArgs& a, hipStream_t s);
int rb23_fwd(int field, Family fam, const FwdArgs& a, hipStream_t s);
int rb23_replay(int field, Family fam, const ReplayArgs& a, hipStream_t s);
int rb23_adj(int field, Family fam, const AdjArgs& a, hipStream_t s);
int rb23_saveat(int field, Family fam, const SaveArgs& a, hipStream_t s);
int r5_launch(int field, Family fam, const char* key, int64_t n, const FwdArgs& a,
              hipStream_t s);
int r5_launch(int field, Family fam, const char* key, int64_t n, const ReplayArgs& a,
              hipStream_t s);
int r5_launch(int field, Family fam, const char* key, int64_t n, const SaveArgs& a,
              hipStream_t s);
int r5_launch(int field, Family fam, const char* key, int64_t n, const SaveVjpArgs& a,
              hipStream_t s);
int r5_adj(int field, Family fam, const AdjArgs& a, hipStream_t s);
int rb23_saveat_vjp(int field, Family fam, const SaveVjpArgs& a, hipStream_t s);
int tsit5w_fwd(int field, const WideFwdArgs& w, hipStream_t s);
int tsit5w_replay(int field, const WideReplayArgs& w, hipStream_t s);
int tsit5w_adj(int field, const WideAdjArgs& w, hipStream_t s);
// Widest state a registered field may have (the wide Tsit5 lane), the widest the register
// lanes serve, and the widest stiff field (each lane holds a D x D W and its LU).
constexpr int kMaxWideD = 4096;
constexpr int kMaxRegisterD = 64;
constexpr int kMaxStiffD = 12;
}  // namespace rt

}  // namespace gsx

namespace gsx_strict {
int tsit5_fwd_dispatch(int field, int D, const ::gsx::FwdArgs& a, hipStream_t s);
int tsit5_replay_dispatch(int field, int D, const ::gsx::ReplayArgs& a, hipStream_t s);
int tsit5_adj_dispatch(int field, int D, const ::gsx::AdjArgs& a, hipStream_t s);
int tsit5_saveat_dispatch(int field, int D, const ::gsx::SaveArgs& a, hipStream_t s);
int tsit5_saveat_vjp_dispatch(int field, int D, const ::gsx::SaveVjpArgs& a, hipStream_t s);
int rb23_fwd_dispatch(int field, int D, const ::gsx::FwdArgs& a, hipStream_t s);
int rb23_replay_dispatch(int field, int D, const ::gsx::ReplayArgs& a, hipStream_t s);
int rb23_adj_dispatch(int field, int D, const ::gsx::AdjArgs& a, hipStream_t s);
int rb23_saveat_dispatch(int field, int D, const ::gsx::SaveArgs& a, hipStream_t s);
int rb23_saveat_vjp_dispatch(int field, int D, const ::gsx::SaveVjpArgs& a, hipStream_t s);
int r5_fwd_dispatch(int field, int D, const ::gsx::FwdArgs& a, hipStream_t s);
int r5_replay_dispatch(int field, int D, const ::gsx::ReplayArgs& a, hipStream_t s);
int r5_adj_dispatch(int field, int D, const ::gsx::AdjArgs& a, hipStream_t s);
int r5_saveat_dispatch(int field, int D, const ::gsx::SaveArgs& a, bool dense, hipStream_t s);
int r5_saveat_vjp_dispatch(int field, int D, const ::gsx::SaveVjpArgs& a, hipStream_t s);
int tsit5_fwd_f32_dispatch(int field, int D, const ::gsx::FwdArgs& a, hipStream_t s);
int tsit5_replay_f32_dispatch(int field, int D, const ::gsx::ReplayArgs& a, hipStream_t s);
int tsit5_adj_f32_dispatch(int field, int D, const ::gsx::AdjArgs& a, hipStream_t s);
}  // namespace gsx_strict
