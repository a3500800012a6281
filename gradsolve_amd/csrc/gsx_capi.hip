// gsx_capi.hip -- the extern "C" boundary (include/gradsolve_amd.h): argument validation,
// error reporting, and the small layout-conversion kernels for tape interchange.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "gsx_internal.h"

// Scratch of gsx_tsit5_adjoint_saveat_noalloc: the re-steps' cotangents mu [nsave][D][n],
// the one buffer the saveat gradient needs beyond its inputs and outputs. 2^23 doubles
// (64 MiB), held by every device the library runs on until the process ends; every call on a
// device shares that device's copy. In host code `scratch` is only a stand-in for the device
// array: the device address comes from hipGetSymbolAddress (gsx::noalloc_scratch).
constexpr int64_t kNoallocScratchDoubles = int64_t(1) << 23;
static __device__ __align__(256) double scratch[kNoallocScratchDoubles];

namespace gsx {

namespace {
thread_local char g_err[8192] = "";  // room for a field compiler log
}

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(GSX_E_HIP, "%s: %s", what, hipGetErrorString(e));
  return GSX_OK;
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// The current device's copy of `scratch`: looked up once per device, with that device
// current, and kept (each device holds its own copy at its own address).
int noalloc_scratch(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_error(GSX_E_HIP, "hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return set_error(GSX_E_UNSUPPORTED, "device %d: the scratch table holds %d devices", dev,
                     kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return set_error(GSX_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                       hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return GSX_OK;
}

bool field_info(int field, int D, FieldInfo* out) {
  if (rt::is_runtime(field)) return rt::info(field, out) && out->D == D;
  switch (field) {
    case GSX_FIELD_LORENZ: *out = {3, 1, true, false}; return D == 3;
    case GSX_FIELD_LINEAR_DECAY: *out = {D, 1, true, false}; return D >= 1 && D <= 3;
    case GSX_FIELD_ROBERTSON: *out = {3, 3, true, true}; return D == 3;
    case GSX_FIELD_KURAMOTO: *out = {32, 32, true, false}; return D == 32;
    case GSX_FIELD_VDP: *out = {2, 1, true, false}; return D == 2;
    case GSX_FIELD_LORENZ3: *out = {3, 3, true, false}; return D == 3;
    case GSX_FIELD_FORCED_STIFF: *out = {2, 3, false, true}; return D == 2;
    default: return false;
  }
}

namespace {

__global__ void tape_to_padded_kernel(int64_t n, const double* __restrict__ tape,
                                      const int32_t* __restrict__ n_acc, int64_t S,
                                      double* __restrict__ out) {
  // One thread per (s, j); j fastest so that reads of the step-major tape coalesce.
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= n * S) return;
  const int64_t s = idx / n;
  const int64_t j = idx - s * n;
  out[j * S + s] = (s < n_acc[j]) ? tape[s * n + j] : 0.0;
}

__global__ void tape_from_padded_kernel(int64_t n, const double* __restrict__ in, int64_t S,
                                        double* __restrict__ tape, int64_t cap,
                                        int32_t* __restrict__ n_acc) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  int32_t last = 0;
  for (int64_t s = 0; s < S; ++s) {
    const double h = in[j * S + s];
    tape[s * n + j] = h;
    if (h != 0.0) last = static_cast<int32_t>(s + 1);
  }
  for (int64_t s = S; s < cap; ++s) tape[s * n + j] = 0.0;
  n_acc[j] = last;
}

// Rows n_acc[j] <= s < S of column j become 0.0, so that the leading S rows of the step-major
// tape are the reference's trimmed [S, n] record (a zero step is the replay's identity).
__global__ void tape_zero_tail_kernel(int64_t n, double* __restrict__ tape,
                                      const int32_t* __restrict__ n_acc, int64_t S) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= n * S) return;
  const int64_t s = idx / n;
  const int64_t j = idx - s * n;
  if (s >= n_acc[j]) tape[idx] = 0.0;
}

// n_acc[j] = 1 + the last row s < S with tape[s][j] != 0 (accepted steps are positive).
__global__ void tape_count_kernel(int64_t n, const double* __restrict__ tape, int64_t S,
                                  int32_t* __restrict__ n_acc) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  int32_t last = 0;
  for (int64_t s = 0; s < S; ++s)
    if (tape[s * n + j] != 0.0) last = static_cast<int32_t>(s + 1);
  n_acc[j] = last;
}

int check_field(int field, int D, int P, FieldInfo* fi) {
  if (!field_info(field, D, fi))
    return set_error(GSX_E_INVALID, "field %d is not built for D=%d", field, D);
  if (P != fi->P)
    return set_error(GSX_E_INVALID, "field %d expects P=%d parameters, got %d", field, fi->P, P);
  return GSX_OK;
}

}  // namespace
}  // namespace gsx

using namespace gsx;

extern "C" {

int gsx_version(void) { return 100; }

const char* gsx_last_error(void) { return g_err; }

int gsx_tsit5_tableau(double* c, double* a, double* b, double* e) {
  if (!c || !a || !b || !e) return set_error(GSX_E_INVALID, "tableau: null output pointer");
  using namespace tab;
  const double cc[5] = {c2, c3, c4, c5, c6};
  const double aa[15] = {a21, a31, a32, a41, a42, a43, a51, a52,
                         a53, a54, a61, a62, a63, a64, a65};
  const double bb[6] = {b1, b2, b3, b4, b5, b6};
  const double ee[7] = {e1, e2, e3, e4, e5, e6, e7};
  std::memcpy(c, cc, sizeof cc);
  std::memcpy(a, aa, sizeof aa);
  std::memcpy(b, bb, sizeof bb);
  std::memcpy(e, ee, sizeof ee);
  return GSX_OK;
}

int gsx_ckpt_rows(int field, int D) {
  FieldInfo fi;
  if (!field_info(field, D, &fi))
    return set_error(GSX_E_INVALID, "field %d is not built for D=%d", field, D);
  return fi.D + (fi.autonomous ? 0 : 1);
}

int64_t gsx_ckpt_slots(int64_t tape_cap, int32_t ckpt_interval) {
  if (tape_cap < 0 || ckpt_interval < 1)
    return set_error(GSX_E_INVALID, "ckpt_slots: tape_cap %lld, interval %d",
                     static_cast<long long>(tape_cap), ckpt_interval);
  return tape_cap / ckpt_interval + 1;
}

int gsx_tsit5_fwd(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                  const double* p, double t0, double t1, const double* t1_per, double rtol,
                  double atol, int64_t max_steps, double* y_final, int32_t* n_acc,
                  int32_t* n_rej, int32_t* status, double* tape, int64_t tape_cap,
                  double* ckpt, int32_t ckpt_interval, int flags, void* stream) {
  FieldInfo fi;
  int rc = check_field(field, D, P, &fi);
  if (rc) return rc;
  if (fi.stiff) return set_error(GSX_E_INVALID, "field %d is stiff; use Rosenbrock23", field);
  if (n < 0 || max_steps < 0) return set_error(GSX_E_INVALID, "negative n or max_steps");
  // n_acc, n_rej and status may be null for a forward without a tape (the reference's
  // forward handler returns y_final and, optionally, the accepted-step count); a recording
  // forward needs n_acc and status (the tape's length and its "tape full" code).
  if (n > 0 && (!y0 || !p || !y_final))
    return set_error(GSX_E_INVALID, "tsit5_fwd: null input/output pointer");
  if (n > 0 && tape && (!n_acc || !status))
    return set_error(GSX_E_INVALID, "tsit5_fwd: a recording forward needs n_acc and status");
  if (tape && tape_cap < 0) return set_error(GSX_E_INVALID, "negative tape_cap");
  if (ckpt && (ckpt_interval < 1 || ckpt_interval > kMaxCkptInterval))
    return set_error(GSX_E_INVALID, "ckpt_interval %d outside [1, %d]", ckpt_interval,
                     kMaxCkptInterval);
  if (ckpt && !tape) return set_error(GSX_E_INVALID, "checkpoints need a tape");
  if (!(rtol > 0.0) || !(atol >= 0.0)) return set_error(GSX_E_INVALID, "bad rtol/atol");
  FwdArgs a{n, y0, p, aux, t0, t1, t1_per, rtol, atol, max_steps, y_final, n_acc, n_rej,
            status, tape, tape ? tape_cap : 0, ckpt, ckpt ? ckpt_interval : 1};
  if (flags & ~(GSX_STRICT | GSX_FP32))
    return set_error(GSX_E_INVALID, "unknown flags 0x%x", flags);
  if ((flags & GSX_STRICT) && (flags & GSX_FP32))
    return set_error(GSX_E_INVALID, "GSX_STRICT and GSX_FP32 are exclusive");
  if (flags & GSX_STRICT)
    return gsx_strict::tsit5_fwd_dispatch(field, D, a, static_cast<hipStream_t>(stream));
  if (flags & GSX_FP32)
    return tsit5_fwd_f32_dispatch(field, D, a, static_cast<hipStream_t>(stream));
  return tsit5_fwd_dispatch(field, D, a, static_cast<hipStream_t>(stream));
}

int gsx_tsit5_replay(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                     const double* p, double t0, const double* tape, const int32_t* n_acc,
                     double* y_final, double* ckpt, int32_t ckpt_interval, int flags,
                     void* stream) {
  FieldInfo fi;
  int rc = check_field(field, D, P, &fi);
  if (rc) return rc;
  if (fi.stiff) return set_error(GSX_E_INVALID, "field %d is stiff; use Rosenbrock23", field);
  if (n < 0) return set_error(GSX_E_INVALID, "negative n");
  if (n > 0 && (!y0 || !p || !tape || !n_acc || !y_final))
    return set_error(GSX_E_INVALID, "tsit5_replay: null pointer");
  if (ckpt && (ckpt_interval < 1 || ckpt_interval > kMaxCkptInterval))
    return set_error(GSX_E_INVALID, "ckpt_interval %d outside [1, %d]", ckpt_interval,
                     kMaxCkptInterval);
  ReplayArgs a{n, y0, p, aux, t0, tape, n_acc, y_final, ckpt, ckpt ? ckpt_interval : 1};
  if (flags & ~(GSX_STRICT | GSX_FP32))
    return set_error(GSX_E_INVALID, "unknown flags 0x%x", flags);
  if ((flags & GSX_STRICT) && (flags & GSX_FP32))
    return set_error(GSX_E_INVALID, "GSX_STRICT and GSX_FP32 are exclusive");
  if (flags & GSX_STRICT)
    return gsx_strict::tsit5_replay_dispatch(field, D, a, static_cast<hipStream_t>(stream));
  if (flags & GSX_FP32)
    return tsit5_replay_f32_dispatch(field, D, a, static_cast<hipStream_t>(stream));
  return tsit5_replay_dispatch(field, D, a, static_cast<hipStream_t>(stream));
}

int gsx_tsit5_adjoint(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                      const double* p, double t0, const double* tape, int64_t tape_cap,
                      const int32_t* n_acc, const double* ckpt, int32_t ckpt_interval,
                      const double* ybar, double* y0bar, double* pbar, int flags,
                      void* stream) {
  FieldInfo fi;
  int rc = check_field(field, D, P, &fi);
  if (rc) return rc;
  if (fi.stiff) return set_error(GSX_E_INVALID, "field %d is stiff; use Rosenbrock23", field);
  if (n < 0) return set_error(GSX_E_INVALID, "negative n");
  if (n > 0 && (!p || !tape || !n_acc || !ckpt || !ybar || !y0bar || !pbar))
    return set_error(GSX_E_INVALID, "tsit5_adjoint: null pointer");
  if (ckpt_interval < 1 || ckpt_interval > kMaxCkptInterval)
    return set_error(GSX_E_INVALID, "ckpt_interval %d outside [1, %d]", ckpt_interval,
                     kMaxCkptInterval);
  AdjArgs a{n, y0, p, aux, t0, tape, tape_cap, n_acc, ckpt, ckpt_interval, ybar, y0bar, pbar};
  if (flags & ~(GSX_STRICT | GSX_FP32))
    return set_error(GSX_E_INVALID, "unknown flags 0x%x", flags);
  if ((flags & GSX_STRICT) && (flags & GSX_FP32))
    return set_error(GSX_E_INVALID, "GSX_STRICT and GSX_FP32 are exclusive");
  if (flags & GSX_STRICT)
    return gsx_strict::tsit5_adj_dispatch(field, D, a, static_cast<hipStream_t>(stream));
  if (flags & GSX_FP32)
    return tsit5_adj_f32_dispatch(field, D, a, static_cast<hipStream_t>(stream));
  return tsit5_adj_dispatch(field, D, a, static_cast<hipStream_t>(stream));
}

// ---- The wide Tsit5 lane (fields compiled at run time, any D up to 4096) ----------------
namespace {
int wide_check(const char* what, int field, int64_t n, int D, int P, int32_t C, bool need_c,
               const double* ws) {
  FieldInfo fi;
  int rc = check_field(field, D, P, &fi);
  if (rc) return rc;
  if (!rt::is_runtime(field))
    return set_error(GSX_E_UNSUPPORTED, "%s: the wide lane serves fields compiled at run time "
                     "(gsx_field_register); field %d is built in", what, field);
  if (fi.stiff) return set_error(GSX_E_INVALID, "field %d is stiff; use Rosenbrock23", field);
  if (n < 0) return set_error(GSX_E_INVALID, "%s: negative n", what);
  if (need_c && (C < 1 || C > kMaxCkptInterval))
    return set_error(GSX_E_INVALID, "ckpt_interval %d outside [1, %d]", C, kMaxCkptInterval);
  if (n > 0 && !ws) return set_error(GSX_E_INVALID, "%s: null workspace", what);
  return GSX_OK;
}
}  // namespace

int64_t gsx_tsit5_wide_ws_rows(int which, int32_t ckpt_interval) {
  if (which == 0) return kWideFwdRows;
  if (which == 1 && ckpt_interval >= 1 && ckpt_interval <= kMaxCkptInterval)
    return wide_adj_rows(ckpt_interval);
  return set_error(GSX_E_INVALID, "tsit5_wide_ws_rows: which %d, ckpt_interval %d", which,
                   ckpt_interval);
}

int gsx_tsit5_wide_fwd(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                       const double* p, double t0, double t1, const double* t1_per,
                       double rtol, double atol, int64_t max_steps, double* y_final,
                       int32_t* n_acc, int32_t* n_rej, int32_t* status, double* tape,
                       int64_t tape_cap, double* ckpt, int32_t ckpt_interval, double* ws,
                       void* stream) {
  int rc = wide_check("tsit5_wide_fwd", field, n, D, P, ckpt_interval, ckpt != nullptr, ws);
  if (rc) return rc;
  if (max_steps < 0) return set_error(GSX_E_INVALID, "negative max_steps");
  if (n > 0 && (!y0 || !p || !y_final))
    return set_error(GSX_E_INVALID, "tsit5_wide_fwd: null input/output pointer");
  if (n > 0 && tape && (!n_acc || !status))
    return set_error(GSX_E_INVALID, "tsit5_wide_fwd: a recording forward needs n_acc and status");
  if (tape && tape_cap < 0) return set_error(GSX_E_INVALID, "negative tape_cap");
  if (ckpt && !tape) return set_error(GSX_E_INVALID, "checkpoints need a tape");
  if (!(rtol > 0.0) || !(atol >= 0.0)) return set_error(GSX_E_INVALID, "bad rtol/atol");
  WideFwdArgs w{{n, y0, p, aux, t0, t1, t1_per, rtol, atol, max_steps, y_final, n_acc, n_rej,
                 status, tape, tape ? tape_cap : 0, ckpt, ckpt ? ckpt_interval : 1}, ws};
  return rt::tsit5w_fwd(field, w, static_cast<hipStream_t>(stream));
}

int gsx_tsit5_wide_replay(int field, int64_t n, int D, int P, const double* aux,
                          const double* y0, const double* p, double t0, const double* tape,
                          const int32_t* n_acc, double* y_final, double* ckpt,
                          int32_t ckpt_interval, double* ws, void* stream) {
  int rc = wide_check("tsit5_wide_replay", field, n, D, P, ckpt_interval, ckpt != nullptr, ws);
  if (rc) return rc;
  if (n > 0 && (!y0 || !p || !tape || !n_acc || !y_final))
    return set_error(GSX_E_INVALID, "tsit5_wide_replay: null pointer");
  WideReplayArgs w{{n, y0, p, aux, t0, tape, n_acc, y_final, ckpt, ckpt ? ckpt_interval : 1},
                   ws};
  return rt::tsit5w_replay(field, w, static_cast<hipStream_t>(stream));
}

int gsx_tsit5_wide_adjoint(int field, int64_t n, int D, int P, const double* aux,
                           const double* y0, const double* p, double t0, const double* tape,
                           int64_t tape_cap, const int32_t* n_acc, const double* ckpt,
                           int32_t ckpt_interval, const double* ybar, double* y0bar,
                           double* pbar, double* ws, void* stream) {
  int rc = wide_check("tsit5_wide_adjoint", field, n, D, P, ckpt_interval, true, ws);
  if (rc) return rc;
  if (n > 0 && (!p || !tape || !n_acc || !ckpt || !ybar || !y0bar || !pbar))
    return set_error(GSX_E_INVALID, "tsit5_wide_adjoint: null pointer");
  WideAdjArgs w{{n, y0, p, aux, t0, tape, tape_cap, n_acc, ckpt, ckpt_interval, ybar, y0bar,
                 pbar},
                ws};
  return rt::tsit5w_adj(field, w, static_cast<hipStream_t>(stream));
}

namespace {

// Dense output and its gradient for one of the two one-step methods; the Tsit5 and
// Rosenbrock23 entry points below differ only in the method and the fields they accept.
enum class DenseMethod { kTsit5, kRb23, kRodas5p };

int saveat_common(DenseMethod m, const char* what, int field, int64_t n, int D, int P,
                  const double* aux, const double* y0, const double* p, double t0,
                  const double* tape, const int32_t* n_acc, const double* y_final,
                  const double* ts, int64_t nsave, double* ys, int32_t* save_step,
                  double* save_t, double* states, int flags, void* stream) {
  FieldInfo fi;
  int rc = check_field(field, D, P, &fi);
  if (rc) return rc;
  if (m == DenseMethod::kTsit5 && fi.stiff)
    return set_error(GSX_E_INVALID, "field %d is stiff; use Rosenbrock23", field);
  if (n < 0 || nsave < 0) return set_error(GSX_E_INVALID, "negative n or nsave");
  if (flags & ~GSX_STRICT)
    return set_error(flags & GSX_FP32 ? GSX_E_UNSUPPORTED : GSX_E_INVALID,
                     "%s: flags 0x%x (0 or GSX_STRICT; no FP32 dense output)", what, flags);
  if (n == 0 || nsave == 0) return GSX_OK;
  if (!y0 || !p || !tape || !n_acc || !ts || !ys || !save_step || !save_t)
    return set_error(GSX_E_INVALID, "%s: null pointer", what);
  SaveArgs a{n, y0, p, aux, t0, tape, n_acc, ts, nsave, ys, save_step, save_t, states, y_final};
  hipStream_t st = static_cast<hipStream_t>(stream);
  const bool strict = (flags & GSX_STRICT) != 0;
  if (m == DenseMethod::kRb23)
    return strict ? gsx_strict::rb23_saveat_dispatch(field, D, a, st)
                  : rb23_saveat_dispatch(field, D, a, st);
  if (m == DenseMethod::kRodas5p)
    return strict ? gsx_strict::r5_saveat_dispatch(field, D, a, false, st)
                  : r5_saveat_dispatch(field, D, a, false, st);
  return strict ? gsx_strict::tsit5_saveat_dispatch(field, D, a, st)
                : tsit5_saveat_dispatch(field, D, a, st);
}

int adjoint_saveat_common(DenseMethod m, const char* what, int field, int64_t n, int D, int P,
                          const double* aux, const double* y0, const double* p, double t0,
                          const double* tape, int64_t tape_cap, const int32_t* n_acc,
                          const double* ckpt, int32_t ckpt_interval, const double* ts,
                          int64_t nsave, const int32_t* save_step, const double* save_t,
                          const double* states, const double* ysbar, const double* ybar,
                          double* mu, double* y0bar, double* pbar, int flags, void* stream) {
  FieldInfo fi;
  int rc = check_field(field, D, P, &fi);
  if (rc) return rc;
  if (m == DenseMethod::kTsit5 && fi.stiff)
    return set_error(GSX_E_INVALID, "field %d is stiff; use Rosenbrock23", field);
  if (n < 0 || nsave < 0) return set_error(GSX_E_INVALID, "negative n or nsave");
  if (flags & ~GSX_STRICT)
    return set_error(flags & GSX_FP32 ? GSX_E_UNSUPPORTED : GSX_E_INVALID,
                     "%s: flags 0x%x (0 or GSX_STRICT)", what, flags);
  if (ckpt_interval < 1 || ckpt_interval > kMaxCkptInterval)
    return set_error(GSX_E_INVALID, "ckpt_interval %d outside [1, %d]", ckpt_interval,
                     kMaxCkptInterval);
  if (n == 0) return GSX_OK;
  if (!p || !tape || !n_acc || !ckpt || !y0bar || !pbar ||
      (nsave > 0 && (!ts || !save_step || !save_t || !states || !ysbar || !mu)))
    return set_error(GSX_E_INVALID, "%s: null pointer", what);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const bool strict = (flags & GSX_STRICT) != 0;
  const bool rb = m == DenseMethod::kRb23;
  // Phase 1: the re-steps' VJPs -> mu, the final-state cotangent (into y0bar) and the
  // re-steps' parameter gradient (into pbar). Phase 2 reads both back in the same lane before
  // writing them, so the buffers are shared.
  const double* ybar_in = ybar;
  if (nsave > 0) {
    SaveVjpArgs v{n, p, aux, n_acc, ts, nsave, save_step, save_t, states, ysbar, ybar,
                  mu, y0bar, pbar};
    if (rb)
      rc = strict ? gsx_strict::rb23_saveat_vjp_dispatch(field, D, v, st)
                  : rb23_saveat_vjp_dispatch(field, D, v, st);
    else if (m == DenseMethod::kRodas5p)
      rc = strict ? gsx_strict::r5_saveat_vjp_dispatch(field, D, v, st)
                  : r5_saveat_vjp_dispatch(field, D, v, st);
    else
      rc = strict ? gsx_strict::tsit5_saveat_vjp_dispatch(field, D, v, st)
                  : tsit5_saveat_vjp_dispatch(field, D, v, st);
    if (rc) return rc;
    ybar_in = y0bar;
  } else if (!ybar) {
    return set_error(GSX_E_INVALID, "%s: no cotangent (ybar and nsave 0)", what);
  }
  AdjArgs a{n, y0, p, aux, t0, tape, tape_cap, n_acc, ckpt, ckpt_interval, ybar_in, y0bar,
            pbar};
  if (nsave > 0) {
    a.nsave = nsave;
    a.save_step = save_step;
    a.mu = mu;
    a.pbar_in = pbar;
  }
  if (rb)
    return strict ? gsx_strict::rb23_adj_dispatch(field, D, a, st)
                  : rb23_adj_dispatch(field, D, a, st);
  if (m == DenseMethod::kRodas5p)
    return strict ? gsx_strict::r5_adj_dispatch(field, D, a, st) : r5_adj_dispatch(field, D, a, st);
  return strict ? gsx_strict::tsit5_adj_dispatch(field, D, a, st)
                : tsit5_adj_dispatch(field, D, a, st);
}

}  // namespace

int gsx_tsit5_saveat(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                     const double* p, double t0, const double* tape, const int32_t* n_acc,
                     const double* y_final, const double* ts, int64_t nsave, double* ys,
                     int32_t* save_step, double* save_t, double* states, int flags,
                     void* stream) {
  return saveat_common(DenseMethod::kTsit5, "tsit5_saveat", field, n, D, P, aux, y0, p, t0,
                       tape, n_acc, y_final, ts, nsave, ys, save_step, save_t, states, flags,
                       stream);
}

int gsx_tsit5_adjoint_saveat(int field, int64_t n, int D, int P, const double* aux,
                             const double* y0, const double* p, double t0, const double* tape,
                             int64_t tape_cap, const int32_t* n_acc, const double* ckpt,
                             int32_t ckpt_interval, const double* ts, int64_t nsave,
                             const int32_t* save_step, const double* save_t,
                             const double* states, const double* ysbar, const double* ybar,
                             double* mu, double* y0bar, double* pbar, int flags,
                             void* stream) {
  return adjoint_saveat_common(DenseMethod::kTsit5, "tsit5_adjoint_saveat", field, n, D, P,
                               aux, y0, p, t0, tape, tape_cap, n_acc, ckpt, ckpt_interval, ts,
                               nsave, save_step, save_t, states, ysbar, ybar, mu, y0bar, pbar,
                               flags, stream);
}

int gsx_tsit5_adjoint_saveat_noalloc(int field, int64_t n, int D, int P, const double* aux,
                                     const double* y0, const double* p, double t0,
                                     const double* tape, int64_t tape_cap, const int32_t* n_acc,
                                     const double* ckpt, int32_t ckpt_interval,
                                     const double* ts, int64_t nsave,
                                     const int32_t* save_step, const double* save_t,
                                     const double* states, const double* ysbar,
                                     const double* ybar, double* y0bar, double* pbar,
                                     int flags, void* stream) {
  const char* what = "tsit5_adjoint_saveat_noalloc";
  if (rt::is_runtime(field))
    return set_error(GSX_E_UNSUPPORTED, "%s: fields compiled at run time load their kernels on "
                     "first use (an allocation); use gsx_tsit5_adjoint_saveat", what);
  if (n < 0 || nsave < 0 || D < 1) return set_error(GSX_E_INVALID, "%s: n %lld, nsave %lld, D %d",
                                                    what, static_cast<long long>(n),
                                                    static_cast<long long>(nsave), D);
  double* mu = nullptr;
  if (n > 0 && nsave > 0) {
    // nsave * D * n doubles must fit the scratch (each factor checked before multiplying)
    const int64_t cap = kNoallocScratchDoubles;
    if (n > cap || static_cast<int64_t>(D) > cap / n || nsave > cap / (static_cast<int64_t>(D) * n))
      return set_error(GSX_E_INVALID, "%s: nsave*D*n = %lld*%d*%lld doubles exceeds the "
                       "library's %lld-double (64 MiB) scratch; use gsx_tsit5_adjoint_saveat "
                       "with a caller buffer", what, static_cast<long long>(nsave), D,
                       static_cast<long long>(n), static_cast<long long>(cap));
    const int rc = noalloc_scratch(&mu);
    if (rc) return rc;
  }
  return adjoint_saveat_common(DenseMethod::kTsit5, what, field, n, D, P, aux, y0, p, t0, tape,
                               tape_cap, n_acc, ckpt, ckpt_interval, ts, nsave, save_step,
                               save_t, states, ysbar, ybar, mu, y0bar, pbar, flags, stream);
}

int gsx_rb23_saveat(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                    const double* p, double t0, const double* tape, const int32_t* n_acc,
                    const double* y_final, const double* ts, int64_t nsave, double* ys,
                    int32_t* save_step, double* save_t, double* states, int flags,
                    void* stream) {
  return saveat_common(DenseMethod::kRb23, "rb23_saveat", field, n, D, P, aux, y0, p, t0,
                       tape, n_acc, y_final, ts, nsave, ys, save_step, save_t, states, flags,
                       stream);
}

int gsx_rb23_adjoint_saveat(int field, int64_t n, int D, int P, const double* aux,
                            const double* y0, const double* p, double t0, const double* tape,
                            int64_t tape_cap, const int32_t* n_acc, const double* ckpt,
                            int32_t ckpt_interval, const double* ts, int64_t nsave,
                            const int32_t* save_step, const double* save_t,
                            const double* states, const double* ysbar, const double* ybar,
                            double* mu, double* y0bar, double* pbar, int flags, void* stream) {
  return adjoint_saveat_common(DenseMethod::kRb23, "rb23_adjoint_saveat", field, n, D, P, aux,
                               y0, p, t0, tape, tape_cap, n_acc, ckpt, ckpt_interval, ts, nsave,
                               save_step, save_t, states, ysbar, ybar, mu, y0bar, pbar, flags,
                               stream);
}

namespace {
// Rosenbrock23 flags: 0, GSX_STRICT, or GSX_FP32 (autonomous fields).
int rb23_flags(int field, int D, int flags) {
  if (flags & ~(GSX_STRICT | GSX_FP32)) return set_error(GSX_E_INVALID, "unknown flags 0x%x", flags);
  if ((flags & GSX_STRICT) && (flags & GSX_FP32))
    return set_error(GSX_E_INVALID, "GSX_STRICT and GSX_FP32 are exclusive");
  if (flags & GSX_FP32) {
    FieldInfo fi;
    if (field_info(field, D, &fi) && !fi.autonomous)
      return set_error(GSX_E_UNSUPPORTED, "rosenbrock23 FP32: field %d reads t; the FP32 "
                       "variant serves autonomous fields", field);
  }
  return GSX_OK;
}
}  // namespace

int gsx_rb23_fwd(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                 const double* p, double t0, double t1, const double* t1_per, double rtol,
                 double atol, int64_t max_steps, double* y_final, int32_t* n_acc,
                 int32_t* n_rej, int32_t* status, double* tape, int64_t tape_cap,
                 double* ckpt, int32_t ckpt_interval, int flags, void* stream) {
  FieldInfo fi;
  int rc = check_field(field, D, P, &fi);
  if (rc) return rc;
  if (n < 0 || max_steps < 0) return set_error(GSX_E_INVALID, "negative n or max_steps");
  // n_acc, n_rej and status may be null for a forward without a tape (the reference's
  // forward handler returns y_final and, optionally, the accepted-step count); a recording
  // forward needs n_acc and status (the tape's length and its "tape full" code).
  if (n > 0 && (!y0 || !p || !y_final))
    return set_error(GSX_E_INVALID, "rb23_fwd: null input/output pointer");
  if (n > 0 && tape && (!n_acc || !status))
    return set_error(GSX_E_INVALID, "rb23_fwd: a recording forward needs n_acc and status");
  if (tape && tape_cap < 0) return set_error(GSX_E_INVALID, "negative tape_cap");
  if (ckpt && (ckpt_interval < 1 || ckpt_interval > kMaxCkptInterval))
    return set_error(GSX_E_INVALID, "ckpt_interval %d outside [1, %d]", ckpt_interval,
                     kMaxCkptInterval);
  if (ckpt && !tape) return set_error(GSX_E_INVALID, "checkpoints need a tape");
  if (!(rtol > 0.0) || !(atol >= 0.0)) return set_error(GSX_E_INVALID, "bad rtol/atol");
  rc = rb23_flags(field, D, flags);
  if (rc) return rc;
  FwdArgs a{n, y0, p, aux, t0, t1, t1_per, rtol, atol, max_steps, y_final, n_acc, n_rej,
            status, tape, tape ? tape_cap : 0, ckpt, ckpt ? ckpt_interval : 1};
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (flags & GSX_FP32) return rb23_fwd_f32_dispatch(field, D, a, st);
  return (flags & GSX_STRICT) ? gsx_strict::rb23_fwd_dispatch(field, D, a, st)
                              : rb23_fwd_dispatch(field, D, a, st);
}

int gsx_rb23_replay(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                    const double* p, double t0, const double* tape, const int32_t* n_acc,
                    double* y_final, double* ckpt, int32_t ckpt_interval, int flags,
                    void* stream) {
  FieldInfo fi;
  int rc = check_field(field, D, P, &fi);
  if (rc) return rc;
  if (n < 0) return set_error(GSX_E_INVALID, "negative n");
  if (n > 0 && (!y0 || !p || !tape || !n_acc || !y_final))
    return set_error(GSX_E_INVALID, "rb23_replay: null pointer");
  if (ckpt && (ckpt_interval < 1 || ckpt_interval > kMaxCkptInterval))
    return set_error(GSX_E_INVALID, "ckpt_interval %d outside [1, %d]", ckpt_interval,
                     kMaxCkptInterval);
  rc = rb23_flags(field, D, flags);
  if (rc) return rc;
  ReplayArgs a{n, y0, p, aux, t0, tape, n_acc, y_final, ckpt, ckpt ? ckpt_interval : 1};
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (flags & GSX_FP32) return rb23_replay_f32_dispatch(field, D, a, st);
  return (flags & GSX_STRICT) ? gsx_strict::rb23_replay_dispatch(field, D, a, st)
                              : rb23_replay_dispatch(field, D, a, st);
}

int gsx_rb23_adjoint(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                     const double* p, double t0, const double* tape, int64_t tape_cap,
                     const int32_t* n_acc, const double* ckpt, int32_t ckpt_interval,
                     const double* ybar, double* y0bar, double* pbar, int flags,
                     void* stream) {
  FieldInfo fi;
  int rc = check_field(field, D, P, &fi);
  if (rc) return rc;
  if (n < 0) return set_error(GSX_E_INVALID, "negative n");
  if (n > 0 && (!p || !tape || !n_acc || !ckpt || !ybar || !y0bar || !pbar))
    return set_error(GSX_E_INVALID, "rb23_adjoint: null pointer");
  if (ckpt_interval < 1 || ckpt_interval > kMaxCkptInterval)
    return set_error(GSX_E_INVALID, "ckpt_interval %d outside [1, %d]", ckpt_interval,
                     kMaxCkptInterval);
  rc = rb23_flags(field, D, flags);
  if (rc) return rc;
  AdjArgs a{n, y0, p, aux, t0, tape, tape_cap, n_acc, ckpt, ckpt_interval, ybar, y0bar, pbar};
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (flags & GSX_FP32) return rb23_adj_f32_dispatch(field, D, a, st);
  return (flags & GSX_STRICT) ? gsx_strict::rb23_adj_dispatch(field, D, a, st)
                              : rb23_adj_dispatch(field, D, a, st);
}

// ---------------------------------------------------------------------------------------
// Rodas5P (gsx_rodas5p.hip): fields with an analytic Jacobian (the built-in stiff fields and
// fields registered with GSX_FIELD_STIFF); a field that reads t also supplies df/dt.
// ---------------------------------------------------------------------------------------
static int r5_check(int field, int D, int P, const char* what) {
  FieldInfo fi;
  int rc = check_field(field, D, P, &fi);
  if (rc) return rc;
  if (::gsx::rt::is_runtime(field) && !fi.stiff)
    return set_error(GSX_E_INVALID, "%s: field %d has no Jacobian (register it with "
                     "GSX_FIELD_STIFF)", what, field);
  return GSX_OK;
}

int gsx_rodas5p_fwd(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                    const double* p, double t0, double t1, const double* t1_per, double rtol,
                    double atol, int64_t max_steps, double* y_final, int32_t* n_acc,
                    int32_t* n_rej, int32_t* status, double* tape, int64_t tape_cap,
                    double* ckpt, int32_t ckpt_interval, int flags, void* stream) {
  int rc = r5_check(field, D, P, "rodas5p_fwd");
  if (rc) return rc;
  if (n < 0 || max_steps < 0) return set_error(GSX_E_INVALID, "negative n or max_steps");
  // n_acc, n_rej and status may be null for a forward without a tape (the reference's
  // forward handler returns y_final and, optionally, the accepted-step count); a recording
  // forward needs n_acc and status (the tape's length and its "tape full" code).
  if (n > 0 && (!y0 || !p || !y_final))
    return set_error(GSX_E_INVALID, "rodas5p_fwd: null input/output pointer");
  if (n > 0 && tape && (!n_acc || !status))
    return set_error(GSX_E_INVALID, "rodas5p_fwd: a recording forward needs n_acc and status");
  if (tape && tape_cap < 0) return set_error(GSX_E_INVALID, "negative tape_cap");
  if (ckpt && (ckpt_interval < 1 || ckpt_interval > kMaxCkptInterval))
    return set_error(GSX_E_INVALID, "ckpt_interval %d outside [1, %d]", ckpt_interval,
                     kMaxCkptInterval);
  if (ckpt && !tape) return set_error(GSX_E_INVALID, "checkpoints need a tape");
  if (!(rtol > 0.0) || !(atol >= 0.0)) return set_error(GSX_E_INVALID, "bad rtol/atol");
  if (flags & ~GSX_STRICT) return set_error(GSX_E_INVALID, "unknown flags 0x%x", flags);
  FwdArgs a{n, y0, p, aux, t0, t1, t1_per, rtol, atol, max_steps, y_final, n_acc, n_rej,
            status, tape, tape ? tape_cap : 0, ckpt, ckpt ? ckpt_interval : 1};
  hipStream_t st = static_cast<hipStream_t>(stream);
  return (flags & GSX_STRICT) ? gsx_strict::r5_fwd_dispatch(field, D, a, st)
                              : r5_fwd_dispatch(field, D, a, st);
}

int gsx_rodas5p_replay(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                       const double* p, double t0, const double* tape, const int32_t* n_acc,
                       double* y_final, double* ckpt, int32_t ckpt_interval, int flags,
                       void* stream) {
  int rc = r5_check(field, D, P, "rodas5p_replay");
  if (rc) return rc;
  if (n < 0) return set_error(GSX_E_INVALID, "negative n");
  if (n > 0 && (!y0 || !p || !tape || !n_acc || !y_final))
    return set_error(GSX_E_INVALID, "rodas5p_replay: null pointer");
  if (ckpt && (ckpt_interval < 1 || ckpt_interval > kMaxCkptInterval))
    return set_error(GSX_E_INVALID, "ckpt_interval %d outside [1, %d]", ckpt_interval,
                     kMaxCkptInterval);
  if (flags & ~GSX_STRICT) return set_error(GSX_E_INVALID, "unknown flags 0x%x", flags);
  ReplayArgs a{n, y0, p, aux, t0, tape, n_acc, y_final, ckpt, ckpt ? ckpt_interval : 1};
  hipStream_t st = static_cast<hipStream_t>(stream);
  return (flags & GSX_STRICT) ? gsx_strict::r5_replay_dispatch(field, D, a, st)
                              : r5_replay_dispatch(field, D, a, st);
}

int gsx_rodas5p_adjoint(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                        const double* p, double t0, const double* tape, int64_t tape_cap,
                        const int32_t* n_acc, const double* ckpt, int32_t ckpt_interval,
                        const double* ybar, double* y0bar, double* pbar, int flags,
                        void* stream) {
  int rc = r5_check(field, D, P, "rodas5p_adjoint");
  if (rc) return rc;
  if (n < 0) return set_error(GSX_E_INVALID, "negative n");
  if (n > 0 && (!p || !tape || !n_acc || !ckpt || !ybar || !y0bar || !pbar))
    return set_error(GSX_E_INVALID, "rodas5p_adjoint: null pointer");
  if (ckpt_interval < 1 || ckpt_interval > kMaxCkptInterval)
    return set_error(GSX_E_INVALID, "ckpt_interval %d outside [1, %d]", ckpt_interval,
                     kMaxCkptInterval);
  if (flags & ~GSX_STRICT) return set_error(GSX_E_INVALID, "unknown flags 0x%x", flags);
  AdjArgs a{n, y0, p, aux, t0, tape, tape_cap, n_acc, ckpt, ckpt_interval, ybar, y0bar, pbar};
  hipStream_t st = static_cast<hipStream_t>(stream);
  return (flags & GSX_STRICT) ? gsx_strict::r5_adj_dispatch(field, D, a, st)
                              : r5_adj_dispatch(field, D, a, st);
}

int gsx_rodas5p_saveat(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                       const double* p, double t0, const double* tape, const int32_t* n_acc,
                       const double* y_final, const double* ts, int64_t nsave, double* ys,
                       int32_t* save_step, double* save_t, double* states, int flags,
                       void* stream) {
  int rc = r5_check(field, D, P, "rodas5p_saveat");
  if (rc) return rc;
  return saveat_common(DenseMethod::kRodas5p, "rodas5p_saveat", field, n, D, P, aux, y0, p, t0,
                       tape, n_acc, y_final, ts, nsave, ys, save_step, save_t, states, flags,
                       stream);
}

int gsx_rodas5p_adjoint_saveat(int field, int64_t n, int D, int P, const double* aux,
                               const double* y0, const double* p, double t0,
                               const double* tape, int64_t tape_cap, const int32_t* n_acc,
                               const double* ckpt, int32_t ckpt_interval, const double* ts,
                               int64_t nsave, const int32_t* save_step, const double* save_t,
                               const double* states, const double* ysbar, const double* ybar,
                               double* mu, double* y0bar, double* pbar, int flags,
                               void* stream) {
  int rc = r5_check(field, D, P, "rodas5p_adjoint_saveat");
  if (rc) return rc;
  return adjoint_saveat_common(DenseMethod::kRodas5p, "rodas5p_adjoint_saveat", field, n, D, P,
                               aux, y0, p, t0, tape, tape_cap, n_acc, ckpt, ckpt_interval, ts,
                               nsave, save_step, save_t, states, ysbar, ybar, mu, y0bar, pbar,
                               flags, stream);
}

int gsx_rodas5p_saveat_dense(int field, int64_t n, int D, int P, const double* aux,
                             const double* y0, const double* p, double t0, const double* tape,
                             const int32_t* n_acc, const double* y_final, const double* ts,
                             int64_t nsave, double* ys, int32_t* save_step, double* save_t,
                             int flags, void* stream) {
  int rc = r5_check(field, D, P, "rodas5p_saveat_dense");
  if (rc) return rc;
  if (n < 0 || nsave < 0) return set_error(GSX_E_INVALID, "negative n or nsave");
  if (flags & ~GSX_STRICT) return set_error(GSX_E_INVALID, "unknown flags 0x%x", flags);
  if (n == 0 || nsave == 0) return GSX_OK;
  if (!y0 || !p || !tape || !n_acc || !ts || !ys || !save_step || !save_t)
    return set_error(GSX_E_INVALID, "rodas5p_saveat_dense: null pointer");
  SaveArgs a{n, y0, p, aux, t0, tape, n_acc, ts, nsave, ys, save_step, save_t, nullptr, y_final};
  hipStream_t st = static_cast<hipStream_t>(stream);
  return (flags & GSX_STRICT) ? gsx_strict::r5_saveat_dispatch(field, D, a, true, st)
                              : r5_saveat_dispatch(field, D, a, true, st);
}

int gsx_rodas5p_adjoint_saveat_dense(int field, int64_t n, int D, int P, const double* aux,
                                     const double* y0, const double* p, double t0,
                                     const double* tape, int64_t tape_cap,
                                     const int32_t* n_acc, const double* ckpt,
                                     int32_t ckpt_interval, const double* ts, int64_t nsave,
                                     const int32_t* save_step, const double* save_t,
                                     const double* ysbar, const double* ybar, double* y0bar,
                                     double* pbar, int flags, void* stream) {
  int rc = r5_check(field, D, P, "rodas5p_adjoint_saveat_dense");
  if (rc) return rc;
  if (n < 0 || nsave < 0) return set_error(GSX_E_INVALID, "negative n or nsave");
  if (flags & ~GSX_STRICT) return set_error(GSX_E_INVALID, "unknown flags 0x%x", flags);
  if (ckpt_interval < 1 || ckpt_interval > kMaxCkptInterval)
    return set_error(GSX_E_INVALID, "ckpt_interval %d outside [1, %d]", ckpt_interval,
                     kMaxCkptInterval);
  if (n == 0) return GSX_OK;
  if (!p || !tape || !n_acc || !ckpt || !y0bar || !pbar ||
      (nsave > 0 && (!ts || !save_step || !save_t || !ysbar)))
    return set_error(GSX_E_INVALID, "rodas5p_adjoint_saveat_dense: null pointer");
  if (nsave == 0 && !ybar)
    return set_error(GSX_E_INVALID, "rodas5p_adjoint_saveat_dense: no cotangent");
  AdjArgs a{n, y0, p, aux, t0, tape, tape_cap, n_acc, ckpt, ckpt_interval, ybar, y0bar, pbar};
  if (nsave > 0) {
    a.nsave = nsave;
    a.save_step = save_step;
    a.ts = ts;
    a.save_t = save_t;
    a.ysbar = ysbar;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  return (flags & GSX_STRICT) ? gsx_strict::r5_adj_dispatch(field, D, a, st)
                              : r5_adj_dispatch(field, D, a, st);
}

int64_t gsx_persist_lanes(int waves_per_simd) {
  if (waves_per_simd < 1 || waves_per_simd > 8)
    return set_error(GSX_E_INVALID, "waves_per_simd %d outside [1, 8]", waves_per_simd);
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return set_error(GSX_E_HIP, "persist_lanes: cannot query the compute-unit count");
  return static_cast<int64_t>(cus) * 4 * waves_per_simd * 64;
}

int32_t gsx_persist_record_len(int field, int D, int32_t ckpt_interval) {
  FieldInfo fi;
  if (!field_info(field, D, &fi))
    return set_error(GSX_E_UNSUPPORTED, "field %d with D=%d is not built", field, D);
  if (ckpt_interval < 1 || ckpt_interval > kMaxCkptInterval)
    return set_error(GSX_E_INVALID, "ckpt_interval %d outside [1, %d]", ckpt_interval,
                     kMaxCkptInterval);
  return gsx::persist_record_len(fi.D + (fi.autonomous ? 0 : 1), ckpt_interval);
}

int gsx_tsit5_persist_fwd(int field, int64_t n, int D, int P, const double* aux,
                          const double* y0, const double* p, double t0, double t1,
                          const double* t1_per, double rtol, double atol, int64_t max_steps,
                          double* y_final, int32_t* n_acc, int32_t* n_rej, int32_t* status,
                          int64_t lanes, double* rec, int64_t rec_cap, int32_t ckpt_interval,
                          int32_t* order, int64_t traj_cap, int32_t* lane_count,
                          int32_t* lane_of, int32_t* rec_start, void* queue,
                          const int32_t* queue_order, int flags, void* stream) {
  FieldInfo fi;
  int rc = check_field(field, D, P, &fi);
  if (rc) return rc;
  if (fi.stiff) return set_error(GSX_E_INVALID, "field %d is stiff", field);
  if (n < 0 || n > INT32_MAX || max_steps < 0)
    return set_error(GSX_E_INVALID, "n must be in [0, 2^31) and max_steps >= 0");
  if (lanes <= 0 || lanes % 256 != 0)
    return set_error(GSX_E_INVALID, "lanes %lld must be a positive multiple of 256",
                     static_cast<long long>(lanes));
  if (rec_cap < 1 || traj_cap < 1 || rec_cap > INT32_MAX || traj_cap > INT32_MAX)
    return set_error(GSX_E_INVALID, "capacities must be in [1, 2^31)");
  if (ckpt_interval < 1 || ckpt_interval > kMaxCkptInterval)
    return set_error(GSX_E_INVALID, "ckpt_interval %d outside [1, %d]", ckpt_interval,
                     kMaxCkptInterval);
  if (flags & ~GSX_FP32) return set_error(GSX_E_INVALID, "unknown flags 0x%x", flags);
  if (!queue || !rec || !order || !lane_count ||
      (n > 0 && (!y0 || !p || !y_final || !n_acc || !n_rej || !status || !lane_of ||
                 !rec_start)))
    return set_error(GSX_E_INVALID, "tsit5_persist_fwd: null pointer");
  if (!(rtol > 0.0) || !(atol >= 0.0)) return set_error(GSX_E_INVALID, "bad rtol/atol");
  PersistFwdArgs pa{};
  pa.f = FwdArgs{n, y0, p, aux, t0, t1, t1_per, rtol, atol, max_steps, y_final, n_acc, n_rej,
                 status, nullptr, 0, nullptr, ckpt_interval};
  pa.lanes = lanes;
  pa.rec = rec;
  pa.rec_cap = rec_cap;
  pa.rec_len = gsx::persist_record_len(fi.D + (fi.autonomous ? 0 : 1), ckpt_interval);
  pa.traj_cap = traj_cap;
  pa.order = order;
  pa.lane_count = lane_count;
  pa.lane_of = lane_of;
  pa.rec_start = rec_start;
  pa.queue = static_cast<unsigned long long*>(queue);
  pa.queue_order = queue_order;
  return persist_fwd_dispatch(field, D, (flags & GSX_FP32) != 0, pa,
                              static_cast<hipStream_t>(stream));
}

int gsx_tsit5_persist_adjoint(int field, int64_t n, int D, int P, const double* aux,
                              const double* p, double t0, const int32_t* n_acc, int64_t lanes,
                              const double* rec, int32_t ckpt_interval, const int32_t* order,
                              const int32_t* lane_count, const int32_t* rec_start,
                              const double* ybar, double* y0bar, double* pbar, int flags,
                              void* stream) {
  FieldInfo fi;
  int rc = check_field(field, D, P, &fi);
  if (rc) return rc;
  if (n < 0 || lanes <= 0 || lanes % 256 != 0)
    return set_error(GSX_E_INVALID, "bad n or lanes");
  if (ckpt_interval < 1 || ckpt_interval > kMaxCkptInterval)
    return set_error(GSX_E_INVALID, "ckpt_interval %d outside [1, %d]", ckpt_interval,
                     kMaxCkptInterval);
  if (flags & ~GSX_FP32) return set_error(GSX_E_INVALID, "unknown flags 0x%x", flags);
  if (!rec || !order || !lane_count ||
      (n > 0 && (!p || !n_acc || !rec_start || !ybar || !y0bar || !pbar)))
    return set_error(GSX_E_INVALID, "tsit5_persist_adjoint: null pointer");
  PersistAdjArgs pa{};
  pa.a = AdjArgs{n, nullptr, p, aux, t0, nullptr, 0, n_acc, nullptr, ckpt_interval, ybar,
                 y0bar, pbar};
  pa.lanes = lanes;
  pa.rec = rec;
  pa.rec_len = gsx::persist_record_len(fi.D + (fi.autonomous ? 0 : 1), ckpt_interval);
  pa.order = order;
  pa.lane_count = lane_count;
  pa.rec_start = rec_start;
  return persist_adj_dispatch(field, D, (flags & GSX_FP32) != 0, pa,
                              static_cast<hipStream_t>(stream));
}

// ---------------------------------------------------------------------------------------
// Kuramoto ensemble (D = P = 32, shared K): gsx_kuramoto.hip.
// ---------------------------------------------------------------------------------------
static int kura_check_flags(int flags) {
  if (flags & ~GSX_FP32) return set_error(GSX_E_INVALID, "unknown flags 0x%x", flags);
  return GSX_OK;
}

static int kura_check_d(int D, const char* what) {
  if (D != 32)
    return set_error(GSX_E_UNSUPPORTED, "%s: the Kuramoto kernels are built for D = 32, got %d",
                     what, D);
  return GSX_OK;
}

int gsx_kuramoto_fwd(int64_t n, int D, const double* K, const double* y0, const double* omega,
                     double t0, double t1, double rtol, double atol, int64_t max_steps,
                     const double* ts, int64_t nsave, double* y_final, int32_t* n_acc,
                     int32_t* n_rej, int32_t* status, double* tape, int64_t tape_cap,
                     double* states, int32_t* save_step, double* save_t, int flags,
                     void* stream) {
  int rc = kura_check_d(D, "kuramoto_fwd");
  if (rc) return rc;
  rc = kura_check_flags(flags);
  if (rc) return rc;
  if (n < 0 || n > INT32_MAX || max_steps < 0 || nsave < 0)
    return set_error(GSX_E_INVALID, "kuramoto_fwd: n in [0, 2^31), max_steps >= 0 and "
                     "nsave >= 0 required");
  if (!(rtol > 0.0) || !(atol >= 0.0)) return set_error(GSX_E_INVALID, "bad rtol/atol");
  if (tape && (tape_cap < 1 || tape_cap > INT32_MAX))
    return set_error(GSX_E_INVALID, "kuramoto_fwd: tape_cap must be in [1, 2^31)");
  if (!tape && (states || nsave > 0))
    return set_error(GSX_E_INVALID, "kuramoto_fwd: states and save times need the tape");
  if (n > 0 && (!K || !y0 || !omega || !y_final || !n_acc || !n_rej || !status))
    return set_error(GSX_E_INVALID, "kuramoto_fwd: null pointer");
  if (nsave > 0 && (!ts || !save_step || !save_t || !states))
    return set_error(GSX_E_INVALID, "kuramoto_fwd: save times need ts, save_step, save_t "
                     "and states");
  if (n == 0) return GSX_OK;
  KuraFwdArgs a{};
  a.n = n;
  a.K = K;
  a.y0 = y0;
  a.omega = omega;
  a.t0 = t0;
  a.t1 = t1;
  a.rtol = rtol;
  a.atol = atol;
  a.max_steps = max_steps;
  a.ts = nsave > 0 ? ts : nullptr;
  a.nsave = nsave;
  a.y_final = y_final;
  a.n_acc = n_acc;
  a.n_rej = n_rej;
  a.status = status;
  a.tape = tape;
  a.tape_cap = tape ? tape_cap : 0;
  a.states = states;
  a.save_step = nsave > 0 ? save_step : nullptr;
  a.save_t = nsave > 0 ? save_t : nullptr;
  return kura_fwd_launch(a, (flags & GSX_FP32) != 0, static_cast<hipStream_t>(stream));
}

int gsx_kuramoto_replay(int64_t n, int D, const double* K, const double* y0,
                        const double* omega, const double* tape, const int32_t* n_acc,
                        double* y_final, double* states, int flags, void* stream) {
  int rc = kura_check_d(D, "kuramoto_replay");
  if (rc) return rc;
  rc = kura_check_flags(flags);
  if (rc) return rc;
  if (n < 0 || n > INT32_MAX) return set_error(GSX_E_INVALID, "kuramoto_replay: bad n");
  if (n > 0 && (!K || !y0 || !omega || !tape || !n_acc || !y_final))
    return set_error(GSX_E_INVALID, "kuramoto_replay: null pointer");
  if (n == 0) return GSX_OK;
  KuraReplayArgs a{n, K, y0, omega, tape, n_acc, y_final, states};
  return kura_replay_launch(a, (flags & GSX_FP32) != 0, static_cast<hipStream_t>(stream));
}

int gsx_kuramoto_saveat(int64_t n, int D, const double* K, const double* y0,
                        const double* omega, const double* y_final, const double* states,
                        const double* ts, int64_t nsave, const int32_t* save_step,
                        const double* save_t, double* ys, int flags, void* stream) {
  int rc = kura_check_d(D, "kuramoto_saveat");
  if (rc) return rc;
  rc = kura_check_flags(flags);
  if (rc) return rc;
  if (n < 0 || n > INT32_MAX || nsave < 0)
    return set_error(GSX_E_INVALID, "kuramoto_saveat: bad n or nsave");
  if (n > 0 && nsave > 0 &&
      (!K || !y0 || !omega || !y_final || !states || !ts || !save_step || !save_t || !ys))
    return set_error(GSX_E_INVALID, "kuramoto_saveat: null pointer");
  if (n == 0 || nsave == 0) return GSX_OK;
  KuraSaveatArgs a{n, K, y0, omega, y_final, states, ts, nsave, save_step, save_t, ys};
  return kura_saveat_launch(a, (flags & GSX_FP32) != 0, static_cast<hipStream_t>(stream));
}

int gsx_kuramoto_adjoint(int64_t n, int D, const double* K, const double* omega,
                         const double* tape, const int32_t* n_acc, const double* states,
                         const double* ts, int64_t nsave, const int32_t* save_step,
                         const double* save_t, const double* ysbar, const double* yfbar,
                         double* mu, double* y0bar, double* omega_bar, int flags,
                         void* stream) {
  int rc = kura_check_d(D, "kuramoto_adjoint");
  if (rc) return rc;
  rc = kura_check_flags(flags);
  if (rc) return rc;
  if (n < 0 || n > INT32_MAX || nsave < 0)
    return set_error(GSX_E_INVALID, "kuramoto_adjoint: bad n or nsave");
  if (n > 0 && (!K || !omega || !tape || !n_acc || !states || !y0bar || !omega_bar))
    return set_error(GSX_E_INVALID, "kuramoto_adjoint: null pointer");
  if (ysbar && (nsave < 1 || !ts || !save_step || !save_t || !mu))
    return set_error(GSX_E_INVALID, "kuramoto_adjoint: save cotangents need ts, save_step, "
                     "save_t and the mu buffer");
  if (ysbar && mu == ysbar)
    return set_error(GSX_E_INVALID, "kuramoto_adjoint: mu must not alias ysbar");
  if (n == 0) return GSX_OK;
  KuraAdjArgs a{};
  a.n = n;
  a.K = K;
  a.omega = omega;
  a.tape = tape;
  a.n_acc = n_acc;
  a.states = states;
  a.ts = ts;
  a.nsave = ysbar ? nsave : 0;
  a.save_step = save_step;
  a.save_t = save_t;
  a.ysbar = ysbar;
  a.yfbar = yfbar;
  a.mu = mu;
  a.y0bar = y0bar;
  a.omega_bar = omega_bar;
  return kura_adj_launch(a, (flags & GSX_FP32) != 0, static_cast<hipStream_t>(stream));
}

// ---------------------------------------------------------------------------------------
// The naive reverse sweep (benchmark plain arm): gsx_plain.hip.
// ---------------------------------------------------------------------------------------
int gsx_plain_tsit5_fwd(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                        const double* p, double t0, double t1, double rtol, double atol,
                        int64_t max_steps, double* y_final, int32_t* n_acc, int32_t* n_rej,
                        int32_t* status, double* tape, double* stages, int64_t cap,
                        void* stream) {
  FieldInfo fi;
  int rc = check_field(field, D, P, &fi);
  if (rc) return rc;
  if (fi.stiff) return set_error(GSX_E_INVALID, "field %d is stiff", field);
  if (n < 0 || n > INT32_MAX || max_steps < 0 || cap < 1 || cap > INT32_MAX)
    return set_error(GSX_E_INVALID, "plain_tsit5_fwd: bad n, max_steps or cap");
  if (!(rtol > 0.0) || !(atol >= 0.0)) return set_error(GSX_E_INVALID, "bad rtol/atol");
  if (n > 0 && (!y0 || !p || !y_final || !n_acc || !n_rej || !status || !tape || !stages))
    return set_error(GSX_E_INVALID, "plain_tsit5_fwd: null pointer");
  PlainArgs a{};
  a.n = n;
  a.aux = aux;
  a.y0 = y0;
  a.p = p;
  a.t0 = t0;
  a.t1 = t1;
  a.rtol = rtol;
  a.atol = atol;
  a.max_steps = max_steps;
  a.y_final = y_final;
  a.n_acc = n_acc;
  a.n_rej = n_rej;
  a.status = status;
  a.tape = tape;
  a.stages = stages;
  a.cap = cap;
  return plain_dispatch(field, false, a, static_cast<hipStream_t>(stream));
}

int gsx_plain_tsit5_adjoint(int field, int64_t n, int D, int P, const double* aux,
                            const double* p, const double* tape, const double* stages,
                            const int32_t* n_acc, const double* ybar, double* y0bar,
                            double* pbar, void* stream) {
  FieldInfo fi;
  int rc = check_field(field, D, P, &fi);
  if (rc) return rc;
  if (n < 0 || n > INT32_MAX) return set_error(GSX_E_INVALID, "plain_tsit5_adjoint: bad n");
  if (n > 0 && (!p || !tape || !stages || !n_acc || !ybar || !y0bar || !pbar))
    return set_error(GSX_E_INVALID, "plain_tsit5_adjoint: null pointer");
  PlainArgs a{};
  a.n = n;
  a.aux = aux;
  a.p = p;
  a.tape = const_cast<double*>(tape);
  a.stages = const_cast<double*>(stages);
  a.n_acc = const_cast<int32_t*>(n_acc);
  a.ybar = ybar;
  a.y0bar = y0bar;
  a.pbar = pbar;
  return plain_dispatch(field, true, a, static_cast<hipStream_t>(stream));
}

int gsx_tape_to_padded(int64_t n, const double* tape, int64_t tape_cap, const int32_t* n_acc,
                       int64_t S, double* out, void* stream) {
  if (n < 0 || S < 0 || S > tape_cap)
    return set_error(GSX_E_INVALID, "tape_to_padded: n %lld, S %lld, cap %lld",
                     static_cast<long long>(n), static_cast<long long>(S),
                     static_cast<long long>(tape_cap));
  if (n == 0 || S == 0) return GSX_OK;
  if (!tape || !n_acc || !out) return set_error(GSX_E_INVALID, "tape_to_padded: null pointer");
  const int64_t total = n * S;
  const unsigned grid = static_cast<unsigned>((total + 255) / 256);
  tape_to_padded_kernel<<<grid, 256, 0, static_cast<hipStream_t>(stream)>>>(n, tape, n_acc, S,
                                                                            out);
  return check_launch("tape_to_padded_kernel");
}

int gsx_tape_from_padded(int64_t n, const double* in, int64_t S, double* tape,
                         int64_t tape_cap, int32_t* n_acc, void* stream) {
  if (n < 0 || S < 0 || S > tape_cap)
    return set_error(GSX_E_INVALID, "tape_from_padded: n %lld, S %lld, cap %lld",
                     static_cast<long long>(n), static_cast<long long>(S),
                     static_cast<long long>(tape_cap));
  if (n == 0) return GSX_OK;
  if ((S > 0 && !in) || !tape || !n_acc)
    return set_error(GSX_E_INVALID, "tape_from_padded: null pointer");
  const unsigned grid = static_cast<unsigned>((n + 255) / 256);
  tape_from_padded_kernel<<<grid, 256, 0, static_cast<hipStream_t>(stream)>>>(n, in, S, tape,
                                                                              tape_cap, n_acc);
  return check_launch("tape_from_padded_kernel");
}

int gsx_tape_zero_tail(int64_t n, double* tape, int64_t tape_cap, const int32_t* n_acc,
                       int64_t S, void* stream) {
  if (n < 0 || S < 0 || S > tape_cap)
    return set_error(GSX_E_INVALID, "tape_zero_tail: n %lld, S %lld, cap %lld",
                     static_cast<long long>(n), static_cast<long long>(S),
                     static_cast<long long>(tape_cap));
  if (n == 0 || S == 0) return GSX_OK;
  if (!tape || !n_acc) return set_error(GSX_E_INVALID, "tape_zero_tail: null pointer");
  const int64_t total = n * S;
  const unsigned grid = static_cast<unsigned>((total + 255) / 256);
  tape_zero_tail_kernel<<<grid, 256, 0, static_cast<hipStream_t>(stream)>>>(n, tape, n_acc, S);
  return check_launch("tape_zero_tail_kernel");
}

int gsx_tape_count(int64_t n, const double* tape, int64_t S, int32_t* n_acc, void* stream) {
  if (n < 0 || S < 0) return set_error(GSX_E_INVALID, "tape_count: negative size");
  if (n == 0) return GSX_OK;
  if ((S > 0 && !tape) || !n_acc) return set_error(GSX_E_INVALID, "tape_count: null pointer");
  const unsigned grid = static_cast<unsigned>((n + 255) / 256);
  tape_count_kernel<<<grid, 256, 0, static_cast<hipStream_t>(stream)>>>(n, tape, S, n_acc);
  return check_launch("tape_count_kernel");
}

}  // extern "C"
