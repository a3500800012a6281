// CPU build of the per-lane kernel bodies (gradsolve_amd/csrc/gsx_core.h), for tests only:
// runs the exact lane code the HIP kernels run, one trajectory at a time, with bounds-checked
// buffers sized as the C ABI sizes them. Lets the CPU test suite check indexing, layout and
// arithmetic against the oracle without a GPU. Never part of the product library.
#include <cstdint>
#include <vector>

#include "../../gradsolve_amd/csrc/gsx_core.h"
#include "../../gradsolve_amd/csrc/gsx_rb23.h"
#include "../../gradsolve_amd/csrc/gsx_rodas5p.h"
#include "../../gradsolve_amd/csrc/gsx_f32.h"

using namespace gsx;

namespace {
template <class F>
int run_fwd(const FwdArgs& a) {
  for (int64_t j = 0; j < a.n; ++j) tsit5_fwd_lane<F>(a, j);
  return 0;
}
template <class F>
int run_replay(const ReplayArgs& a) {
  for (int64_t j = 0; j < a.n; ++j) tsit5_replay_lane<F>(a, j);
  return 0;
}
template <class F>
int run_adj(const AdjArgs& a) {
  std::vector<double> seg(static_cast<size_t>(a.ckpt_interval) * ckpt_rows<F>());
  for (int64_t j = 0; j < a.n; ++j) tsit5_adj_lane<F>(a, j, seg.data(), 1);
  return 0;
}
template <class F>
int run_adj_pair(const AdjArgs& a) {
  for (int64_t j = 0; j < a.n; ++j) {
    const int32_t na = a.n_acc[j];
    tsit5_adj_pair_lane<F>(a, j, na, (na + 1) / 2);
  }
  return 0;
}
template <class F>
int run_saveat(const SaveArgs& a) {
  for (int64_t j = 0; j < a.n; ++j) tsit5_saveat_lane<F>(a, j);
  return 0;
}
// The saveat gradient as gsx_tsit5_adjoint_saveat runs it: the re-steps' VJPs, then the
// segment (C != 2) or pair (C == 2) adjoint with the saves injected.
template <class F>
int run_adj_saveat(const SaveVjpArgs& v, AdjArgs a) {
  for (int64_t j = 0; j < v.n; ++j) tsit5_saveat_vjp_lane<F>(v, j);
  a.ybar = v.ybar_eff;
  a.nsave = v.nsave;
  a.save_step = v.save_step;
  a.mu = v.mu;
  a.pbar_in = v.pbar;
  if (a.ckpt_interval == 2) return run_adj_pair<F>(a);
  return run_adj<F>(a);
}
template <class F>
int run_rb_fwd(const FwdArgs& a) {
  for (int64_t j = 0; j < a.n; ++j) rb23_fwd_lane<F>(a, j);
  return 0;
}
template <class F>
int run_rb_replay(const ReplayArgs& a) {
  for (int64_t j = 0; j < a.n; ++j) rb23_replay_lane<F>(a, j);
  return 0;
}
// As gsx_rb23_adjoint dispatches: the pair adjoint at C = 2 (with the save injection when save
// cotangents or an incoming pbar are given), the segment adjoint otherwise.
template <class F>
int run_rb_adj(const AdjArgs& a) {
  if (a.ckpt_interval == 2) {
    const bool saves = a.nsave > 0 || a.pbar_in;
    for (int64_t j = 0; j < a.n; ++j) {
      const int32_t na = a.n_acc[j];
      if (saves)
        rb23_adj_pair_lane<F, true>(a, j, na, (na + 1) / 2);
      else
        rb23_adj_pair_lane<F, false>(a, j, na, (na + 1) / 2);
    }
    return 0;
  }
  std::vector<double> seg(static_cast<size_t>(a.ckpt_interval) * ckpt_rows<F>());
  for (int64_t j = 0; j < a.n; ++j) rb23_adj_lane<F>(a, j, seg.data(), 1);
  return 0;
}
template <class F>
int run_rb_saveat(const SaveArgs& a) {
  for (int64_t j = 0; j < a.n; ++j) rb23_saveat_lane<F>(a, j);
  return 0;
}
// As gsx_rb23_adjoint_saveat runs it: the re-steps' VJPs, then the segment adjoint with the
// saves injected.
template <class F>
int run_rb_adj_saveat(const SaveVjpArgs& v, AdjArgs a) {
  for (int64_t j = 0; j < v.n; ++j) rb23_saveat_vjp_lane<F>(v, j);
  a.ybar = v.ybar_eff;
  a.nsave = v.nsave;
  a.save_step = v.save_step;
  a.mu = v.mu;
  a.pbar_in = v.pbar;
  return run_rb_adj<F>(a);
}
template <class F>
int run_r5_adj(const AdjArgs& a) {
  std::vector<double> seg(static_cast<size_t>(a.ckpt_interval) * ckpt_rows<F>());
  for (int64_t j = 0; j < a.n; ++j) r5_adj_lane<F>(a, j, seg.data(), 1);
  return 0;
}
}  // namespace

extern "C" {
int lh_tsit5_f32(int field, int64_t n, const double* y0, const double* p, double t0, double t1,
                 const double* t1_per, double rtol, double atol, int64_t max_steps,
                 double* y_final, int32_t* n_acc, int32_t* n_rej, int32_t* status, double* tape,
                 int64_t cap, double* ckpt, int32_t C, const double* ybar, double* y0bar,
                 double* pbar) {
  if (field != 1) return -1;
  FwdArgs a{n, y0, p, nullptr, t0, t1, t1_per, rtol, atol, max_steps, y_final, n_acc, n_rej,
            status, tape, cap, ckpt, C};
  for (int64_t j = 0; j < n; ++j) tsit5_fwd_lane_f32<FLorenz>(a, j);
  if (!ybar) return 0;
  AdjArgs b{n, nullptr, p, nullptr, t0, tape, cap, n_acc, ckpt, C, ybar, y0bar, pbar};
  std::vector<double> seg(static_cast<size_t>(C) * 3);
  for (int64_t j = 0; j < n; ++j) tsit5_adj_lane_f32<FLorenz>(b, j, seg.data(), 1);
  return 0;
}
// The controller's clamped step-size factor clamp(0.9 * max(err, 1e-10)^e, 0.2, 5) with the
// build's power (which: 0 Tsit5, 1 Rosenbrock23), and the raw power.
double lh_ctrl_fac(int which, double err) {
  const double x = py_max(err, kErrFloor);
  const double pw = which == 0 ? ctrl_pow_tsit5(x) : ctrl_pow_rb23(x);
  return py_min(py_max(kSafety * pw, kFacMin), kFacMax);
}
void lh_sincos(int64_t n, const double* x, double* s, double* c) {
  for (int64_t i = 0; i < n; ++i) gsx_sincos(x[i], s + i, c + i);
}

void lh_sincosf(int64_t n, const float* x, float* s, float* c) {
  for (int64_t i = 0; i < n; ++i) gsx_sincosf(x[i], s + i, c + i);
}

// The default variant's Tsit5 factor from the squared norm m = err^2, clamped.
double lh_ctrl_fac_sq(double m) {
  return py_min(py_max(tsit5_fac(m), kFacMin), kFacMax);
}
double lh_norm_div(double a, double b) { return norm_div(a, b); }
// (q0^2 + q1^2 + q2^2) / 3 as the default variant's squared norm forms it (unit scales).
double lh_mean_sq3(const double* q) {
  const double z[3] = {0.0, 0.0, 0.0};
  return wrms_sq<3>(q, z, z, 0.0, 1.0);
}
double lh_ctrl_pow(int which, double x) {
  return which == 0 ? ctrl_pow_tsit5(x) : ctrl_pow_rb23(x);
}
int lh_rb23_fwd(int field, int64_t n, const double* y0, const double* p, double t0, double t1,
                const double* t1_per, double rtol, double atol, int64_t max_steps,
                double* y_final, int32_t* n_acc, int32_t* n_rej, int32_t* status, double* tape,
                int64_t cap, double* ckpt, int32_t C) {
  FwdArgs a{n, y0, p, nullptr, t0, t1, t1_per, rtol, atol, max_steps, y_final, n_acc, n_rej,
            status, tape, tape ? cap : 0, ckpt, ckpt ? C : 1};
  if (field == 3) return run_rb_fwd<FRobertson>(a);
  if (field == 8) return run_rb_fwd<FForcedStiff>(a);
  if (field == 2) return run_rb_fwd<FLinearDecay<2>>(a);
  return -1;
}
int lh_rb23_replay(int field, int64_t n, const double* y0, const double* p, double t0,
                   const double* tape, const int32_t* n_acc, double* y_final, double* ckpt,
                   int32_t C) {
  ReplayArgs a{n, y0, p, nullptr, t0, tape, n_acc, y_final, ckpt, ckpt ? C : 1};
  if (field == 3) return run_rb_replay<FRobertson>(a);
  if (field == 8) return run_rb_replay<FForcedStiff>(a);
  if (field == 2) return run_rb_replay<FLinearDecay<2>>(a);
  return -1;
}
int lh_rb23_adjoint(int field, int64_t n, const double* p, double t0, const double* tape,
                    int64_t cap, const int32_t* n_acc, const double* ckpt, int32_t C,
                    const double* ybar, double* y0bar, double* pbar) {
  AdjArgs a{n, nullptr, p, nullptr, t0, tape, cap, n_acc, ckpt, C, ybar, y0bar, pbar};
  if (field == 3) return run_rb_adj<FRobertson>(a);
  if (field == 8) return run_rb_adj<FForcedStiff>(a);
  if (field == 2) return run_rb_adj<FLinearDecay<2>>(a);
  return -1;
}
int lh_rb23_saveat(int field, int64_t n, const double* y0, const double* p, double t0,
                   const double* tape, const int32_t* n_acc, const double* ts, int64_t k,
                   double* ys, int32_t* save_step, double* save_t, double* states) {
  SaveArgs a{n, y0, p, nullptr, t0, tape, n_acc, ts, k, ys, save_step, save_t, states,
             nullptr};
  if (field == 3) return run_rb_saveat<FRobertson>(a);
  if (field == 8) return run_rb_saveat<FForcedStiff>(a);
  if (field == 2) return run_rb_saveat<FLinearDecay<2>>(a);
  return -1;
}
int lh_rb23_adjoint_saveat(int field, int64_t n, const double* p, double t0,
                           const double* tape, int64_t cap, const int32_t* n_acc,
                           const double* ckpt, int32_t C, const double* ts, int64_t k,
                           const int32_t* save_step, const double* save_t,
                           const double* states, const double* ysbar, const double* ybar,
                           double* mu, double* y0bar, double* pbar) {
  SaveVjpArgs v{n, p, nullptr, n_acc, ts, k, save_step, save_t, states, ysbar, ybar, mu,
                y0bar, pbar};
  AdjArgs a{n, nullptr, p, nullptr, t0, tape, cap, n_acc, ckpt, C, ybar, y0bar, pbar};
  if (field == 3) return run_rb_adj_saveat<FRobertson>(v, a);
  if (field == 8) return run_rb_adj_saveat<FForcedStiff>(v, a);
  if (field == 2) return run_rb_adj_saveat<FLinearDecay<2>>(v, a);
  return -1;
}
}  // extern "C"
namespace {
template <class F>
int r5_fwd_t(const FwdArgs& a) {
  for (int64_t j = 0; j < a.n; ++j) r5_fwd_lane<F>(a, j);
  return 0;
}
template <class F>
int r5_replay_t(const ReplayArgs& a) {
  for (int64_t j = 0; j < a.n; ++j) r5_replay_lane<F>(a, j);
  return 0;
}
template <class F>
int r5_saveat_t(const SaveArgs& a, int dense) {
  for (int64_t j = 0; j < a.n; ++j) {
    if (dense)
      r5_saveat_dense_lane<F>(a, j);
    else
      r5_saveat_lane<F>(a, j);
  }
  return 0;
}
template <class F>
int r5_adjoint_t(AdjArgs a, const SaveVjpArgs* v) {
  if (v) {
    for (int64_t j = 0; j < a.n; ++j) r5_saveat_vjp_lane<F>(*v, j);
  }
  return run_r5_adj<F>(a);
}
}  // namespace
extern "C" {
// Rodas5P fields of the harness: 3 Robertson, 8 the forced stiff field (reads t).
#define LH_R5(CALL, ...)                                          \
  if (field == 3) return CALL<FRobertson>(__VA_ARGS__);           \
  if (field == 8) return CALL<FForcedStiff>(__VA_ARGS__);         \
  return -1;

int lh_r5_fwd(int field, int64_t n, const double* y0, const double* p, double t0, double t1,
              double rtol, double atol, int64_t max_steps, double* y_final, int32_t* n_acc,
              int32_t* n_rej, int32_t* status, double* tape, int64_t cap, double* ckpt,
              int32_t C) {
  FwdArgs a{n, y0, p, nullptr, t0, t1, nullptr, rtol, atol, max_steps, y_final, n_acc, n_rej,
            status, tape, tape ? cap : 0, ckpt, ckpt ? C : 1};
  if (field == 2) return r5_fwd_t<FLinearDecay<1>>(a);
  LH_R5(r5_fwd_t, a)
}
int lh_r5_replay(int field, int64_t n, const double* y0, const double* p, double t0,
                 const double* tape, const int32_t* n_acc, double* y_final) {
  ReplayArgs a{n, y0, p, nullptr, t0, tape, n_acc, y_final, nullptr, 1};
  LH_R5(r5_replay_t, a)
}
// dense: 0 re-step saveat, 1 the continuous extension
int lh_r5_saveat(int field, int dense, int64_t n, const double* y0, const double* p, double t0,
                 const double* tape, const int32_t* n_acc, const double* ts, int64_t k,
                 double* ys, int32_t* save_step, double* save_t, double* states) {
  SaveArgs a{n, y0, p, nullptr, t0, tape, n_acc, ts, k, ys, save_step, save_t, states, nullptr};
  LH_R5(r5_saveat_t, a, dense)
}
// The gradient as gsx_rodas5p_adjoint[_saveat[_dense]] runs it (k = 0: final state only).
int lh_r5_adjoint(int field, int dense, int64_t n, const double* p, double t0,
                  const double* tape, int64_t cap, const int32_t* n_acc, const double* ckpt,
                  int32_t C, const double* ts, int64_t k, const int32_t* save_step,
                  const double* save_t, const double* states, const double* ysbar,
                  const double* ybar, double* mu, double* y0bar, double* pbar) {
  AdjArgs a{n, nullptr, p, nullptr, t0, tape, cap, n_acc, ckpt, C, ybar, y0bar, pbar};
  SaveVjpArgs v{n, p, nullptr, n_acc, ts, k, save_step, save_t, states, ysbar, ybar, mu,
                y0bar, pbar};
  const SaveVjpArgs* vp = nullptr;
  if (k > 0 && dense) {
    a.nsave = k;
    a.save_step = save_step;
    a.ts = ts;
    a.save_t = save_t;
    a.ysbar = ysbar;
  } else if (k > 0) {
    vp = &v;
    a.ybar = y0bar;
    a.nsave = k;
    a.save_step = save_step;
    a.mu = mu;
    a.pbar_in = pbar;
  }
  LH_R5(r5_adjoint_t, a, vp)
}
#undef LH_R5
int lh_tsit5_fwd(int field, int64_t n, const double* y0, const double* p, double t0, double t1,
                 const double* t1_per, double rtol, double atol, int64_t max_steps,
                 double* y_final, int32_t* n_acc, int32_t* n_rej, int32_t* status, double* tape,
                 int64_t cap, double* ckpt, int32_t C) {
  FwdArgs a{n, y0, p, nullptr, t0, t1, t1_per, rtol, atol, max_steps, y_final, n_acc, n_rej,
            status, tape, tape ? cap : 0, ckpt, ckpt ? C : 1};
  if (field == 1) return run_fwd<FLorenz>(a);
  if (field == 6) return run_fwd<FLorenz3>(a);
  if (field == 5) return run_fwd<FVdp>(a);
  if (field == 2) return run_fwd<FLinearDecay<2>>(a);
  return -1;
}
int lh_tsit5_replay(int field, int64_t n, const double* y0, const double* p, double t0,
                    const double* tape, const int32_t* n_acc, double* y_final, double* ckpt,
                    int32_t C) {
  ReplayArgs a{n, y0, p, nullptr, t0, tape, n_acc, y_final, ckpt, ckpt ? C : 1};
  if (field == 1) return run_replay<FLorenz>(a);
  if (field == 6) return run_replay<FLorenz3>(a);
  if (field == 5) return run_replay<FVdp>(a);
  if (field == 2) return run_replay<FLinearDecay<2>>(a);
  return -1;
}
// The checkpoint-interval-2 pair adjoint (the kernel's per-lane body, segments walked from
// each trajectory's last; the kernel walks them in step across the wavefront, which only
// changes which lanes idle).
int lh_tsit5_adjoint_pair(int field, int64_t n, const double* p, double t0, const double* tape,
                          int64_t cap, const int32_t* n_acc, const double* ckpt,
                          const double* ybar, double* y0bar, double* pbar) {
  AdjArgs a{n, nullptr, p, nullptr, t0, tape, cap, n_acc, ckpt, 2, ybar, y0bar, pbar};
  if (field == 1) return run_adj_pair<FLorenz>(a);
  if (field == 6) return run_adj_pair<FLorenz3>(a);
  if (field == 5) return run_adj_pair<FVdp>(a);
  if (field == 2) return run_adj_pair<FLinearDecay<2>>(a);
  return -1;
}
int lh_tsit5_saveat(int field, int64_t n, const double* y0, const double* p, double t0,
                    const double* tape, const int32_t* n_acc, const double* ts, int64_t k,
                    double* ys, int32_t* save_step, double* save_t, double* states) {
  SaveArgs a{n, y0, p, nullptr, t0, tape, n_acc, ts, k, ys, save_step, save_t, states,
             nullptr};
  if (field == 1) return run_saveat<FLorenz>(a);
  if (field == 2) return run_saveat<FLinearDecay<2>>(a);
  return -1;
}
int lh_tsit5_adjoint_saveat(int field, int64_t n, const double* p, double t0,
                            const double* tape, int64_t cap, const int32_t* n_acc,
                            const double* ckpt, int32_t C, const double* ts, int64_t k,
                            const int32_t* save_step, const double* save_t,
                            const double* states, const double* ysbar, const double* ybar,
                            double* mu, double* y0bar, double* pbar) {
  SaveVjpArgs v{n, p, nullptr, n_acc, ts, k, save_step, save_t, states, ysbar, ybar, mu,
                y0bar, pbar};
  AdjArgs a{n, nullptr, p, nullptr, t0, tape, cap, n_acc, ckpt, C, ybar, y0bar, pbar};
  if (field == 1) return run_adj_saveat<FLorenz>(v, a);
  if (field == 2) return run_adj_saveat<FLinearDecay<2>>(v, a);
  return -1;
}
int lh_tsit5_adjoint(int field, int64_t n, const double* p, double t0, const double* tape,
                     int64_t cap, const int32_t* n_acc, const double* ckpt, int32_t C,
                     const double* ybar, double* y0bar, double* pbar) {
  AdjArgs a{n, nullptr, p, nullptr, t0, tape, cap, n_acc, ckpt, C, ybar, y0bar, pbar};
  if (field == 1) return run_adj<FLorenz>(a);
  if (field == 6) return run_adj<FLorenz3>(a);
  if (field == 5) return run_adj<FVdp>(a);
  if (field == 2) return run_adj<FLinearDecay<2>>(a);
  return -1;
}
}
