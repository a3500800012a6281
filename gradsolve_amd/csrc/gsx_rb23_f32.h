// gsx_rb23_f32.h -- the FP32 variant of the Rosenbrock23 lanes.
//
// The field evaluations and the stage vectors (k1, k2, the stage inputs and the third stage's
// right-hand side) run in FP32 on a compensated state (s, lo) with value s + lo: each accepted
// increment h k2 is added with Kahan's compensated summation. W = I - d h J, its LU
// factorisation with partial pivoting and the triangular solves stay FP64 (J evaluated in FP64
// at s + lo; each solve takes its FP32 right-hand side exactly and rounds its result to FP32),
// as do the error estimate (h/6)(k1 - 2 k2 + k3), the WRMS norm, the controller, time and the
// tape. Checkpoints hold the pair (s, lo) as two FP32 rows per component, in the bytes an FP64
// checkpoint takes. The adjoint is the VJP of this FP32 step: FP32 through the field, FP64
// through the solves and J, with compensated cotangent and parameter-gradient accumulators.
#pragma once

#include "gsx_core.h"
#include "gsx_f32.h"
#include "gsx_rb23.h"

namespace GSX_NS {

// h * k with contraction off: the product is rounded on its own, so the forward (where the
// increment also forms the trial state) and the replay (where it only feeds the compensated
// sum) add the same increment and reproduce each other's states bit for bit.
GSX_HD float mul_rn(float a, float b) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  return a * b;
}

// W = I - d h J at the compensated state, factorised in FP64.
template <class F>
GSX_HD void rb23_w_f32(const float* y, const float* lo, double h, const double* pd,
                       const double* aux, double (&W)[F::D * F::D], int (&piv)[F::D]) {
  constexpr int D = F::D;
  double yd[D];
  GSX_UNROLL for (int d = 0; d < D; ++d) yd[d] = static_cast<double>(y[d]) +
                                                 static_cast<double>(lo[d]);
  rb23_build_w<F>(0.0, yd, h, pd, aux, W, piv);
}

// x = W^-1 b for an FP32 right-hand side, solved in FP64 and rounded to FP32.
template <int D>
GSX_HD void lu_solve_f32(const double (&W)[D * D], const int (&piv)[D], const float* b,
                         float* x) {
  double bd[D], xd[D];
  GSX_UNROLL for (int i = 0; i < D; ++i) bd[i] = static_cast<double>(b[i]);
  lu_solve<D>(W, piv, bd, xd);
  GSX_UNROLL for (int i = 0; i < D; ++i) x[i] = static_cast<float>(xd[i]);
}

// The two stages every use of the step shares (forward, replay, VJP): F0 = f(y),
// k1 = W^-1 F0, u = y + (h/2) k1, F1 = f(u), x2 = W^-1 (F1 - k1), k2 = x2 + k1.
template <class F>
GSX_HD void rb23_stages_f32(const float* y, float hf, const float* pf,
                            const double (&W)[F::D * F::D], const int (&piv)[F::D],
                            const double* aux, float* F0, float* k1, float* u, float* F1,
                            float* x2, float* k2) {
  constexpr int D = F::D;
  F::f(0.0f, y, pf, F0, aux);
  lu_solve_f32<D>(W, piv, F0, k1);
  const float hh = 0.5f * hf;
  GSX_UNROLL for (int i = 0; i < D; ++i) u[i] = y[i] + hh * k1[i];
  F::f(0.0f, u, pf, F1, aux);
  float r[D];
  GSX_UNROLL for (int i = 0; i < D; ++i) r[i] = F1[i] - k1[i];
  lu_solve_f32<D>(W, piv, r, x2);
  GSX_UNROLL for (int i = 0; i < D; ++i) k2[i] = x2[i] + k1[i];
}

// One FP32 replay step (rosenbrock23_step, gradsolve/solvers/rosenbrock23_step.py:47-66):
// (y, lo) += h k2.
template <class F>
GSX_HD void rb23_advance_f32(float* y, float* lo, double h, const float* pf, const double* pd,
                             const double* aux) {
  constexpr int D = F::D;
  double W[D * D];
  int piv[D];
  rb23_w_f32<F>(y, lo, h, pd, aux, W, piv);
  const float hf = static_cast<float>(h);
  float F0[D], k1[D], u[D], F1[D], x2[D], k2[D];
  rb23_stages_f32<F>(y, hf, pf, W, piv, aux, F0, k1, u, F1, x2, k2);
  GSX_UNROLL for (int i = 0; i < D; ++i) kahan_add(y[i], lo[i], mul_rn(hf, k2[i]));
}

// Forward + record with the tape and checkpoint contract of rb23_fwd_lane.
template <class F>
GSX_HD void rb23_fwd_lane_f32(const FwdArgs& a, int64_t j) {
  static_assert(F::autonomous, "the FP32 variant is built for autonomous fields");
  constexpr int D = F::D, P = F::P;
  const int64_t n = a.n;
  float y[D], lo[D], pf[P];
  double pd[P];
  GSX_UNROLL for (int d = 0; d < D; ++d) {
    const double v = a.y0[d * n + j];
    y[d] = static_cast<float>(v);
    lo[d] = static_cast<float>(v - static_cast<double>(y[d]));
  }
  GSX_UNROLL for (int q = 0; q < P; ++q) {
    pd[q] = a.p[q * n + j];
    pf[q] = static_cast<float>(pd[q]);
  }
  double t = a.t0;
  const double t1 = a.t1_per ? a.t1_per[j] : a.t1;
  double dt = (t1 - t) / 100.0;
  int32_t acc = 0, rej = 0, status = kReached;
  const bool rec = a.tape != nullptr;
  const bool ck = a.ckpt != nullptr;
  int32_t ck_left = a.ckpt_interval;
  double* tp = rec ? a.tape + j : nullptr;
  float* cp = ck ? reinterpret_cast<float*>(a.ckpt) + j : nullptr;
  const int64_t cstride = static_cast<int64_t>(2 * D) * n;
  if (ck) ckpt_put_f32<F>(cp, n, y, lo);
  const float e32 = static_cast<float>(rb_e32());
  for (int64_t it = 0; it < a.max_steps; ++it) {
    if (t >= t1) break;
    const double h = py_min(dt, t1 - t);
    const float hf = static_cast<float>(h);
    double W[D * D];
    int piv[D];
    rb23_w_f32<F>(y, lo, h, pd, a.aux, W, piv);
    float F0[D], k1[D], u[D], F1[D], x2[D], k2[D], inc[D], yn[D], F2[D], r3[D];
    rb23_stages_f32<F>(y, hf, pf, W, piv, a.aux, F0, k1, u, F1, x2, k2);
    GSX_UNROLL for (int i = 0; i < D; ++i) {
      inc[i] = mul_rn(hf, k2[i]);
      yn[i] = y[i] + inc[i];
    }
    F::f(0.0f, yn, pf, F2, a.aux);
    GSX_UNROLL for (int i = 0; i < D; ++i)
      r3[i] = F2[i] - e32 * (k2[i] - F1[i]) - 2.0f * (k1[i] - F0[i]);
    double r3d[D], k3[D], e[D], yd[D], ynd[D];
    GSX_UNROLL for (int i = 0; i < D; ++i) r3d[i] = static_cast<double>(r3[i]);
    lu_solve<D>(W, piv, r3d, k3);
    const double h6 = h * (1.0 / 6.0);
    GSX_UNROLL for (int i = 0; i < D; ++i) {
      e[i] = h6 * (static_cast<double>(k1[i]) - 2.0 * static_cast<double>(k2[i]) + k3[i]);
      yd[i] = static_cast<double>(y[i]) + static_cast<double>(lo[i]);
      ynd[i] = yd[i] + static_cast<double>(inc[i]);
    }
    const double err = wrms<D>(e, yd, ynd, a.rtol, a.atol);
    if (err != err) {
      rej += static_cast<int32_t>(a.max_steps - it);
      break;
    }
    if (err <= 1.0) {
      if (rec && acc >= a.tape_cap) {
        status = kTapeFull;
        break;
      }
      t = t + h;
      GSX_UNROLL for (int d = 0; d < D; ++d) kahan_add(y[d], lo[d], inc[d]);
      if (rec) {
        *tp = h;
        tp += n;
      }
      ++acc;
      if (ck && --ck_left == 0) {
        cp += cstride;
        ckpt_put_f32<F>(cp, n, y, lo);
        ck_left = a.ckpt_interval;
      }
    } else {
      ++rej;
    }
    const double fac = kSafety * ctrl_pow_rb23(py_max(err, kErrFloor));
    dt = h * py_min(py_max(fac, kFacMin), kFacMax);
  }
  if (status == kReached && !(t >= t1)) status = kExhausted;
  GSX_UNROLL for (int d = 0; d < D; ++d)
    a.y_final[d * n + j] = static_cast<double>(y[d]) + static_cast<double>(lo[d]);
  // the counts and the status are optional outputs of a forward without a tape
  if (a.n_acc) a.n_acc[j] = acc;
  if (a.n_rej) a.n_rej[j] = rej;
  if (a.status) a.status[j] = status;
}

template <class F>
GSX_HD void rb23_replay_lane_f32(const ReplayArgs& a, int64_t j) {
  constexpr int D = F::D, P = F::P;
  const int64_t n = a.n;
  float y[D], lo[D], pf[P];
  double pd[P];
  GSX_UNROLL for (int d = 0; d < D; ++d) {
    const double v = a.y0[d * n + j];
    y[d] = static_cast<float>(v);
    lo[d] = static_cast<float>(v - static_cast<double>(y[d]));
  }
  GSX_UNROLL for (int q = 0; q < P; ++q) {
    pd[q] = a.p[q * n + j];
    pf[q] = static_cast<float>(pd[q]);
  }
  const int32_t na = a.n_acc[j];
  const bool ck = a.ckpt != nullptr;
  if (ck) ckpt_store_f32<F>(a.ckpt, 0, n, j, y, lo);
  int32_t ck_left = a.ckpt_interval;
  for (int32_t s = 0; s < na; ++s) {
    rb23_advance_f32<F>(y, lo, a.tape[static_cast<int64_t>(s) * n + j], pf, pd, a.aux);
    if (ck && --ck_left == 0) {
      ckpt_store_f32<F>(a.ckpt, (s + 1) / a.ckpt_interval, n, j, y, lo);
      ck_left = a.ckpt_interval;
    }
  }
  GSX_UNROLL for (int d = 0; d < D; ++d)
    a.y_final[d * n + j] = static_cast<double>(y[d]) + static_cast<double>(lo[d]);
}

// Adjoint of one FP32 replay step, the map (y, lo) -> (y, lo) + h k2(y): lam += (d h k2 /
// dy)^T lam and pb += (d h k2 / dp)^T lam, both compensated. Through x = W^-1 r:
// rbar = W^-T xbar and Wbar -= rbar x^T; through W = I - d h J(y, p): Jbar = -d h Wbar,
// contracted with dJ/dy and dJ/dp in FP64.
template <class F>
GSX_HD void rb23_step_vjp_f32(const float* y, const float* lo, double h, const float* pf,
                              const double* pd, float* lam, float* lam_lo, float* pb,
                              float* pb_lo, const double* aux) {
  constexpr int D = F::D, P = F::P;
  double W[D * D];
  int piv[D];
  rb23_w_f32<F>(y, lo, h, pd, aux, W, piv);
  const float hf = static_cast<float>(h);
  float F0[D], k1[D], u[D], F1[D], x2[D], k2[D];
  rb23_stages_f32<F>(y, hf, pf, W, piv, aux, F0, k1, u, F1, x2, k2);
  float dlam[D], dp[P], k1b[D];
  double Wb[D * D], xb[D], rb[D];
  // Accumulators start at -0.0, the exact additive identity, so their first adds fold away.
  GSX_UNROLL for (int i = 0; i < D; ++i) dlam[i] = -0.0f;
  GSX_UNROLL for (int q = 0; q < P; ++q) dp[q] = -0.0f;
  // k2 = x2 + k1 with increment hf k2: x2bar = k1bar = hf lam
  GSX_UNROLL for (int i = 0; i < D; ++i) {
    k1b[i] = hf * lam[i];
    xb[i] = static_cast<double>(k1b[i]);
  }
  lu_solve_t<D>(W, piv, xb, rb);  // rbar of r = F1 - k1
  float r2b[D];
  GSX_UNROLL for (int i = 0; i < D; ++i) {
    GSX_UNROLL for (int c = 0; c < D; ++c) Wb[i * D + c] = -rb[i] * static_cast<double>(x2[c]);
    r2b[i] = static_cast<float>(rb[i]);
    k1b[i] -= r2b[i];
  }
  float ub[D];
  GSX_UNROLL for (int i = 0; i < D; ++i) ub[i] = -0.0f;
  F::vjp(0.0f, u, pf, r2b, ub, dp, aux);  // F1 = f(u)
  const float hh = 0.5f * hf;
  GSX_UNROLL for (int i = 0; i < D; ++i) {
    dlam[i] += ub[i];
    k1b[i] += hh * ub[i];
  }
  GSX_UNROLL for (int i = 0; i < D; ++i) xb[i] = static_cast<double>(k1b[i]);
  lu_solve_t<D>(W, piv, xb, rb);  // F0bar, through k1 = W^-1 F0
  float F0b[D];
  GSX_UNROLL for (int i = 0; i < D; ++i) {
    GSX_UNROLL for (int c = 0; c < D; ++c) Wb[i * D + c] -= rb[i] * static_cast<double>(k1[c]);
    F0b[i] = static_cast<float>(rb[i]);
  }
  F::vjp(0.0f, y, pf, F0b, dlam, dp, aux);  // F0 = f(y)
  const double dh = h * rb_d();
  double Jb[D * D], yd[D], ybd[D], pbd[P];
  GSX_UNROLL for (int i = 0; i < D * D; ++i) Jb[i] = -dh * Wb[i];
  GSX_UNROLL for (int d = 0; d < D; ++d) {
    yd[d] = static_cast<double>(y[d]) + static_cast<double>(lo[d]);
    ybd[d] = -0.0;
  }
  GSX_UNROLL for (int q = 0; q < P; ++q) pbd[q] = -0.0;
  F::jac_vjp(0.0, yd, pd, Jb, ybd, pbd, aux);
  GSX_UNROLL for (int i = 0; i < D; ++i) {
    dlam[i] += static_cast<float>(ybd[i]);
    kahan_add(lam[i], lam_lo[i], dlam[i]);
  }
  GSX_UNROLL for (int q = 0; q < P; ++q) {
    dp[q] += static_cast<float>(pbd[q]);
    kahan_add(pb[q], pb_lo[q], dp[q]);
  }
}

// The segment buffer is the FP64 lane's: each double cell holds one component's compensated
// pair (s, lo) as two floats, so a slot takes D doubles per lane.
template <class F>
GSX_HD void rb23_adj_lane_f32(const AdjArgs& a, int64_t j, double* seg, int64_t sstride,
                              int32_t mtop = 0) {
  constexpr int D = F::D, P = F::P;
  const int64_t n = a.n;
  float pf[P], pb[P], pb_lo[P], lam[D], lam_lo[D];
  double pd[P];
  GSX_UNROLL for (int q = 0; q < P; ++q) {
    pd[q] = a.p[q * n + j];
    pf[q] = static_cast<float>(pd[q]);
    pb[q] = 0.0f;
    pb_lo[q] = 0.0f;
  }
  GSX_UNROLL for (int d = 0; d < D; ++d) {
    const double v = a.ybar[d * n + j];
    lam[d] = static_cast<float>(v);
    lam_lo[d] = static_cast<float>(v - static_cast<double>(lam[d]));
  }
  const int32_t na = a.n_acc[j];
  const int32_t C = a.ckpt_interval;
  const int32_t nseg = (na + C - 1) / C;
  for (int32_t m = (mtop > nseg ? mtop : nseg) - 1; m >= 0; --m) {
    if (m >= nseg) continue;  // the wavefront's walk has not reached this lane
    const int32_t s0 = m * C;
    const int32_t s1 = (s0 + C < na) ? s0 + C : na;
    float y[D], lo[D];
    const float* base =
        reinterpret_cast<const float*>(a.ckpt) + static_cast<int64_t>(m) * 2 * D * n + j;
    GSX_UNROLL for (int d = 0; d < D; ++d) {
      y[d] = base[d * n];
      lo[d] = base[(D + d) * n];
    }
    for (int32_t s = s0; s < s1; ++s) {
      double* slot = seg + static_cast<int64_t>(s - s0) * D * sstride;
      GSX_UNROLL for (int d = 0; d < D; ++d) {
        float* cell = reinterpret_cast<float*>(slot + d * sstride);
        cell[0] = y[d];
        cell[1] = lo[d];
      }
      if (s + 1 < s1)
        rb23_advance_f32<F>(y, lo, a.tape[static_cast<int64_t>(s) * n + j], pf, pd, a.aux);
    }
    for (int32_t s = s1 - 1; s >= s0; --s) {
      const double* slot = seg + static_cast<int64_t>(s - s0) * D * sstride;
      GSX_UNROLL for (int d = 0; d < D; ++d) {
        const float* cell = reinterpret_cast<const float*>(slot + d * sstride);
        y[d] = cell[0];
        lo[d] = cell[1];
      }
      rb23_step_vjp_f32<F>(y, lo, a.tape[static_cast<int64_t>(s) * n + j], pf, pd, lam,
                           lam_lo, pb, pb_lo, a.aux);
    }
  }
  GSX_UNROLL for (int d = 0; d < D; ++d)
    a.y0bar[d * n + j] = static_cast<double>(lam[d]) + static_cast<double>(lam_lo[d]);
  GSX_UNROLL for (int q = 0; q < P; ++q)
    a.pbar[q * n + j] = static_cast<double>(pb[q]) + static_cast<double>(pb_lo[q]);
}

}  // namespace GSX_NS
