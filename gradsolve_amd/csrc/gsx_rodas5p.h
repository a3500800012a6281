// gsx_rodas5p.h -- per-lane Rodas5P (Steinebach 2023: order-5(4) Rosenbrock-Wanner pair, 8
// stages, order-4 continuous extension) for stiff fields with an analytic Jacobian: forward +
// record, replay, adjoint, and the dense output of the continuous extension with its adjoint.
// gradsolve's Rodas5P lane is rodas5p_step.py / rodas5p_replay.py, recorded by
// record_jax.py:83-132 (the recorder's status 2 at the underflow floor and its FAC_MIN shrink
// after a non-finite trial are kept). The tableau (gsx_rodas5p_tab.h) is in alpha/Gamma form:
//   J = df/dy(t, y), f_t = df/dt(t, y) (fields that read t; rodas5p_step.py:119-146);
//   W = I - (h gamma) J;  u_i = y + sum_{j<i} alpha_ij k_j,  a_i = sum_{j<i} Gamma_ij k_j,
//   k_i = W \ (h f(t + c_i h, u_i) + h J a_i + h^2 g_i f_t),  g_i = sum_j Gamma_ij;
//   y_next = y + sum m_i k_i;  error = sum (m_i - mhat_i) k_i.
// One LU of W per step (in registers, partial pivoting), reused by the 8 stage solves.
#pragma once

#include "gsx_core.h"
#include "gsx_rb23.h"
#include "gsx_rodas5p_tab.h"

namespace GSX_NS {

namespace r5t = ::gsx_rodas5p_tab;

// The stages of one step at (y, h): k_i, J and the LU of W.
template <class F>
struct R5Stages {
  static constexpr int D = F::D;
  double k[8][D];
  double J[D * D];
  double W[D * D];
  int piv[D];
};

// Stage input u_i and J-coupling sum a_i from k_0..k_{i-1} (i a compile-time index after
// unrolling, so k stays in registers).
template <class F>
GSX_HD void r5_stage_input(const double* y, const double (&k)[8][F::D], int i, double* u,
                           double* acc) {
  constexpr int D = F::D;
  GSX_UNROLL for (int d = 0; d < D; ++d) {
    u[d] = y[d];
    acc[d] = 0.0;
  }
  GSX_UNROLL for (int j = 0; j < 8; ++j) {
    if (j < i) {
      GSX_UNROLL for (int d = 0; d < D; ++d) {
        u[d] = u[d] + r5t::kAlpha[i * 8 + j] * k[j][d];
        acc[d] = acc[d] + r5t::kGammaOff[i * 8 + j] * k[j][d];
      }
    }
  }
}

template <class F>
GSX_HD void r5_stages(double t, const double* y, double h, const double* p, const double* aux,
                      R5Stages<F>& w) {
  constexpr int D = F::D;
  F::jac(t, y, p, w.J, aux);
  const double hg = h * r5t::kGamma;
  GSX_UNROLL for (int a = 0; a < D; ++a) {
    GSX_UNROLL for (int b = 0; b < D; ++b)
      w.W[a * D + b] = (a == b ? 1.0 : 0.0) - hg * w.J[a * D + b];
  }
  lu_factor<D>(w.W, w.piv);
  double ft[D];
  if constexpr (!F::autonomous) F::dfdt(t, y, p, ft, aux);
  GSX_UNROLL for (int i = 0; i < 8; ++i) {
    double u[D], acc[D], fv[D], rhs[D];
    r5_stage_input<F>(y, w.k, i, u, acc);
    F::f(t + r5t::kC[i] * h, u, p, fv, aux);
    GSX_UNROLL for (int r = 0; r < D; ++r) {
      double ja = 0.0;
      GSX_UNROLL for (int c = 0; c < D; ++c) ja = ja + w.J[r * D + c] * acc[c];
      rhs[r] = h * fv[r] + h * ja;
      if constexpr (!F::autonomous) rhs[r] = rhs[r] + (h * h * r5t::kGammaSum[i]) * ft[r];
    }
    lu_solve<D>(w.W, w.piv, rhs, w.k[i]);
  }
}

template <class F>
GSX_HD void r5_combine(const double* y, const R5Stages<F>& w, double* ynew) {
  constexpr int D = F::D;
  GSX_UNROLL for (int d = 0; d < D; ++d) ynew[d] = y[d];
  GSX_UNROLL for (int i = 0; i < 8; ++i) {
    GSX_UNROLL for (int d = 0; d < D; ++d) ynew[d] = ynew[d] + r5t::kM[i] * w.k[i][d];
  }
}

// Value-only step rodas5p_advance (rodas5p_step.py:149-155).
template <class F>
GSX_HD void r5_advance(double t, double* y, double h, const double* p, const double* aux) {
  R5Stages<F> w;
  r5_stages<F>(t, y, h, p, aux, w);
  double yn[F::D];
  r5_combine<F>(y, w, yn);
  GSX_UNROLL for (int d = 0; d < F::D; ++d) y[d] = yn[d];
}

// One trial step: y_next and the error measure of tsit5_err (the same WRMS norm; the
// controller exponent -1/5 is Tsit5's too). finite: y_next, the error vector and the measure
// are all finite.
template <class F>
GSX_HD double r5_trial(double t, const double* y, double h, const double* p, const double* aux,
                       double* ynew, double rtol, double atol, bool& finite) {
  constexpr int D = F::D;
  R5Stages<F> w;
  r5_stages<F>(t, y, h, p, aux, w);
  double e[D];
  GSX_UNROLL for (int d = 0; d < D; ++d) {
    ynew[d] = y[d];
    e[d] = 0.0;
  }
  GSX_UNROLL for (int i = 0; i < 8; ++i) {
    GSX_UNROLL for (int d = 0; d < D; ++d) {
      ynew[d] = ynew[d] + r5t::kM[i] * w.k[i][d];
      e[d] = e[d] + r5t::kE[i] * w.k[i][d];
    }
  }
  const double err = tsit5_err<D>(e, y, ynew, rtol, atol);
  bool ok = gsx_finite(err);
  GSX_UNROLL for (int d = 0; d < D; ++d) ok = ok && gsx_finite(ynew[d]) && gsx_finite(e[d]);
  finite = ok;
  return err;
}

// The continuous extension's vectors D_r = sum_i hd[r][i] k_i (rodas5p_step.py:198-226) and
// its value (1 - th) y_l + th (y_r + (1 - th) (D1 + th (D2 + th D3))) (rodas5p_step.py:171-195).
template <class F>
GSX_HD void r5_dense_vectors(const R5Stages<F>& w, double (&dv)[3][F::D]) {
  constexpr int D = F::D;
  GSX_UNROLL for (int r = 0; r < 3; ++r) {
    GSX_UNROLL for (int d = 0; d < D; ++d) dv[r][d] = 0.0;
    GSX_UNROLL for (int i = 0; i < 8; ++i) {
      GSX_UNROLL for (int d = 0; d < D; ++d) dv[r][d] = dv[r][d] + r5t::kHD[r * 8 + i] * w.k[i][d];
    }
  }
}

template <class F>
GSX_HD void r5_dense_eval(const double* yl, const double* yr, const double (&dv)[3][F::D],
                          double th, double* out) {
  const double om = 1.0 - th;
  GSX_UNROLL for (int d = 0; d < F::D; ++d)
    out[d] = om * yl[d] + th * (yr[d] + om * (dv[0][d] + th * (dv[1][d] + th * dv[2][d])));
}

// Adjoint of r5_advance at (t, y, h): lam <- (dy_next/dy)^T lam (+ the dense vectors'
// cotangents dbar, when given, carried to the stages), pb += (d/dp)^T. Reverse over the
// stages: each solve k_i = W^-1 rhs_i sends W^-T kbar_i to rhs_i and -W^-T kbar_i k_i^T to W;
// rhs_i sends its cotangent to f at u_i (then to y and the earlier k via alpha), to J (outer
// product with a_i), to the earlier k via Gamma J^T, and h^2 g_i of it to f_t; W = I - h gamma J
// sends -h gamma Wbar to J; J's cotangent reaches y and p through jac_vjp, f_t's through
// dfdt_vjp.
template <class F>
GSX_HD void r5_step_vjp(double t, const double* y, double h, const double* p, double* lam,
                        double* pb, const double* aux, const double (*dbar)[F::D] = nullptr) {
  constexpr int D = F::D;
  R5Stages<F> w;
  r5_stages<F>(t, y, h, p, aux, w);
  double kb[8][D], yb[D], Wb[D * D], Jb[D * D], ftb[D];
  GSX_UNROLL for (int d = 0; d < D; ++d) yb[d] = lam[d];
  GSX_UNROLL for (int i = 0; i < 8; ++i) {
    GSX_UNROLL for (int d = 0; d < D; ++d) {
      kb[i][d] = r5t::kM[i] * lam[d];
      if (dbar) {
        GSX_UNROLL for (int r = 0; r < 3; ++r) kb[i][d] = kb[i][d] + r5t::kHD[r * 8 + i] * dbar[r][d];
      }
    }
  }
  GSX_UNROLL for (int x = 0; x < D * D; ++x) {
    Wb[x] = 0.0;
    Jb[x] = 0.0;
  }
  GSX_UNROLL for (int d = 0; d < D; ++d) ftb[d] = 0.0;
  GSX_UNROLL for (int i = 7; i >= 0; --i) {
    double u[D], acc[D], rb[D], fb[D], ub[D], accb[D];
    r5_stage_input<F>(y, w.k, i, u, acc);
    lu_solve_t<D>(w.W, w.piv, kb[i], rb);
    if constexpr (!F::autonomous) {
      GSX_UNROLL for (int d = 0; d < D; ++d) ftb[d] = ftb[d] + (h * h * r5t::kGammaSum[i]) * rb[d];
    }
    GSX_UNROLL for (int a = 0; a < D; ++a) {
      GSX_UNROLL for (int b = 0; b < D; ++b) Wb[a * D + b] = Wb[a * D + b] - rb[a] * w.k[i][b];
    }
    GSX_UNROLL for (int d = 0; d < D; ++d) {
      fb[d] = h * rb[d];
      ub[d] = -0.0;  // x + (-0.0) == x: the first add folds
    }
    F::vjp(t + r5t::kC[i] * h, u, p, fb, ub, pb, aux);
    GSX_UNROLL for (int a = 0; a < D; ++a) {
      GSX_UNROLL for (int b = 0; b < D; ++b) Jb[a * D + b] = Jb[a * D + b] + h * rb[a] * acc[b];
    }
    GSX_UNROLL for (int c = 0; c < D; ++c) {
      double s = 0.0;
      GSX_UNROLL for (int r = 0; r < D; ++r) s = s + w.J[r * D + c] * rb[r];
      accb[c] = h * s;
    }
    GSX_UNROLL for (int d = 0; d < D; ++d) yb[d] = yb[d] + ub[d];
    GSX_UNROLL for (int j = 0; j < 8; ++j) {
      if (j < i) {
        GSX_UNROLL for (int d = 0; d < D; ++d)
          kb[j][d] = kb[j][d] + r5t::kAlpha[i * 8 + j] * ub[d] +
                     r5t::kGammaOff[i * 8 + j] * accb[d];
      }
    }
  }
  const double hg = h * r5t::kGamma;
  GSX_UNROLL for (int x = 0; x < D * D; ++x) Jb[x] = Jb[x] - hg * Wb[x];
  F::jac_vjp(t, y, p, Jb, yb, pb, aux);
  if constexpr (!F::autonomous) F::dfdt_vjp(t, y, p, ftb, yb, pb, aux);
  GSX_UNROLL for (int d = 0; d < D; ++d) lam[d] = yb[d];
}

// Rodas5P as the step policy of the re-step saveat lanes (saveat_lane / saveat_vjp_lane).
struct MethodRodas5P {
  template <class F>
  GSX_HD static void advance(double t, double* y, double h, const double* p, const double* aux) {
    r5_advance<F>(t, y, h, p, aux);
  }
  template <class F>
  GSX_HD static void step_vjp(double t, const double* y, double h, const double* p,
                              double* lam, double* pb, const double* aux) {
    r5_step_vjp<F>(t, y, h, p, lam, pb, aux);
  }
};

// Forward + record (the Tsit5 lane's tape / checkpoint contract) with the recorder semantics
// of record_jax.py:83-132 for Rodas5P: a step h with t + h == t or h <= floor (4 eps
// max(|t0|, |t1 - t0|)) ends the trajectory with status 2; a non-finite trial counts as a
// rejection and shrinks h by FAC_MIN without a controller update; max_steps counts attempts.
template <class F>
GSX_HD void r5_fwd_lane(const FwdArgs& a, int64_t j) {
  constexpr int D = F::D, P = F::P;
  const int64_t n = a.n;
  double y[D], p[P], ynew[D];
  GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = a.y0[d * n + j];
  GSX_UNROLL for (int q = 0; q < P; ++q) p[q] = a.p[q * n + j];
  double t = a.t0;
  const double t1 = a.t1_per ? a.t1_per[j] : a.t1;
  double dt = (t1 - t) / 100.0;
  const double floor = 4.0 * 2.220446049250313e-16 * py_max(std::fabs(a.t0), std::fabs(t1 - a.t0));
  int32_t acc = 0, rej = 0, status = kReached;
  int64_t attempts = 0;
  const bool rec = a.tape != nullptr;
  const bool ck = a.ckpt != nullptr;
  int32_t ck_left = a.ckpt_interval;
  double* tp = rec ? a.tape + j : nullptr;
  double* cp = ck ? a.ckpt + j : nullptr;
  const int64_t cstride = static_cast<int64_t>(ckpt_rows<F>()) * n;
  if (ck) ckpt_put<F>(cp, n, y, t);
  while (t < t1 && status == kReached) {
    const double h = py_min(dt, t1 - t);
    if (t + h == t || h <= floor) {
      status = kDegenerate;
      break;
    }
    bool finite = true;
    const double err = r5_trial<F>(t, y, h, p, a.aux, ynew, a.rtol, a.atol, finite);
    ++attempts;
    if (finite && err <= 1.0) {
      if (rec && acc >= a.tape_cap) {
        status = kTapeFull;
        break;
      }
      t = t + h;
      GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = ynew[d];
      if (rec) {
        *tp = h;
        tp += n;
      }
      ++acc;
      if (ck && --ck_left == 0) {
        cp += cstride;
        ckpt_put<F>(cp, n, y, t);
        ck_left = a.ckpt_interval;
      }
      dt = tsit5_next_dt(h, err);
    } else {
      ++rej;
      dt = finite ? tsit5_next_dt(h, err) : h * kFacMin;
    }
    if (attempts >= a.max_steps && t < t1) status = kExhausted;
  }
  GSX_UNROLL for (int d = 0; d < D; ++d) a.y_final[d * n + j] = y[d];
  // the counts and the status are optional outputs of a forward without a tape
  if (a.n_acc) a.n_acc[j] = acc;
  if (a.n_rej) a.n_rej[j] = rej;
  if (a.status) a.status[j] = status;
}

template <class F>
GSX_HD void r5_replay_lane(const ReplayArgs& a, int64_t j) {
  constexpr int D = F::D, P = F::P;
  const int64_t n = a.n;
  double y[D], p[P];
  GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = a.y0[d * n + j];
  GSX_UNROLL for (int q = 0; q < P; ++q) p[q] = a.p[q * n + j];
  double t = a.t0;
  const int32_t na = a.n_acc[j];
  const bool ck = a.ckpt != nullptr;
  if (ck) ckpt_store<F>(a.ckpt, 0, n, j, y, t);
  int32_t ck_left = a.ckpt_interval;
  for (int32_t s = 0; s < na; ++s) {
    const double h = a.tape[static_cast<int64_t>(s) * n + j];
    r5_advance<F>(t, y, h, p, a.aux);
    t = t + h;
    if (ck && --ck_left == 0) {
      ckpt_store<F>(a.ckpt, (s + 1) / a.ckpt_interval, n, j, y, t);
      ck_left = a.ckpt_interval;
    }
  }
  GSX_UNROLL for (int d = 0; d < D; ++d) a.y_final[d * n + j] = y[d];
}

// The dense part of the cotangent at step s: saves q with save_step == s, walked downwards
// from the cursor qs. Each sends th * g to y_{s+1} (into lam, before the step's VJP),
// th (1 - th) th^r * g to the step's dense vector D_{r+1} (dbar), and (1 - th) * g to y_s
// (after the VJP; returned in post).
template <int D>
GSX_HD bool r5_dense_cotangents(const AdjArgs& a, int64_t j, int32_t s, double h, int32_t& qs,
                                double* lam, double (&dbar)[3][D], double* post) {
  const int64_t n = a.n;
  bool any = false;
  GSX_UNROLL for (int d = 0; d < D; ++d) {
    post[d] = 0.0;
    GSX_UNROLL for (int r = 0; r < 3; ++r) dbar[r][d] = 0.0;
  }
  while (qs >= 0) {
    const int64_t q = qs;
    const int32_t st = a.save_step[q * n + j];
    if (st < s) break;
    if (st == s) {
      any = true;
      const double th = (a.ts[q] - a.save_t[q * n + j]) / h, om = 1.0 - th;
      GSX_UNROLL for (int d = 0; d < D; ++d) {
        const double g = a.ysbar[(q * D + d) * n + j];
        lam[d] += th * g;
        dbar[0][d] += th * om * g;
        dbar[1][d] += th * om * th * g;
        dbar[2][d] += th * om * th * th * g;
        post[d] += om * g;
      }
    }
    --qs;
  }
  return any;
}

// Adjoint of the replay, one segment of C steps per iteration (re-integrated from its
// checkpoint into `seg`, then reversed). With a.nsave and a.mu: re-step saveat cotangents (as
// the Tsit5 lane). With a.ysbar: dense-output cotangents (save_step / save_t / ts as the dense
// saveat lane wrote them; saves at the final time and at or before t0 are handled here).
template <class F>
GSX_HD void r5_adj_lane(const AdjArgs& a, int64_t j, double* seg, int64_t sstride,
                        int32_t mtop = 0) {
  constexpr int D = F::D, P = F::P;
  constexpr int DCK = ckpt_rows<F>();
  const int64_t n = a.n;
  const bool dense = a.ysbar != nullptr;
  double p[P], pb[P], lam[D];
  GSX_UNROLL for (int q = 0; q < P; ++q) {
    p[q] = a.p[q * n + j];
    pb[q] = a.pbar_in ? a.pbar_in[q * n + j] : 0.0;
  }
  GSX_UNROLL for (int d = 0; d < D; ++d) lam[d] = a.ybar ? a.ybar[d * n + j] : 0.0;
  const int32_t na = a.n_acc[j];
  int32_t qs = static_cast<int32_t>(a.nsave) - 1;  // save cursor, walked downwards
  if (dense) {
    // saves at or past the final time are y_final itself
    while (qs >= 0 && a.save_step[static_cast<int64_t>(qs) * n + j] >= na) {
      GSX_UNROLL for (int d = 0; d < D; ++d)
        lam[d] += a.ysbar[(static_cast<int64_t>(qs) * D + d) * n + j];
      --qs;
    }
  }
  const int32_t C = a.ckpt_interval;
  const int32_t nseg = (na + C - 1) / C;
  for (int32_t m = (mtop > nseg ? mtop : nseg) - 1; m >= 0; --m) {
    if (m >= nseg) continue;  // the wavefront's walk has not reached this lane
    const int32_t s0 = m * C;
    const int32_t s1 = (s0 + C < na) ? s0 + C : na;
    double y[D];
    double t = 0.0;
    const double* base = a.ckpt + static_cast<int64_t>(m) * DCK * n + j;
    GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = base[d * n];
    if constexpr (!F::autonomous) t = base[D * n];
    for (int32_t s = s0; s < s1; ++s) {
      double* slot = seg + static_cast<int64_t>(s - s0) * DCK * sstride;
      GSX_UNROLL for (int d = 0; d < D; ++d) slot[d * sstride] = y[d];
      if constexpr (!F::autonomous) slot[D * sstride] = t;
      if (s + 1 < s1) {
        const double h = a.tape[static_cast<int64_t>(s) * n + j];
        r5_advance<F>(t, y, h, p, a.aux);
        t = t + h;
      }
    }
    for (int32_t s = s1 - 1; s >= s0; --s) {
      const double* slot = seg + static_cast<int64_t>(s - s0) * DCK * sstride;
      GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = slot[d * sstride];
      if constexpr (!F::autonomous) t = slot[D * sstride];
      const double h = a.tape[static_cast<int64_t>(s) * n + j];
      if (dense) {
        double dbar[3][D], post[D];
        const bool any = r5_dense_cotangents<D>(a, j, s, h, qs, lam, dbar, post);
        r5_step_vjp<F>(t, y, h, p, lam, pb, a.aux, any ? dbar : nullptr);
        GSX_UNROLL for (int d = 0; d < D; ++d) lam[d] += post[d];
      } else {
        r5_step_vjp<F>(t, y, h, p, lam, pb, a.aux);
        inject_saves<D>(a, j, s, qs, lam);
      }
    }
  }
  if (dense) {
    // the rest are at or before t0: y0 itself
    for (; qs >= 0; --qs) {
      GSX_UNROLL for (int d = 0; d < D; ++d)
        lam[d] += a.ysbar[(static_cast<int64_t>(qs) * D + d) * n + j];
    }
  } else {
    inject_saves<D>(a, j, -1, qs, lam);
  }
  GSX_UNROLL for (int d = 0; d < D; ++d) a.y0bar[d * n + j] = lam[d];
  GSX_UNROLL for (int q = 0; q < P; ++q) a.pbar[q * n + j] = pb[q];
}

// Dense output of the continuous extension for one trajectory (gradsolve/solvers/dense.py:
// 166-227): the replay forms each step's stages once; a save time in (t_s, t_s + h_s] is the
// step's interpolant at theta = (ts - t_s) / h_s; at or past the final time it is y_final
// (save_step = n_acc), and at or before t0 it is y0 (save_step = -1). save_step / save_t
// record the bracket for the adjoint.
template <class F>
GSX_HD void r5_saveat_dense_lane(const SaveArgs& a, int64_t j) {
  constexpr int D = F::D, P = F::P;
  const int64_t n = a.n, k = a.nsave;
  double y[D], p[P];
  GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = a.y0[d * n + j];
  GSX_UNROLL for (int q = 0; q < P; ++q) p[q] = a.p[q * n + j];
  auto put = [&](int64_t q, const double* v, int32_t st, double tq) {
    GSX_UNROLL for (int d = 0; d < D; ++d) a.ys[(q * D + d) * n + j] = v[d];
    a.save_step[q * n + j] = st;
    a.save_t[q * n + j] = tq;
  };
  double t = a.t0;
  int64_t q = 0;
  while (q < k && a.ts[q] <= a.t0) {
    put(q, y, -1, t);
    ++q;
  }
  const int32_t na = a.n_acc[j];
  for (int32_t s = 0; s < na; ++s) {
    const double h = a.tape[static_cast<int64_t>(s) * n + j];
    const double tn = t + h;
    R5Stages<F> w;
    r5_stages<F>(t, y, h, p, a.aux, w);
    double yn[D];
    r5_combine<F>(y, w, yn);
    if (q < k && a.ts[q] <= tn && !(s == na - 1 && a.ts[q] >= tn)) {
      double dv[3][D];
      r5_dense_vectors<F>(w, dv);
      while (q < k && a.ts[q] <= tn) {
        if (s == na - 1 && a.ts[q] >= tn) break;  // at the final time: y_final, below
        double yq[D];
        r5_dense_eval<F>(y, yn, dv, (a.ts[q] - t) / h, yq);
        put(q, yq, s, t);
        ++q;
      }
    }
    GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = yn[d];
    t = tn;
  }
  if (a.y_final) {
    GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = a.y_final[d * n + j];
  }
  for (; q < k; ++q) put(q, y, na, t);
}

template <class F>
GSX_HD void r5_saveat_lane(const SaveArgs& a, int64_t j) {
  saveat_lane<F, MethodRodas5P>(a, j);
}

template <class F>
GSX_HD void r5_saveat_vjp_lane(const SaveVjpArgs& a, int64_t j) {
  saveat_vjp_lane<F, MethodRodas5P>(a, j);
}

}  // namespace GSX_NS
