// gsx_rb23.h -- per-lane Rosenbrock23 (ode23s, autonomous) forward + record, replay and
// adjoint. W = I - d*h*J(y, p) is LU-factorised once per step in registers and reused by the
// stage solves (gradsolve/cuda/kernels/rosenbrock23_template.cu.in:61-146 factors once per
// step; the Warp kernel re-eliminates per solve, gradsolve/warp/_warp_rosenbrock.py:406-445).
// Included from gsx_core.h's namespace conventions: compiled in GSX_NS.
#pragma once

#include "gsx_core.h"

namespace GSX_NS {

constexpr double kRb23Exp = -1.0 / 3.0;

// d = 1/(2 + sqrt 2), e32 = 6 + sqrt 2 (Shampine & Reichelt 1997), in binary64 as the
// reference computes them (gradsolve/warp/_warp_rosenbrock.py:61-62).
GSX_HD double rb_d() { return 1.0 / (2.0 + std::sqrt(2.0)); }
GSX_HD double rb_e32() { return 6.0 + std::sqrt(2.0); }

// In-place LU with partial pivoting of a D x D row-major matrix held in registers. Row swaps
// and the pivot search use compile-time indices with data-dependent selects, so nothing is
// indexed dynamically (no scratch memory). piv[c] is the row swapped with row c at step c.
// The default variant keeps the reciprocal of U's diagonal in place of the diagonal after
// factoring: one refined reciprocal per pivot (pivot_recip), and the multipliers and every
// triangular solve multiply by it (within an ulp or so of the quotients). The strict variant keeps U's diagonal and
// divides, as LAPACK's factorisation and triangular solves do.
#if defined(GSX_STRICT_ARITH)
constexpr bool kLuRecipDiag = false;
#else
constexpr bool kLuRecipDiag = true;
#endif

// 1/b for a pivot of the default variant: the hardware reciprocal refined by two Newton steps
// (each squares the relative error, so the result is within an ulp or so of 1/b) in place of
// the correctly rounded division's scale-and-fixup sequence. A zero pivot gives a non-finite
// result, as the division does.
GSX_HD double pivot_recip(double b) {
  double r = recip_seed(b);
  r = std::fma(r, std::fma(-b, r, 1.0), r);
  return std::fma(r, std::fma(-b, r, 1.0), r);
}

template <int D>
GSX_HD void lu_factor(double (&a)[D * D], int (&piv)[D]) {
  GSX_UNROLL for (int c = 0; c < D; ++c) {
    int pr = c;
    double big = std::fabs(a[c * D + c]);
    GSX_UNROLL for (int r = c + 1; r < D; ++r) {
      const double v = std::fabs(a[r * D + c]);
      if (v > big) {
        big = v;
        pr = r;
      }
    }
    piv[c] = pr;
    GSX_UNROLL for (int r = c + 1; r < D; ++r) {
      const bool sw = (pr == r);
      GSX_UNROLL for (int k = 0; k < D; ++k) {
        const double x = a[c * D + k], z = a[r * D + k];
        a[c * D + k] = sw ? z : x;
        a[r * D + k] = sw ? x : z;
      }
    }
    const double inv = kLuRecipDiag ? pivot_recip(a[c * D + c]) : 0.0;
    GSX_UNROLL for (int r = c + 1; r < D; ++r) {
      const double l = kLuRecipDiag ? a[r * D + c] * inv : a[r * D + c] / a[c * D + c];
      a[r * D + c] = l;
      GSX_UNROLL for (int k = c + 1; k < D; ++k) a[r * D + k] = a[r * D + k] - l * a[c * D + k];
    }
    if (kLuRecipDiag) a[c * D + c] = inv;
  }
}

// Solve W x = b with P W = L U.
template <int D>
GSX_HD void lu_solve(const double (&a)[D * D], const int (&piv)[D], const double* b, double* x) {
  double w[D];
  GSX_UNROLL for (int i = 0; i < D; ++i) w[i] = b[i];
  GSX_UNROLL for (int c = 0; c < D; ++c) {
    GSX_UNROLL for (int r = c + 1; r < D; ++r) {
      const bool sw = (piv[c] == r);
      const double u = w[c], v = w[r];
      w[c] = sw ? v : u;
      w[r] = sw ? u : v;
    }
  }
  GSX_UNROLL for (int i = 0; i < D; ++i) {
    GSX_UNROLL for (int k = 0; k < i; ++k) w[i] = w[i] - a[i * D + k] * w[k];
  }
  GSX_UNROLL for (int i = D - 1; i >= 0; --i) {
    double s = w[i];
    GSX_UNROLL for (int k = i + 1; k < D; ++k) s = s - a[i * D + k] * x[k];
    x[i] = kLuRecipDiag ? s * a[i * D + i] : s / a[i * D + i];
  }
}

// Solve W^T x = b: U^T z = b, L^T v = z, then undo the row swaps in reverse order.
template <int D>
GSX_HD void lu_solve_t(const double (&a)[D * D], const int (&piv)[D], const double* b,
                       double* x) {
  double z[D];
  GSX_UNROLL for (int i = 0; i < D; ++i) {
    double s = b[i];
    GSX_UNROLL for (int k = 0; k < i; ++k) s = s - a[k * D + i] * z[k];
    z[i] = kLuRecipDiag ? s * a[i * D + i] : s / a[i * D + i];
  }
  GSX_UNROLL for (int i = D - 1; i >= 0; --i) {
    double s = z[i];
    GSX_UNROLL for (int k = i + 1; k < D; ++k) s = s - a[k * D + i] * x[k];
    x[i] = s;
  }
  GSX_UNROLL for (int c = D - 1; c >= 0; --c) {
    GSX_UNROLL for (int r = c + 1; r < D; ++r) {
      const bool sw = (piv[c] == r);
      const double u = x[c], v = x[r];
      x[c] = sw ? v : u;
      x[r] = sw ? u : v;
    }
  }
}

template <class F>
GSX_HD void rb23_build_w(double t, const double* y, double h, const double* p,
                         const double* aux, double (&W)[F::D * F::D], int (&piv)[F::D]) {
  constexpr int D = F::D;
  double J[D * D];
  F::jac(t, y, p, J, aux);
  const double dh = h * rb_d();
  GSX_UNROLL for (int i = 0; i < D; ++i) {
    GSX_UNROLL for (int j = 0; j < D; ++j) W[i * D + j] = (i == j ? 1.0 : 0.0) - dh * J[i * D + j];
  }
  lu_factor<D>(W, piv);
}

// One adaptive trial step (gradsolve tests/_oracles/gradsolve_g2_spike_oracle.py:172-183).
template <class F>
GSX_HD double rb23_trial(double t, const double* y, double h, const double* p,
                         const double* aux, double* ynew, double rtol, double atol) {
  constexpr int D = F::D;
  double W[D * D];
  int piv[D];
  rb23_build_w<F>(t, y, h, p, aux, W, piv);
  double F0[D], F1[D], F2[D], k1[D], k2[D], k3[D], u[D], r[D];
  F::f(t, y, p, F0, aux);
  lu_solve<D>(W, piv, F0, k1);
  GSX_UNROLL for (int i = 0; i < D; ++i) u[i] = y[i] + 0.5 * h * k1[i];
  F::f(t, u, p, F1, aux);
  GSX_UNROLL for (int i = 0; i < D; ++i) r[i] = F1[i] - k1[i];
  lu_solve<D>(W, piv, r, k2);
  GSX_UNROLL for (int i = 0; i < D; ++i) k2[i] = k2[i] + k1[i];
  GSX_UNROLL for (int i = 0; i < D; ++i) ynew[i] = y[i] + h * k2[i];
  F::f(t, ynew, p, F2, aux);
  const double e32 = rb_e32();
  GSX_UNROLL for (int i = 0; i < D; ++i)
    r[i] = F2[i] - e32 * (k2[i] - F1[i]) - 2.0 * (k1[i] - F0[i]);
  lu_solve<D>(W, piv, r, k3);
  double e[D];
#if defined(GSX_STRICT_ARITH)
  const double h6 = h / 6.0;
#else
  const double h6 = h * (1.0 / 6.0);  // within an ulp of h / 6, without a division
#endif
  GSX_UNROLL for (int i = 0; i < D; ++i) e[i] = h6 * (k1[i] - 2.0 * k2[i] + k3[i]);
  return wrms<D>(e, y, ynew, rtol, atol);
}

// Replay step rosenbrock23_step (gradsolve/solvers/rosenbrock23_step.py:47-66): no k3.
template <class F>
GSX_HD void rb23_advance(double t, double* y, double h, const double* p, const double* aux) {
  constexpr int D = F::D;
  double W[D * D];
  int piv[D];
  rb23_build_w<F>(t, y, h, p, aux, W, piv);
  double F0[D], F1[D], k1[D], k2[D], u[D], r[D];
  F::f(t, y, p, F0, aux);
  lu_solve<D>(W, piv, F0, k1);
  GSX_UNROLL for (int i = 0; i < D; ++i) u[i] = y[i] + 0.5 * h * k1[i];
  F::f(t, u, p, F1, aux);
  GSX_UNROLL for (int i = 0; i < D; ++i) r[i] = F1[i] - k1[i];
  lu_solve<D>(W, piv, r, k2);
  GSX_UNROLL for (int i = 0; i < D; ++i) y[i] = y[i] + h * (k2[i] + k1[i]);
}

// What the reverse of one step needs from its forward: the LU of W, k1 = W^-1 f(y) and
// x2 = W^-1 (f(u) - k1) (so that k2 = x2 + k1 and y_next = y + h (x2 + k1)).
template <class F>
struct Rb23Step {
  static constexpr int D = F::D;
  double W[D * D];
  int piv[D];
  double k1[D], x2[D];
};

// The forward of one replay step at (t, y, h), keeping what its VJP needs; y_next (when
// given) is rb23_advance's result bit for bit.
template <class F>
GSX_HD void rb23_step_state(double t, const double* y, double h, const double* p,
                            const double* aux, Rb23Step<F>& w, double* y_next) {
  constexpr int D = F::D;
  rb23_build_w<F>(t, y, h, p, aux, w.W, w.piv);
  double F0[D], F1[D], u[D], r[D];
  F::f(t, y, p, F0, aux);
  lu_solve<D>(w.W, w.piv, F0, w.k1);
  GSX_UNROLL for (int i = 0; i < D; ++i) u[i] = y[i] + 0.5 * h * w.k1[i];
  F::f(t, u, p, F1, aux);
  GSX_UNROLL for (int i = 0; i < D; ++i) r[i] = F1[i] - w.k1[i];
  lu_solve<D>(w.W, w.piv, r, w.x2);
  if (y_next) {
    GSX_UNROLL for (int i = 0; i < D; ++i) y_next[i] = y[i] + h * (w.x2[i] + w.k1[i]);
  }
}

// Adjoint of rb23_advance at (t, y, h) from its forward state: lam <- (dy_next/dy)^T lam;
// pb += (dy_next/dp)^T lam. Differentiates through the solves (x = W^-1 r: rbar = W^-T xbar,
// Wbar -= rbar x^T) and through W = I - d h J(y, p) (Jbar = -d h Wbar, contracted with dJ/dy
// and dJ/dp). With h = 0 it leaves lam and pb unchanged (every cotangent it adds is a zero;
// only the sign of a zero may change).
template <class F>
GSX_HD void rb23_step_vjp_state(double t, const double* y, double h, const double* p,
                                const Rb23Step<F>& w, double* lam, double* pb,
                                const double* aux) {
  constexpr int D = F::D;
  double u[D];
  GSX_UNROLL for (int i = 0; i < D; ++i) u[i] = y[i] + 0.5 * h * w.k1[i];
  // ynew = y + h (x2 + k1)
  double yb[D], k1b[D], x2b[D], r2b[D], ub[D], F0b[D], Wb[D * D];
  GSX_UNROLL for (int i = 0; i < D; ++i) {
    yb[i] = lam[i];
    x2b[i] = h * lam[i];
    k1b[i] = h * lam[i];
  }
  lu_solve_t<D>(w.W, w.piv, x2b, r2b);
  GSX_UNROLL for (int i = 0; i < D; ++i) {
    GSX_UNROLL for (int j = 0; j < D; ++j) Wb[i * D + j] = -r2b[i] * w.x2[j];
  }
  GSX_UNROLL for (int i = 0; i < D; ++i) {
    k1b[i] -= r2b[i];
    ub[i] = -0.0;  // x + (-0.0) == x: the first add folds
  }
  F::vjp(t, u, p, r2b, ub, pb, aux);
  GSX_UNROLL for (int i = 0; i < D; ++i) {
    yb[i] += ub[i];
    k1b[i] += 0.5 * h * ub[i];
  }
  lu_solve_t<D>(w.W, w.piv, k1b, F0b);
  GSX_UNROLL for (int i = 0; i < D; ++i) {
    GSX_UNROLL for (int j = 0; j < D; ++j) Wb[i * D + j] -= F0b[i] * w.k1[j];
  }
  F::vjp(t, y, p, F0b, yb, pb, aux);
  const double dh = h * rb_d();
  double Jb[D * D];
  GSX_UNROLL for (int i = 0; i < D * D; ++i) Jb[i] = -dh * Wb[i];
  F::jac_vjp(t, y, p, Jb, yb, pb, aux);
  GSX_UNROLL for (int i = 0; i < D; ++i) lam[i] = yb[i];
}

// Adjoint of rb23_advance at (t, y, h): the step's forward state, then its reverse.
template <class F>
GSX_HD void rb23_step_vjp(double t, const double* y, double h, const double* p, double* lam,
                          double* pb, const double* aux) {
  Rb23Step<F> w;
  rb23_step_state<F>(t, y, h, p, aux, w, nullptr);
  rb23_step_vjp_state<F>(t, y, h, p, w, lam, pb, aux);
}

// Rosenbrock23 as the step policy of the dense-output lanes (saveat_lane / saveat_vjp_lane):
// the re-step to a save time is rosenbrock23_step of size ts - t_s, as the reference's
// replay_solve_rosenbrock_saveat_jax takes it (gradsolve/warp/warp_replay.py:78-84).
struct MethodRb23 {
  template <class F>
  GSX_HD static void advance(double t, double* y, double h, const double* p, const double* aux) {
    rb23_advance<F>(t, y, h, p, aux);
  }
  template <class F>
  GSX_HD static void step_vjp(double t, const double* y, double h, const double* p,
                              double* lam, double* pb, const double* aux) {
    rb23_step_vjp<F>(t, y, h, p, lam, pb, aux);
  }
};

template <class F>
GSX_HD void rb23_saveat_lane(const SaveArgs& a, int64_t j) {
  saveat_lane<F, MethodRb23>(a, j);
}

template <class F>
GSX_HD void rb23_saveat_vjp_lane(const SaveVjpArgs& a, int64_t j) {
  saveat_vjp_lane<F, MethodRb23>(a, j);
}

// Forward + record (same tape/checkpoint contract as the Tsit5 lane).
template <class F>
GSX_HD void rb23_fwd_lane(const FwdArgs& a, int64_t j) {
  constexpr int D = F::D, P = F::P;
  const int64_t n = a.n;
  double y[D], p[P], ynew[D];
  GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = a.y0[d * n + j];
  GSX_UNROLL for (int q = 0; q < P; ++q) p[q] = a.p[q * n + j];
  double t = a.t0;
  const double t1 = a.t1_per ? a.t1_per[j] : a.t1;
  double dt = (t1 - t) / 100.0;
  int32_t acc = 0, rej = 0, status = kReached;
  const bool rec = a.tape != nullptr;
  const bool ck = a.ckpt != nullptr;
  int32_t ck_left = a.ckpt_interval;
  double* tp = rec ? a.tape + j : nullptr;
  double* cp = ck ? a.ckpt + j : nullptr;
  const int64_t cstride = static_cast<int64_t>(ckpt_rows<F>()) * n;
  if (ck) ckpt_put<F>(cp, n, y, t);
  for (int64_t it = 0; it < a.max_steps; ++it) {
    if (t >= t1) break;
    const double h = py_min(dt, t1 - t);
    const double err = rb23_trial<F>(t, y, h, p, a.aux, ynew, a.rtol, a.atol);
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
    } else {
      ++rej;
    }
    const double fac = kSafety * ctrl_pow_rb23(py_max(err, kErrFloor));
    dt = h * py_min(py_max(fac, kFacMin), kFacMax);
  }
  if (status == kReached && !(t >= t1)) status = kExhausted;
  GSX_UNROLL for (int d = 0; d < D; ++d) a.y_final[d * n + j] = y[d];
  // the counts and the status are optional outputs of a forward without a tape
  if (a.n_acc) a.n_acc[j] = acc;
  if (a.n_rej) a.n_rej[j] = rej;
  if (a.status) a.status[j] = status;
}

template <class F>
GSX_HD void rb23_replay_lane(const ReplayArgs& a, int64_t j) {
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
    rb23_advance<F>(t, y, h, p, a.aux);
    t = t + h;
    if (ck && --ck_left == 0) {
      ckpt_store<F>(a.ckpt, (s + 1) / a.ckpt_interval, n, j, y, t);
      ck_left = a.ckpt_interval;
    }
  }
  GSX_UNROLL for (int d = 0; d < D; ++d) a.y_final[d * n + j] = y[d];
}

template <class F>
GSX_HD void rb23_adj_lane(const AdjArgs& a, int64_t j, double* seg, int64_t sstride,
                          int32_t mtop = 0) {
  constexpr int D = F::D, P = F::P;
  constexpr int DCK = ckpt_rows<F>();
  const int64_t n = a.n;
  double p[P], pb[P], lam[D];
  GSX_UNROLL for (int q = 0; q < P; ++q) {
    p[q] = a.p[q * n + j];
    pb[q] = a.pbar_in ? a.pbar_in[q * n + j] : 0.0;
  }
  GSX_UNROLL for (int d = 0; d < D; ++d) lam[d] = a.ybar[d * n + j];
  int32_t qs = static_cast<int32_t>(a.nsave) - 1;  // save cursor, walked downwards
  const int32_t na = a.n_acc[j];
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
        rb23_advance<F>(t, y, h, p, a.aux);
        t = t + h;
      }
    }
    for (int32_t s = s1 - 1; s >= s0; --s) {
      const double* slot = seg + static_cast<int64_t>(s - s0) * DCK * sstride;
      GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = slot[d * sstride];
      if constexpr (!F::autonomous) t = slot[D * sstride];
      const double h = a.tape[static_cast<int64_t>(s) * n + j];
      rb23_step_vjp<F>(t, y, h, p, lam, pb, a.aux);
      inject_saves<D>(a, j, s, qs, lam);
    }
  }
  inject_saves<D>(a, j, -1, qs, lam);
  GSX_UNROLL for (int d = 0; d < D; ++d) a.y0bar[d * n + j] = lam[d];
  GSX_UNROLL for (int q = 0; q < P; ++q) a.pbar[q * n + j] = pb[q];
}

// The adjoint with checkpoints every 2 accepted steps (the Tsit5 pair adjoint's schedule,
// tsit5_adj_pair_segment): for segment m one forward state of step 2m (W's LU, k1, x2) gives
// y_{2m+1}; step 2m+1's forward state is formed for its own VJP, then step 2m's VJP reuses the
// state kept from the advance, so each step's W is built and factored once and nothing is
// re-integrated. A missing second step (odd n_acc) runs with h = 0, which leaves lam and pb
// unchanged, so every lane follows the same control flow; `m` is wavefront-uniform in the
// kernel. kSaves as in the Tsit5 pair adjoint.
template <class F, bool kSaves = true>
GSX_HD void rb23_adj_pair_segment(const AdjArgs& a, int64_t j, int32_t m, int32_t na,
                                  const double* p, double* lam, double* pb, int32_t& qs) {
  constexpr int D = F::D;
  constexpr int DCK = ckpt_rows<F>();
  const int64_t n = a.n;
  const int32_t s0 = 2 * m;
  const double* base = a.ckpt + static_cast<int64_t>(m) * DCK * n + j;
  double y0[D], y1[D];
  GSX_UNROLL for (int d = 0; d < D; ++d) y0[d] = base[d * n];
  double t0 = 0.0;
  if constexpr (!F::autonomous) t0 = base[D * n];
  const double h0 = a.tape[static_cast<int64_t>(s0) * n + j];
  double h1 = 0.0;
  if (s0 + 1 < na) h1 = a.tape[static_cast<int64_t>(s0 + 1) * n + j];
  Rb23Step<F> w0;
  rb23_step_state<F>(t0, y0, h0, p, a.aux, w0, y1);
  {
    const double t1 = t0 + h0;
    Rb23Step<F> w1;
    rb23_step_state<F>(t1, y1, h1, p, a.aux, w1, nullptr);
    rb23_step_vjp_state<F>(t1, y1, h1, p, w1, lam, pb, a.aux);
  }
  if constexpr (kSaves) {
    if (s0 + 1 < na) inject_saves<D>(a, j, s0 + 1, qs, lam);
  }
  rb23_step_vjp_state<F>(t0, y0, h0, p, w0, lam, pb, a.aux);
  if constexpr (kSaves) inject_saves<D>(a, j, s0, qs, lam);
}

// The pair adjoint of one lane; `mmax` as in tsit5_adj_pair_lane.
template <class F, bool kSaves = true>
GSX_HD void rb23_adj_pair_lane(const AdjArgs& a, int64_t j, int32_t na, int32_t mmax) {
  constexpr int D = F::D, P = F::P;
  const int64_t n = a.n;
  double p[P], pb[P], lam[D];
  GSX_UNROLL for (int q = 0; q < P; ++q) {
    p[q] = a.p[q * n + j];
    pb[q] = (kSaves && a.pbar_in) ? a.pbar_in[q * n + j] : 0.0;
  }
  GSX_UNROLL for (int d = 0; d < D; ++d) lam[d] = a.ybar[d * n + j];
  int32_t qs = kSaves ? static_cast<int32_t>(a.nsave) - 1 : -1;
  const int32_t nseg = (na + 1) / 2;
  for (int32_t m = mmax - 1; m >= 0; --m) {
    if (m < nseg) rb23_adj_pair_segment<F, kSaves>(a, j, m, na, p, lam, pb, qs);
  }
  if constexpr (kSaves) inject_saves<D>(a, j, -1, qs, lam);
  GSX_UNROLL for (int d = 0; d < D; ++d) a.y0bar[d * n + j] = lam[d];
  GSX_UNROLL for (int q = 0; q < P; ++q) a.pbar[q * n + j] = pb[q];
}

}  // namespace GSX_NS
