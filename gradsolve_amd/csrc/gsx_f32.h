// gsx_f32.h -- the FP32 variant of the Tsit5 lanes.
//
// Stages, the field and its vector-Jacobian products run in FP32. The state is a compensated
// pair (s, lo) with value s + lo: each accepted increment is added with Kahan's compensated
// summation, so the state does not accumulate one FP32 rounding per step. Time, the step size,
// the WRMS error norm and the step-size controller stay in FP64, as does the tape (the
// recorded h is the FP64 step; stages use it rounded to FP32). Checkpoints hold the pair
// (s, lo) as two FP32 rows per component, in the same bytes an FP64 checkpoint takes. The
// adjoint keeps its cotangent and parameter-gradient accumulators compensated the same way.
#pragma once

#include "gsx_core.h"

namespace GSX_NS {

// value += d, with (s, lo) a compensated pair (Kahan summation).
GSX_HD void kahan_add(float& s, float& lo, float d) {
  const float yk = d + lo;
  const float t = s + yk;
  lo = yk - (t - s);
  s = t;
}

// One FP32 checkpoint (value rows, then compensation rows) at `base` = &ck[slot][0][j].
template <class F>
GSX_HD void ckpt_put_f32(float* base, int64_t n, const float* y, const float* lo) {
  constexpr int D = F::D;
  GSX_UNROLL for (int d = 0; d < D; ++d) {
    base[d * n] = y[d];
    base[(D + d) * n] = lo[d];
  }
}

template <class F>
GSX_HD void ckpt_store_f32(double* ck, int64_t slot, int64_t n, int64_t j, const float* y,
                           const float* lo) {
  ckpt_put_f32<F>(reinterpret_cast<float*>(ck) + slot * (2 * F::D) * n + j, n, y, lo);
}

template <class F>
GSX_HD void tsit5_fwd_lane_f32(const FwdArgs& a, int64_t j) {
  static_assert(F::autonomous, "the FP32 variant is built for autonomous fields");
  constexpr int D = F::D, P = F::P;
  using namespace tab;
  const int64_t n = a.n;
  float y[D], lo[D], p[P], y5[D], inc[D];
  float k[7][D];
  GSX_UNROLL for (int d = 0; d < D; ++d) {
    const double v = a.y0[d * n + j];
    y[d] = static_cast<float>(v);
    lo[d] = static_cast<float>(v - static_cast<double>(y[d]));
  }
  GSX_UNROLL for (int q = 0; q < P; ++q) p[q] = static_cast<float>(a.p[q * n + j]);
  double t = a.t0;
  const double t1 = a.t1_per ? a.t1_per[j] : a.t1;
  double dt = (t1 - t) / 100.0;
  F::f(0.0f, y, p, k[0], a.aux);
  int32_t acc = 0, rej = 0, status = kReached;
  const bool rec = a.tape != nullptr;
  const bool ck = a.ckpt != nullptr;
  int32_t ck_left = a.ckpt_interval;
  double* tp = rec ? a.tape + j : nullptr;
  float* cp = ck ? reinterpret_cast<float*>(a.ckpt) + j : nullptr;
  const int64_t cstride = static_cast<int64_t>(2 * D) * n;
  if (ck) ckpt_put_f32<F>(cp, n, y, lo);
  const float E1 = float(e1), E2 = float(e2), E3 = float(e3), E4 = float(e4), E5 = float(e5),
              E6 = float(e6), E7 = float(e7);
  for (int64_t it = 0; it < a.max_steps; ++it) {
    if (t >= t1) break;
    const double h = nn_min(dt, t1 - t);
    const float hf = static_cast<float>(h);
    tsit5_stages<F, float>(0.0f, y, hf, p, k, y5, a.aux, inc);
    F::f(0.0f, y5, p, k[6], a.aux);
    double e[D], yd[D], y5d[D];
    GSX_UNROLL for (int i = 0; i < D; ++i) {
      e[i] = static_cast<double>(hf * (E1 * k[0][i] + E2 * k[1][i] + E3 * k[2][i] +
                                       E4 * k[3][i] + E5 * k[4][i] + E6 * k[5][i] +
                                       E7 * k[6][i]));
      yd[i] = static_cast<double>(y[i]) + static_cast<double>(lo[i]);
      y5d[i] = static_cast<double>(y5[i]);
    }
    const double err = tsit5_err<D>(e, yd, y5d, a.rtol, a.atol);
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
      GSX_UNROLL for (int d = 0; d < D; ++d) {
        kahan_add(y[d], lo[d], inc[d]);
        k[0][d] = k[6][d];
      }
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
    dt = tsit5_next_dt(h, err);
  }
  if (status == kReached && !(t >= t1)) status = kExhausted;
  GSX_UNROLL for (int d = 0; d < D; ++d)
    a.y_final[d * n + j] = static_cast<double>(y[d]) + static_cast<double>(lo[d]);
  // the counts and the status are optional outputs of a forward without a tape
  if (a.n_acc) a.n_acc[j] = acc;
  if (a.n_rej) a.n_rej[j] = rej;
  if (a.status) a.status[j] = status;
}

// One FP32 replay step of the compensated state.
template <class F>
GSX_HD void tsit5_advance_f32(float* y, float* lo, float hf, const float* p,
                              const double* aux) {
  constexpr int D = F::D;
  float k[7][D], y5[D], inc[D];
  F::f(0.0f, y, p, k[0], aux);
  tsit5_stages<F, float>(0.0f, y, hf, p, k, y5, aux, inc);
  GSX_UNROLL for (int d = 0; d < D; ++d) kahan_add(y[d], lo[d], inc[d]);
}

template <class F>
GSX_HD void tsit5_replay_lane_f32(const ReplayArgs& a, int64_t j) {
  constexpr int D = F::D, P = F::P;
  const int64_t n = a.n;
  float y[D], lo[D], p[P];
  GSX_UNROLL for (int d = 0; d < D; ++d) {
    const double v = a.y0[d * n + j];
    y[d] = static_cast<float>(v);
    lo[d] = static_cast<float>(v - static_cast<double>(y[d]));
  }
  GSX_UNROLL for (int q = 0; q < P; ++q) p[q] = static_cast<float>(a.p[q * n + j]);
  const int32_t na = a.n_acc[j];
  const bool ck = a.ckpt != nullptr;
  if (ck) ckpt_store_f32<F>(a.ckpt, 0, n, j, y, lo);
  int32_t ck_left = a.ckpt_interval;
  for (int32_t s = 0; s < na; ++s) {
    const float hf = static_cast<float>(a.tape[static_cast<int64_t>(s) * n + j]);
    tsit5_advance_f32<F>(y, lo, hf, p, a.aux);
    if (ck && --ck_left == 0) {
      ckpt_store_f32<F>(a.ckpt, (s + 1) / a.ckpt_interval, n, j, y, lo);
      ck_left = a.ckpt_interval;
    }
  }
  GSX_UNROLL for (int d = 0; d < D; ++d)
    a.y_final[d * n + j] = static_cast<double>(y[d]) + static_cast<double>(lo[d]);
}

// Adjoint of one FP32 replay step (the map y -> y + inc(y)); lam and pb are compensated.
template <class F>
GSX_HD void tsit5_step_vjp_f32(const float* y, float h, const float* p, float* lam,
                               float* lam_lo, float* pb, float* pb_lo, const double* aux) {
  constexpr int D = F::D, P = F::P;
  using namespace tab;
  float z[6][D];
  float k[5][D];
  const float A21 = float(a21), A31 = float(a31), A32 = float(a32), A41 = float(a41),
              A42 = float(a42), A43 = float(a43), A51 = float(a51), A52 = float(a52),
              A53 = float(a53), A54 = float(a54), A61 = float(a61), A62 = float(a62),
              A63 = float(a63), A64 = float(a64), A65 = float(a65);
  GSX_UNROLL for (int i = 0; i < D; ++i) z[0][i] = y[i];
  F::f(0.0f, z[0], p, k[0], aux);
  GSX_UNROLL for (int i = 0; i < D; ++i) z[1][i] = y[i] + h * (A21 * k[0][i]);
  F::f(0.0f, z[1], p, k[1], aux);
  GSX_UNROLL for (int i = 0; i < D; ++i) z[2][i] = y[i] + h * (A31 * k[0][i] + A32 * k[1][i]);
  F::f(0.0f, z[2], p, k[2], aux);
  GSX_UNROLL for (int i = 0; i < D; ++i)
    z[3][i] = y[i] + h * (A41 * k[0][i] + A42 * k[1][i] + A43 * k[2][i]);
  F::f(0.0f, z[3], p, k[3], aux);
  GSX_UNROLL for (int i = 0; i < D; ++i)
    z[4][i] = y[i] + h * (A51 * k[0][i] + A52 * k[1][i] + A53 * k[2][i] + A54 * k[3][i]);
  F::f(0.0f, z[4], p, k[4], aux);
  GSX_UNROLL for (int i = 0; i < D; ++i)
    z[5][i] = y[i] + h * (A61 * k[0][i] + A62 * k[1][i] + A63 * k[2][i] + A64 * k[3][i] +
                          A65 * k[4][i]);
  const float A[6][6] = {{0, 0, 0, 0, 0, 0},      {A21, 0, 0, 0, 0, 0},
                         {A31, A32, 0, 0, 0, 0},  {A41, A42, A43, 0, 0, 0},
                         {A51, A52, A53, A54, 0, 0}, {A61, A62, A63, A64, A65, 0}};
  const float B[6] = {float(b1), float(b2), float(b3), float(b4), float(b5), float(b6)};
  float hl[D];
  GSX_UNROLL for (int i = 0; i < D; ++i) hl[i] = h * lam[i];
  float kb[6][D];
  GSX_UNROLL for (int s = 0; s < 6; ++s) {
    GSX_UNROLL for (int i = 0; i < D; ++i) kb[s][i] = B[s] * hl[i];
  }
  float dlam[D];
  float dp[P];
  // Accumulators start at -0.0, the exact additive identity, so their first adds fold away.
  GSX_UNROLL for (int i = 0; i < D; ++i) dlam[i] = -0.0f;
  GSX_UNROLL for (int q = 0; q < P; ++q) dp[q] = -0.0f;
  GSX_UNROLL for (int s = 5; s >= 0; --s) {
    float zb[D];
    GSX_UNROLL for (int i = 0; i < D; ++i) zb[i] = -0.0f;
    F::vjp(0.0f, z[s], p, kb[s], zb, dp, aux);
    GSX_UNROLL for (int i = 0; i < D; ++i) dlam[i] += zb[i];
    float hz[D];
    GSX_UNROLL for (int i = 0; i < D; ++i) hz[i] = h * zb[i];
    GSX_UNROLL for (int r = 0; r < s; ++r) {
      GSX_UNROLL for (int i = 0; i < D; ++i) kb[r][i] += A[s][r] * hz[i];
    }
  }
  GSX_UNROLL for (int i = 0; i < D; ++i) kahan_add(lam[i], lam_lo[i], dlam[i]);
  GSX_UNROLL for (int q = 0; q < P; ++q) kahan_add(pb[q], pb_lo[q], dp[q]);
}

// The segment buffer is the FP64 one of the FP64 lane: each double cell holds one component's
// compensated pair (s, lo) as two floats, so a slot takes D doubles per lane.
template <class F>
GSX_HD void tsit5_adj_lane_f32(const AdjArgs& a, int64_t j, double* seg, int64_t sstride,
                               int32_t mtop = 0) {
  constexpr int D = F::D, P = F::P;
  const int64_t n = a.n;
  float p[P], pb[P], pb_lo[P], lam[D], lam_lo[D];
  GSX_UNROLL for (int q = 0; q < P; ++q) {
    p[q] = static_cast<float>(a.p[q * n + j]);
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
      if (s + 1 < s1) {
        const float hf = static_cast<float>(a.tape[static_cast<int64_t>(s) * n + j]);
        tsit5_advance_f32<F>(y, lo, hf, p, a.aux);
      }
    }
    for (int32_t s = s1 - 1; s >= s0; --s) {
      const double* slot = seg + static_cast<int64_t>(s - s0) * D * sstride;
      GSX_UNROLL for (int d = 0; d < D; ++d)
        y[d] = reinterpret_cast<const float*>(slot + d * sstride)[0];
      const float hf = static_cast<float>(a.tape[static_cast<int64_t>(s) * n + j]);
      tsit5_step_vjp_f32<F>(y, hf, p, lam, lam_lo, pb, pb_lo, a.aux);
    }
  }
  GSX_UNROLL for (int d = 0; d < D; ++d)
    a.y0bar[d * n + j] = static_cast<double>(lam[d]) + static_cast<double>(lam_lo[d]);
  GSX_UNROLL for (int q = 0; q < P; ++q)
    a.pbar[q * n + j] = static_cast<double>(pb[q]) + static_cast<double>(pb_lo[q]);
}

}  // namespace GSX_NS
