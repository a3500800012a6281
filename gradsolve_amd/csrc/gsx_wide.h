// gsx_wide.h -- Tsit5 lanes for fields wider than a lane's registers hold (fields compiled at
// run time with D above 64: the reference's high-NVAR cells, where gradsolve's router picks its
// general record-and-replay engine, gradsolve/dispatch.py:256-258 with
// gradsolve/solvers/record_jax.py:171-213 and tsit5_replay.py:192-232).
//
// One trajectory per lane, as in the register lanes, but the state, the stage derivatives and
// the stage inputs live in a per-trajectory workspace in HBM laid out [row][D][n] (trajectory
// index fastest): every component access of a wavefront is one 512-byte row segment, and the
// field reads and writes its vectors through strided views. Stages run in a loop with one call
// site of the field (and one of its VJP in the adjoint), so a wide generated field is compiled
// once per kernel, and loops over components are not unrolled. The arithmetic is the register
// lanes' (the tableau, the stage sums' left-to-right order, the NumPy-order WRMS norm, the
// controller, the tape and checkpoint layouts); only the rounding of contracted products may
// differ.
//
// Workspace rows (gsx_args.h):
//   forward / replay: y, y5 (exchanged on accept), z (stage input), k1..k7 (k1 and k7
//   exchanged on accept: FSAL without a copy) -- kWideFwdRows
//   adjoint: for each of the segment's C steps its start state and its stage inputs z2..z6
//   (6 rows), then k1..k6, kbar1..kbar6, lam and zbar -- wide_adj_rows(C)
#pragma once

#include "gsx_core.h"

namespace GSX_NS {

using ::gsx::kWideFwdRows;
using ::gsx::wide_adj_rows;
using ::gsx::WideAdjArgs;
using ::gsx::WideFwdArgs;
using ::gsx::WideReplayArgs;

// A vector of one trajectory in the workspace: element i at b[i * s] (s = n).
template <class T>
struct Strided {
  T* b;
  int64_t s;
  GSX_HD T& operator[](int i) const { return b[static_cast<int64_t>(i) * s]; }
};

// Stage weights by row: rows 1..5 are a_s1..a_s5 (stage inputs z2..z6), row 6 is b1..b6 (the
// step's result); stage times c_s (row 6: the end of the step, where k7 is taken).
static __constant__ double kWideW[7][6] = {
    {0, 0, 0, 0, 0, 0},
    {tab::a21, 0, 0, 0, 0, 0},
    {tab::a31, tab::a32, 0, 0, 0, 0},
    {tab::a41, tab::a42, tab::a43, 0, 0, 0},
    {tab::a51, tab::a52, tab::a53, tab::a54, 0, 0},
    {tab::a61, tab::a62, tab::a63, tab::a64, tab::a65, 0},
    {tab::b1, tab::b2, tab::b3, tab::b4, tab::b5, tab::b6}};
static __constant__ double kWideC[7] = {0.0, tab::c2, tab::c3, tab::c4, tab::c5, tab::c6, 1.0};
static __constant__ double kWideE[7] = {tab::e1, tab::e2, tab::e3, tab::e4,
                                        tab::e5, tab::e6, tab::e7};

// NumPy's pairwise summation order for N terms q(OFF)..q(OFF + N - 1): left to right below 8,
// eight interleaved accumulators up to 128, and above that the two halves (split at a multiple
// of 8) summed recursively (numpy's pairwise_sum). The terms are formed on demand.
template <int N, int OFF, class Q>
__device__ double np_pairwise(const Q& q) {
  if constexpr (N < 8) {
    double r = q(OFF);
    for (int i = 1; i < N; ++i) r = r + q(OFF + i);
    return r;
  } else if constexpr (N <= 128) {
    double r[8];
    GSX_UNROLL for (int k = 0; k < 8; ++k) r[k] = q(OFF + k);
    constexpr int full = N - (N % 8);
#pragma nounroll
    for (int i = 8; i < full; i += 8) {
      GSX_UNROLL for (int k = 0; k < 8; ++k) r[k] = r[k] + q(OFF + i + k);
    }
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (int i = full; i < N; ++i) res = res + q(OFF + i);
    return res;
  } else {
    constexpr int n2 = N / 2 - (N / 2) % 8;
    return np_pairwise<n2, OFF>(q) + np_pairwise<N - n2, OFF + n2>(q);
  }
}

// The error measure of tsit5_err from the sum of the squared scaled errors.
template <int D>
__device__ double wide_err_from_sum(double s) {
#if defined(GSX_STRICT_ARITH)
  return std::sqrt(s / double(D));
#else
  constexpr double inv = 1.0 / D;
  if constexpr ((D & (D - 1)) == 0) return s * inv;
  const double m = s * inv;
  return std::fma(std::fma(-m, double(D), s), inv, m);
#endif
}

// Per-trajectory views of the workspace rows.
template <int D>
struct WideRows {
  double* base;  // &ws[0][0][j]
  int64_t n;
  __device__ Strided<double> row(int r) const {
    return {base + static_cast<int64_t>(r) * D * n, n};
  }
};

// out = y + h (w_s1 k1 + ... + w_sS k_S) for weight row s (stage input z_{s+1}, or the step's
// result for s = 6): the sum runs left to right, as tsit5_step.py:88-102 writes it. krow(r) is
// the workspace row of k_{r+1}, computed from the row indices (an array of views indexed by
// the dynamic r would live in per-lane scratch); four components per iteration keep that many
// independent loads in flight.
template <int D, class KRow>
__device__ void wide_combine(const Strided<double>& y, const KRow& krow, int s, double h,
                             const Strided<double>& out) {
#pragma unroll 4
  for (int i = 0; i < D; ++i) {
    double acc = kWideW[s][0] * krow(0)[i];
    for (int r = 1; r < s; ++r) acc = acc + kWideW[s][r] * krow(r)[i];
    out[i] = y[i] + h * acc;
  }
}

// Stages s0..s1 of the step (t, y, h): stage s >= 1 forms its input z_s (row zrow(s); the
// step's result for s = 6) from y and k_1..k_s, and evaluates k_{s+1} = f(t + c_s h, z_s) into
// k[s] (k7 at the step's result for s = 6); stage 0 evaluates k1 = f(t, y).
template <class F, class KRow, class ZRow>
__device__ void wide_stage_loop(double t, const Strided<double>& y, double h, const double* p,
                                const KRow& krow, const ZRow& zrow, int s0, int s1,
                                const double* aux) {
  constexpr int D = F::D;
#pragma nounroll
  for (int s = s0; s <= s1; ++s) {
    Strided<double> in = y;
    if (s > 0) {
      in = zrow(s);
      wide_combine<D>(y, krow, s, h, in);
    }
    F::f(t + kWideC[s] * h, Strided<const double>{in.b, in.s}, p, krow(s), aux);
  }
}

template <class F>
__device__ void wide_ckpt_put(double* base, int64_t n, const Strided<double>& y, double t) {
  constexpr int D = F::D;
#pragma nounroll
  for (int d = 0; d < D; ++d) base[static_cast<int64_t>(d) * n] = y[d];
  if constexpr (!F::autonomous) base[static_cast<int64_t>(D) * n] = t;
  (void)t;
}

// Forward + accepted-h record (tsit5_fwd_lane's semantics).
template <class F>
__device__ void tsit5w_fwd_lane(const WideFwdArgs& w, int64_t j) {
  constexpr int D = F::D, P = F::P;
  const FwdArgs& a = w.f;
  const int64_t n = a.n;
  const WideRows<D> R{w.ws + j, n};
  double p[P];
  GSX_UNROLL for (int q = 0; q < P; ++q) p[q] = a.p[q * n + j];
  int ry = 0, ry5 = 1, rk1 = 3, rk7 = 9;
  {
    const Strided<double> y = R.row(ry);
#pragma nounroll
    for (int d = 0; d < D; ++d) y[d] = a.y0[static_cast<int64_t>(d) * n + j];
  }
  double t = a.t0;
  const double t1 = a.t1_per ? a.t1_per[j] : a.t1;
  double dt = (t1 - t) / 100.0;
  int32_t acc = 0, rej = 0, status = kReached;
  const bool rec = a.tape != nullptr;
  const bool ck = a.ckpt != nullptr;
  const int32_t cap = rec ? static_cast<int32_t>(py_min<int64_t>(a.tape_cap, INT32_MAX))
                          : INT32_MAX;
  int32_t ck_left = a.ckpt_interval;
  double* tp = rec ? a.tape + j : nullptr;
  double* cp = ck ? a.ckpt + j : nullptr;
  const int64_t cstride = static_cast<int64_t>(ckpt_rows<F>()) * n;
  if (ck) wide_ckpt_put<F>(cp, n, R.row(ry), t);
  bool have_k1 = false;  // k1 = f(t, y) is evaluated once, then carried (FSAL)
  for (int64_t it = 0; it < a.max_steps; ++it) {
    if (t >= t1) break;
    const double h = nn_min(dt, t1 - t);
    const Strided<double> y = R.row(ry), y5 = R.row(ry5);
    // k1 and k7 exchange rows on acceptance (FSAL); k2..k6 are rows 4..8
    auto krow = [&](int r) { return R.row(r == 0 ? rk1 : (r == 6 ? rk7 : 3 + r)); };
    auto zrow = [&](int s) { return s == 6 ? y5 : R.row(2); };
    wide_stage_loop<F>(t, y, h, p, krow, zrow, have_k1 ? 1 : 0, 6, a.aux);
    have_k1 = true;
    const double rtol = a.rtol, atol = a.atol;
    auto q = [&](int i) {
      double e = kWideE[0] * krow(0)[i];
      for (int r = 1; r < 7; ++r) e = e + kWideE[r] * krow(r)[i];
      e = h * e;
#if defined(GSX_STRICT_ARITH)
      const double sc = atol + rtol * py_max(std::fabs(y[i]), std::fabs(y5[i]));
#else
      const double sc = atol + rtol * nn_max(std::fabs(y[i]), std::fabs(y5[i]));
#endif
      const double r = norm_div(e, sc);
      return r * r;
    };
    const double err = wide_err_from_sum<D>(np_pairwise<D, 0>(q));
    if (err != err) {
      rej += static_cast<int32_t>(a.max_steps - it);
      break;
    }
    if (err <= 1.0) {
      if (acc >= cap) {
        status = kTapeFull;
        break;
      }
      t = t + h;
      const int sy = ry, sk = rk1;
      ry = ry5;
      ry5 = sy;
      rk1 = rk7;
      rk7 = sk;
      if (rec) {
        *tp = h;
        tp += n;
      }
      ++acc;
      if (ck && --ck_left == 0) {
        cp += cstride;
        wide_ckpt_put<F>(cp, n, R.row(ry), t);
        ck_left = a.ckpt_interval;
      }
    } else {
      ++rej;
    }
    dt = tsit5_next_dt(h, err);
  }
  if (status == kReached && !(t >= t1)) status = kExhausted;
  const Strided<double> y = R.row(ry);
#pragma nounroll
  for (int d = 0; d < D; ++d) a.y_final[static_cast<int64_t>(d) * n + j] = y[d];
  if (a.n_acc) a.n_acc[j] = acc;
  if (a.n_rej) a.n_rej[j] = rej;
  if (a.status) a.status[j] = status;
}

// Frozen-mesh replay (tsit5_replay_lane's semantics), writing the checkpoints.
template <class F>
__device__ void tsit5w_replay_lane(const WideReplayArgs& w, int64_t j) {
  constexpr int D = F::D, P = F::P;
  const ReplayArgs& a = w.r;
  const int64_t n = a.n;
  const WideRows<D> R{w.ws + j, n};
  double p[P];
  GSX_UNROLL for (int q = 0; q < P; ++q) p[q] = a.p[q * n + j];
  int ry = 0, rn = 1;
  {
    const Strided<double> y = R.row(ry);
#pragma nounroll
    for (int d = 0; d < D; ++d) y[d] = a.y0[static_cast<int64_t>(d) * n + j];
  }
  double t = a.t0;
  const int32_t na = a.n_acc[j];
  const bool ck = a.ckpt != nullptr;
  const int64_t cstride = static_cast<int64_t>(ckpt_rows<F>()) * n;
  double* cp = ck ? a.ckpt + j : nullptr;
  if (ck) wide_ckpt_put<F>(cp, n, R.row(ry), t);
  int32_t ck_left = a.ckpt_interval;
  auto krow = [&](int r) { return R.row(3 + r); };
  for (int32_t s = 0; s < na; ++s) {
    const double h = a.tape[static_cast<int64_t>(s) * n + j];
    const Strided<double> y = R.row(ry);
    auto zrow = [&](int) { return R.row(2); };
    wide_stage_loop<F>(t, y, h, p, krow, zrow, 0, 5, a.aux);
    wide_combine<D>(y, krow, 6, h, R.row(rn));
    const int sw = ry;
    ry = rn;
    rn = sw;
    t = t + h;
    if (ck && --ck_left == 0) {
      cp += cstride;
      wide_ckpt_put<F>(cp, n, R.row(ry), t);
      ck_left = a.ckpt_interval;
    }
  }
  const Strided<double> y = R.row(ry);
#pragma nounroll
  for (int d = 0; d < D; ++d) a.y_final[static_cast<int64_t>(d) * n + j] = y[d];
}

// Adjoint (the VJP of the frozen-h replay, tsit5_step_vjp): each segment of C accepted steps
// is re-integrated from its checkpoint, keeping every step's start state and stage inputs
// z2..z6, then walked in reverse: kbar_s = h b_s lam, and for s = 6..1
// zbar = (df/dz_s)^T kbar_s, lam += zbar, kbar_r += h a_sr zbar (r < s). The segments are
// walked in step across the wavefront from mtop, its largest segment count (tsit5_adj_lane).
template <class F>
__device__ void tsit5w_adj_lane(const WideAdjArgs& w, int64_t j, int32_t mtop) {
  constexpr int D = F::D, P = F::P;
  constexpr int DCK = ckpt_rows<F>();
  const AdjArgs& a = w.a;
  const int64_t n = a.n;
  const int32_t C = a.ckpt_interval;
  const WideRows<D> R{w.ws + j, n};
  // rows: step i of the segment at 6i (its start state) .. 6i+5 (z6) | k1..k6 | kbar1..6 |
  // lam | zbar
  const int rk = 6 * C, rkb = rk + 6, rlam = rkb + 6, rzb = rlam + 1;
  auto krow = [&](int r) { return R.row(rk + r); };
  auto kbrow = [&](int r) { return R.row(rkb + r); };
  const Strided<double> lam = R.row(rlam), zb = R.row(rzb);
  double p[P], pb[P];
  GSX_UNROLL for (int q = 0; q < P; ++q) {
    p[q] = a.p[q * n + j];
    pb[q] = 0.0;
  }
#pragma nounroll
  for (int d = 0; d < D; ++d) lam[d] = a.ybar[static_cast<int64_t>(d) * n + j];
  const int32_t na = a.n_acc[j];
  const int32_t nseg = (na + C - 1) / C;
  for (int32_t m = (mtop > nseg ? mtop : nseg) - 1; m >= 0; --m) {
    if (m >= nseg) continue;  // the wavefront's walk has not reached this lane
    const int32_t s0 = m * C;
    const int32_t s1 = (s0 + C < na) ? s0 + C : na;
    const double* base = a.ckpt + static_cast<int64_t>(m) * DCK * n + j;
    double t = 0.0;
    {
      const Strided<double> y = R.row(0);
#pragma nounroll
      for (int d = 0; d < D; ++d) y[d] = base[static_cast<int64_t>(d) * n];
      if constexpr (!F::autonomous) t = base[static_cast<int64_t>(D) * n];
    }
    const double tseg = t;
    // forward over the segment: every step's stage inputs, and the next step's start state
    // (z_{s+1} of weight row s lands in row b + s; the step's result in row b + 6, the next
    // step's start state)
    for (int32_t s = s0; s < s1; ++s) {
      const int b = 6 * (s - s0);
      const double h = a.tape[static_cast<int64_t>(s) * n + j];
      const Strided<double> y = R.row(b);
      auto zrow = [&](int st) { return R.row(b + st); };
      wide_stage_loop<F>(t, y, h, p, krow, zrow, 0, s + 1 < s1 ? 5 : 4, a.aux);
      if (s + 1 < s1) {
        wide_combine<D>(y, krow, 6, h, R.row(b + 6));
      } else {
        wide_combine<D>(y, krow, 5, h, R.row(b + 5));  // z6 of the segment's last step
      }
      t = t + h;
    }
    // reverse over the segment's steps
    for (int32_t s = s1 - 1; s >= s0; --s) {
      const int b = 6 * (s - s0);
      double ts = tseg;
      if constexpr (!F::autonomous) {
        for (int32_t r = s0; r < s; ++r) ts = ts + a.tape[static_cast<int64_t>(r) * n + j];
      }
      const double h = a.tape[static_cast<int64_t>(s) * n + j];
#pragma unroll 4
      for (int d = 0; d < D; ++d) {
        const double l = lam[d];
        GSX_UNROLL for (int st = 0; st < 6; ++st) kbrow(st)[d] = h * (kWideW[6][st] * l);
      }
#pragma nounroll
      for (int st = 5; st >= 0; --st) {
#pragma nounroll
        for (int d = 0; d < D; ++d) zb[d] = 0.0;
        const Strided<double> zs = R.row(b + st);
        F::vjp(ts + kWideC[st] * h, Strided<const double>{zs.b, n}, p,
               Strided<const double>{kbrow(st).b, n}, zb, pb, a.aux);
#pragma unroll 4
        for (int d = 0; d < D; ++d) {
          const double zbd = zb[d];
          lam[d] = lam[d] + zbd;
          for (int r = 0; r < st; ++r) kbrow(r)[d] = kbrow(r)[d] + h * (kWideW[st][r] * zbd);
        }
      }
    }
  }
#pragma nounroll
  for (int d = 0; d < D; ++d) a.y0bar[static_cast<int64_t>(d) * n + j] = lam[d];
  GSX_UNROLL for (int q = 0; q < P; ++q) a.pbar[q * n + j] = pb[q];
}

template <class F>
__global__ __launch_bounds__(kFwdBlock) void tsit5w_fwd_kernel(WideFwdArgs w) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  if (j >= w.f.n) return;
  tsit5w_fwd_lane<F>(w, j);
}

template <class F>
__global__ __launch_bounds__(kFwdBlock) void tsit5w_replay_kernel(WideReplayArgs w) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  if (j >= w.r.n) return;
  tsit5w_replay_lane<F>(w, j);
}

template <class F>
__global__ __launch_bounds__(kFwdBlock) void tsit5w_adj_kernel(WideAdjArgs w) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kFwdBlock + threadIdx.x;
  const int32_t mtop = wave_max(j < w.a.n ? segment_count(w.a, j) : 0);
  if (j >= w.a.n) return;
  tsit5w_adj_lane<F>(w, j, mtop);
}

}  // namespace GSX_NS
