// gsx_plain.hip -- the naive HIP reverse sweep, the plain arm the benchmark measures the fused
// path against (it is not used by the drop-in path). One trajectory per lane, as the fused
// kernels, but with no checkpoints and no recomputation: the forward stores the six stage
// inputs z_1..z_6 of every accepted step, and the reverse reads them back and applies the
// stage VJPs (the VJP of gradsolve/solvers/tsit5_step.py:80-102 without re-evaluating the
// field). Layout: stages [cap][6][D][n], element (s, i, d, j) at ((s*6 + i)*D + d)*n + j.
// The reverse sweep walks a wavefront's lanes in step, from the wavefront's largest count
// down (a lane joins once the walk reaches its own last step), so that every stage-tape read
// of the wavefront is one contiguous row segment, as in the fused adjoints.
#include <hip/hip_runtime.h>

#include "gsx_internal.h"

namespace gsx {
namespace {

constexpr int kPlainBlock = 256;

template <class F>
__device__ void plain_fwd_lane(const PlainArgs& a, int64_t j) {
  constexpr int D = F::D, P = F::P;
  using namespace tab;
  const int64_t n = a.n;
  double y[D], p[P], k[7][D], z[6][D], y5[D];
  GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = a.y0[d * n + j];
  GSX_UNROLL for (int q = 0; q < P; ++q) p[q] = a.p[q * n + j];
  double t = a.t0;
  const double t1 = a.t1;
  double dt = (t1 - t) / 100.0;
  F::f(t, y, p, k[0], a.aux);
  int32_t acc = 0, rej = 0, status = kReached;
  for (int64_t it = 0; it < a.max_steps; ++it) {
    if (t >= t1) break;
    const double h = nn_min(dt, t1 - t);
    GSX_UNROLL for (int i = 0; i < D; ++i) {
      z[0][i] = y[i];
      z[1][i] = y[i] + h * (a21 * k[0][i]);
    }
    F::f(t, z[1], p, k[1], a.aux);
    GSX_UNROLL for (int i = 0; i < D; ++i) z[2][i] = y[i] + h * (a31 * k[0][i] + a32 * k[1][i]);
    F::f(t, z[2], p, k[2], a.aux);
    GSX_UNROLL for (int i = 0; i < D; ++i)
      z[3][i] = y[i] + h * (a41 * k[0][i] + a42 * k[1][i] + a43 * k[2][i]);
    F::f(t, z[3], p, k[3], a.aux);
    GSX_UNROLL for (int i = 0; i < D; ++i)
      z[4][i] = y[i] + h * (a51 * k[0][i] + a52 * k[1][i] + a53 * k[2][i] + a54 * k[3][i]);
    F::f(t, z[4], p, k[4], a.aux);
    GSX_UNROLL for (int i = 0; i < D; ++i)
      z[5][i] = y[i] + h * (a61 * k[0][i] + a62 * k[1][i] + a63 * k[2][i] + a64 * k[3][i] +
                            a65 * k[4][i]);
    F::f(t, z[5], p, k[5], a.aux);
    GSX_UNROLL for (int i = 0; i < D; ++i)
      y5[i] = y[i] + h * (b1 * k[0][i] + b2 * k[1][i] + b3 * k[2][i] + b4 * k[3][i] +
                          b5 * k[4][i] + b6 * k[5][i]);
    F::f(t, y5, p, k[6], a.aux);
    double e[D];
    GSX_UNROLL for (int i = 0; i < D; ++i)
      e[i] = h * (e1 * k[0][i] + e2 * k[1][i] + e3 * k[2][i] + e4 * k[3][i] + e5 * k[4][i] +
                  e6 * k[5][i] + e7 * k[6][i]);
    const double err = tsit5_err<D>(e, y, y5, a.rtol, a.atol);
    if (err != err) {
      rej += static_cast<int32_t>(a.max_steps - it);
      break;
    }
    if (err <= 1.0) {
      if (acc >= a.cap) {
        status = kTapeFull;
        break;
      }
      double* row = a.stages + static_cast<int64_t>(acc) * 6 * D * n + j;
      GSX_UNROLL for (int i = 0; i < 6; ++i) {
        GSX_UNROLL for (int d = 0; d < D; ++d) row[(i * D + d) * n] = z[i][d];
      }
      a.tape[static_cast<int64_t>(acc) * n + j] = h;
      t = t + h;
      GSX_UNROLL for (int d = 0; d < D; ++d) {
        y[d] = y5[d];
        k[0][d] = k[6][d];
      }
      ++acc;
    } else {
      ++rej;
    }
    dt = tsit5_next_dt(h, err);
  }
  if (status == kReached && !(t >= t1)) status = kExhausted;
  GSX_UNROLL for (int d = 0; d < D; ++d) a.y_final[d * n + j] = y[d];
  a.n_acc[j] = acc;
  a.n_rej[j] = rej;
  a.status[j] = status;
}

template <class F>
__device__ void plain_adj_lane(const PlainArgs& a, int64_t j, int32_t smax) {
  constexpr int D = F::D, P = F::P;
  using namespace tab;
  constexpr double A[6][6] = {{0, 0, 0, 0, 0, 0},
                              {a21, 0, 0, 0, 0, 0},
                              {a31, a32, 0, 0, 0, 0},
                              {a41, a42, a43, 0, 0, 0},
                              {a51, a52, a53, a54, 0, 0},
                              {a61, a62, a63, a64, a65, 0}};
  constexpr double B[6] = {b1, b2, b3, b4, b5, b6};
  const int64_t n = a.n;
  double p[P], pb[P], lam[D];
  GSX_UNROLL for (int q = 0; q < P; ++q) {
    p[q] = a.p[q * n + j];
    pb[q] = 0.0;
  }
  GSX_UNROLL for (int d = 0; d < D; ++d) lam[d] = a.ybar[d * n + j];
  const int32_t na = a.n_acc[j];
  for (int32_t s = smax - 1; s >= 0; --s) {
    if (s >= na) continue;
    const double h = a.tape[static_cast<int64_t>(s) * n + j];
    const double* row = a.stages + static_cast<int64_t>(s) * 6 * D * n + j;
    double hl[D], lam_new[D], kb[6][D];
    GSX_UNROLL for (int i = 0; i < D; ++i) {
      hl[i] = h * lam[i];
      lam_new[i] = lam[i];
    }
    GSX_UNROLL for (int st = 0; st < 6; ++st) {
      GSX_UNROLL for (int i = 0; i < D; ++i) kb[st][i] = B[st] * hl[i];
    }
    GSX_UNROLL for (int st = 5; st >= 0; --st) {
      double z[D], zb[D];
      GSX_UNROLL for (int d = 0; d < D; ++d) {
        z[d] = row[(st * D + d) * n];
        zb[d] = -0.0;  // x + (-0.0) == x: the first add folds
      }
      F::vjp(0.0, z, p, kb[st], zb, pb, a.aux);
      GSX_UNROLL for (int i = 0; i < D; ++i) {
        lam_new[i] += zb[i];
        const double hz = h * zb[i];
        GSX_UNROLL for (int r = 0; r < st; ++r) kb[r][i] += A[st][r] * hz;
      }
    }
    GSX_UNROLL for (int i = 0; i < D; ++i) lam[i] = lam_new[i];
  }
  GSX_UNROLL for (int d = 0; d < D; ++d) a.y0bar[d * n + j] = lam[d];
  GSX_UNROLL for (int q = 0; q < P; ++q) a.pbar[q * n + j] = pb[q];
}

template <class F>
__global__ __launch_bounds__(kPlainBlock) void plain_fwd_kernel(PlainArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kPlainBlock + threadIdx.x;
  if (j < a.n) plain_fwd_lane<F>(a, j);
}

template <class F>
__global__ __launch_bounds__(kPlainBlock) void plain_adj_kernel(PlainArgs a) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kPlainBlock + threadIdx.x;
  const bool valid = j < a.n;
  int32_t smax = valid ? a.n_acc[j] : 0;  // the wavefront's largest count
  GSX_UNROLL for (int off = 32; off >= 1; off >>= 1) {
    const int32_t o = __shfl_xor(smax, off);
    smax = o > smax ? o : smax;
  }
  if (valid) plain_adj_lane<F>(a, j, smax);
}

unsigned plain_grid(int64_t n) {
  return static_cast<unsigned>((n + kPlainBlock - 1) / kPlainBlock);
}

}  // namespace

int plain_dispatch(int field, bool adjoint, const PlainArgs& a, hipStream_t s) {
  if (a.n == 0) return GSX_OK;
  switch (field) {
    case GSX_FIELD_LORENZ:
      if (adjoint) {
        plain_adj_kernel<FLorenz><<<plain_grid(a.n), kPlainBlock, 0, s>>>(a);
      } else {
        plain_fwd_kernel<FLorenz><<<plain_grid(a.n), kPlainBlock, 0, s>>>(a);
      }
      return check_launch(adjoint ? "plain_adj_kernel" : "plain_fwd_kernel");
    default: break;
  }
  return set_error(GSX_E_UNSUPPORTED, "plain arm: field %d is not built", field);
}

}  // namespace gsx
