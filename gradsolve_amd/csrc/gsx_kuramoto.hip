// gsx_kuramoto.hip -- Kuramoto ensemble on gfx950: D = 32 phases per trajectory, P = 32
// natural frequencies, one coupling matrix K shared by the ensemble,
//   f(theta) = omega + (1/32) * ((K sin theta) * cos theta - (K cos theta) * sin theta),
// with the coupling products on FP64 MFMA. Forward + record (accepted step sizes, the state at
// the start of every accepted step, and the step bracketing each save time), frozen-mesh
// replay, saveat (one re-step per save time, gradsolve/solvers/dense.py:106-163), and the
// adjoint with saveat cotangents.
//
// Wavefront layout. A wavefront integrates 16 trajectories. Lane l serves trajectory
// j = 16*w + l%16 and holds 8 of its 32 phases: slot m of lane group g = l/16 is phase
// kdim(g, m) = 8*(m%4) + 2*g + m/4. The sums (K sin)_i and (K cos)_i of the 16 trajectories are
// two [16 x 32] x [32 x 32] products, each 2 output tiles x 8 k-steps of
// v_mfma_f64_16x16x4f64. That instruction takes A[i][k] from lane i + 16k, B[k][c] from lane
// c + 16k, and returns D[4r + l/16][l%16] in item r of lane l. With tile T's row rho holding
// K row krow(T, rho) = 8*(rho/4) + 2*(rho%4) + T and k-step kk of lane group g contracting
// phase kdim(g, kk), the B operand of k-step kk is the lane's own slot kk and item r of tile
// T is the lane's own slot 4T + r: sin/cos go into the MFMA from the registers that hold
// theta and the sums come back where they are used, with no LDS transposes or shuffles. The
// adjoint's products K^T (v cos), K^T (v sin) use the same permutation on K^T. The per-lane A
// operands (16 forward, 16 transposed) sit in LDS, shared by the block's 4 wavefronts.
//
// The WRMS norm sums (e/sc)^2 in NumPy's order for 32 elements (eight interleaved
// accumulators r_q = x_q + x_{q+8} + x_{q+16} + x_{q+24}, then
// ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7))): lane group g holds exactly r_{2g} and
// r_{2g+1}, and two xor-shuffles finish the tree in the same order in every lane.
//
// HBM layout (n trajectories, k save times, cap tape rows):
//   y0, omega, y_final, yfbar, y0bar, omega_bar : [32][n]
//   tape                 : [cap][n]        accepted step s of trajectory j at s*n + j
//   states               : [cap][n][32]    the state at the start of accepted step s
//   save_step, save_t    : [k][n]          bracketing step of save time q (-1: ts[q] <= t0,
//                                          the save is y0; -2: ts[q] >= the final time, the
//                                          save is y_final) and that step's start time
//   ys, ysbar, mu        : [n][k][32]      (the reference's y_saved layout)
#include <hip/hip_runtime.h>

#include "gsx_internal.h"

namespace gsx {
namespace {

typedef double kd4 __attribute__((ext_vector_type(4)));
// LDS-qualified double: accesses through it compile to ds_read/ds_write (a generic pointer to
// LDS becomes FLAT instructions when the access is volatile).
typedef __attribute__((address_space(3))) double lds_double;

constexpr int kKD = 32;            // phases per trajectory
constexpr int kSlots = 8;          // phases per lane
constexpr int kTrajPerWave = 16;
constexpr int kKBlock = 256;       // 4 wavefronts, 64 trajectories
constexpr int kOpTable = 16 * 64;  // A operands of one product direction, [T*8 + kk][lane]
constexpr double kInvD = 1.0 / 32.0;

__device__ __forceinline__ int kdim(int g, int m) { return 8 * (m & 3) + 2 * g + (m >> 2); }
__device__ __forceinline__ int krow(int T, int rho) {
  return 8 * (rho >> 2) + 2 * (rho & 3) + T;
}

// Fills the block's operand tables: op[e*64 + l] for e = T*8 + kk, forward then transposed.
__device__ void fill_operands(double* op, const double* K) {
  for (int x = threadIdx.x; x < kOpTable; x += blockDim.x) {
    const int e = x >> 6, l = x & 63;
    const int T = e >> 3, kk = e & 7;
    const int row = krow(T, l & 15), col = kdim(l >> 4, kk);
    op[x] = K[row * kKD + col];
    op[kOpTable + x] = K[col * kKD + row];
  }
  __syncthreads();
}

// a = (K s) and b = (K c) at the lane's 8 slots (tab: forward or transposed table).
__device__ __forceinline__ void coupling(const double* tab, const double* s, const double* c,
                                         double* a, double* b) {
  const int l = static_cast<int>(__lane_id());
  // volatile keeps the operand reads next to their MFMAs instead of 32 hoisted registers
  const volatile lds_double* tv = (const volatile lds_double*)tab;
  kd4 as0 = {0.0, 0.0, 0.0, 0.0}, as1 = as0, ac0 = as0, ac1 = as0;
  GSX_UNROLL for (int kk = 0; kk < 8; ++kk) {
    const double k0 = tv[kk * 64 + l], k1 = tv[(8 + kk) * 64 + l];
    as0 = __builtin_amdgcn_mfma_f64_16x16x4f64(k0, s[kk], as0, 0, 0, 0);
    ac0 = __builtin_amdgcn_mfma_f64_16x16x4f64(k0, c[kk], ac0, 0, 0, 0);
    as1 = __builtin_amdgcn_mfma_f64_16x16x4f64(k1, s[kk], as1, 0, 0, 0);
    ac1 = __builtin_amdgcn_mfma_f64_16x16x4f64(k1, c[kk], ac1, 0, 0, 0);
  }
  GSX_UNROLL for (int r = 0; r < 4; ++r) {
    a[r] = as0[r];
    a[4 + r] = as1[r];
    b[r] = ac0[r];
    b[4 + r] = ac1[r];
  }
}

// The field as the per-lane Tsit5 templates see it: D = P = 8 slots of one trajectory; every
// lane of the wavefront must call it together (the MFMA spans the wavefront). `aux` is the
// block's operand tables.
struct FKuramotoWave {
  __device__ static void f(double, const double* y, const double* om, double* dy,
                           const double* op) {
    double s[kSlots], c[kSlots], a[kSlots], b[kSlots];
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) gsx_sincos(y[m], &s[m], &c[m]);
    coupling(op, s, c, a, b);
    GSX_UNROLL for (int m = 0; m < kSlots; ++m)
      dy[m] = om[m] + kInvD * (a[m] * c[m] - b[m] * s[m]);
  }

  // yb += J^T v, pb += v: (J^T v)_m = (1/32) [c_m (K^T (v c))_m + s_m (K^T (v s))_m
  //                                            - v_m (a_m s_m + b_m c_m)].
  __device__ static void vjp(double, const double* y, const double* om, const double* v,
                             double* yb, double* pb, const double* op) {
    (void)om;
    double s[kSlots], c[kSlots], a[kSlots], b[kSlots];
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) gsx_sincos(y[m], &s[m], &c[m]);
    coupling(op, s, c, a, b);
    double dg[kSlots], vc[kSlots], vs[kSlots], u[kSlots], w[kSlots];
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
      dg[m] = a[m] * s[m] + b[m] * c[m];
      vc[m] = v[m] * c[m];
      vs[m] = v[m] * s[m];
    }
    coupling(op + kOpTable, vc, vs, u, w);
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
      yb[m] += kInvD * (c[m] * u[m] + s[m] * w[m] - v[m] * dg[m]);
      pb[m] += v[m];
    }
  }
};

// WRMS error norm of the lane's trajectory in NumPy's summation order (see the header).
__device__ __forceinline__ double kura_wrms(const double* e, const double* y, const double* y5,
                                            double rtol, double atol) {
  double q[kSlots];
  GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
    const double sc = atol + rtol * py_max(std::fabs(y[m]), std::fabs(y5[m]));
    const double r = norm_div(e[m], sc);
    q[m] = r * r;
  }
  const double r_even = ((q[0] + q[1]) + q[2]) + q[3];  // NumPy accumulator 2g
  const double r_odd = ((q[4] + q[5]) + q[6]) + q[7];   // NumPy accumulator 2g + 1
  double x = r_even + r_odd;
  x = x + __shfl_xor(x, 16);
  x = x + __shfl_xor(x, 32);
  return std::sqrt(x / double(kKD));
}

// Row of 32 phases <-> the lane's 8 slots (slots r and 4 + r are adjacent phases).
__device__ __forceinline__ void store_row(double* row, const double* v) {
  const int g = static_cast<int>(__lane_id()) >> 4;
  GSX_UNROLL for (int r = 0; r < 4; ++r)
    *reinterpret_cast<double2*>(row + 8 * r + 2 * g) = make_double2(v[r], v[4 + r]);
}
__device__ __forceinline__ void load_row(const double* row, double* v) {
  const int g = static_cast<int>(__lane_id()) >> 4;
  GSX_UNROLL for (int r = 0; r < 4; ++r) {
    const double2 x = *reinterpret_cast<const double2*>(row + 8 * r + 2 * g);
    v[r] = x.x;
    v[4 + r] = x.y;
  }
}
__device__ __forceinline__ void load_soa(const double* a, int64_t n, int64_t j, double* v) {
  const int g = static_cast<int>(__lane_id()) >> 4;
  GSX_UNROLL for (int m = 0; m < kSlots; ++m) v[m] = a[kdim(g, m) * n + j];
}
__device__ __forceinline__ void store_soa(double* a, int64_t n, int64_t j, const double* v) {
  const int g = static_cast<int>(__lane_id()) >> 4;
  GSX_UNROLL for (int m = 0; m < kSlots; ++m) a[kdim(g, m) * n + j] = v[m];
}

struct LaneId {
  int64_t j;
  bool valid, leader;  // leader: lane group 0 writes the per-trajectory scalars
};
__device__ __forceinline__ LaneId lane_id(int64_t n) {
  const int64_t wave = (static_cast<int64_t>(blockIdx.x) * kKBlock + threadIdx.x) >> 6;
  const int l = static_cast<int>(__lane_id());
  LaneId id;
  id.j = wave * kTrajPerWave + (l & 15);
  id.valid = id.j < n;
  id.leader = (l >> 4) == 0;
  if (!id.valid) id.j = 0;  // in-bounds reads for idle slots; nothing is written for them
  return id;
}

// Stage derivatives live in LDS slots (this thread's column of a [slots][8][kKBlock] block
// array) and the stage loops are not unrolled, so each kernel holds one copy of the field (its
// 8 sincos and 32 MFMAs) instead of one per stage: unrolled, the kernels exceeded 110 KB of
// code, far past the instruction cache. Accesses are volatile: each thread reads back only its
// own column, and the compiler would otherwise forward the stores into registers.
struct StageSlots {
  volatile lds_double* base;
  __device__ __forceinline__ explicit StageSlots(double* column)
      : base((volatile lds_double*)column) {}
  __device__ __forceinline__ volatile lds_double& at(int s, int m) const {
    return base[(s * kSlots + m) * kKBlock];
  }
};

// Row st of the stage weights: rows 1..5 the inputs of k2..k6 (a_{st+1, 1..st}), row 6 the
// solution weights b_1..b_6, row 7 the error weights e_1..e_7 (b_i - bhat_i).
__device__ __forceinline__ double stage_w(int row, int r) {
  using namespace tab;
  constexpr double W[8][7] = {{0, 0, 0, 0, 0, 0, 0},
                              {a21, 0, 0, 0, 0, 0, 0},
                              {a31, a32, 0, 0, 0, 0, 0},
                              {a41, a42, a43, 0, 0, 0, 0},
                              {a51, a52, a53, a54, 0, 0, 0},
                              {a61, a62, a63, a64, a65, 0, 0},
                              {b1, b2, b3, b4, b5, b6, 0},
                              {e1, e2, e3, e4, e5, e6, e7}};
  return W[row][r];
}

// sum_{r < nk} W[row][r] k_r, summed left to right (nk >= 1).
__device__ __forceinline__ void stage_sum(const StageSlots& ks, int row, int nk, double* acc) {
  const double w0 = stage_w(row, 0);
  GSX_UNROLL for (int m = 0; m < kSlots; ++m) acc[m] = w0 * ks.at(0, m);
  _Pragma("unroll 1") for (int r = 1; r < nk; ++r) {
    const double w = stage_w(row, r);
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) acc[m] = acc[m] + w * ks.at(r, m);
  }
}

// Input of stage st (0: y itself; 1..5: y + h * sum_{r<st} a k_r; 6: y5).
__device__ __forceinline__ void stage_input(const StageSlots& ks, const double* y, double h,
                                            int st, double* z) {
  if (st == 0) {
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) z[m] = y[m];
    return;
  }
  double acc[kSlots];
  stage_sum(ks, st, st, acc);
  GSX_UNROLL for (int m = 0; m < kSlots; ++m) z[m] = y[m] + h * acc[m];
}

// Stages first..6 of a trial step: k_st = f(stage input st) into slot st; y5 = input 6.
__device__ __forceinline__ void trial_stages(const StageSlots& ks, const double* y, double h,
                                             const double* om, const double* op, int first,
                                             double* y5) {
  _Pragma("unroll 1") for (int st = first; st <= 6; ++st) {
    double z[kSlots], kk[kSlots];
    stage_input(ks, y, h, st, z);
    if (st == 6) {
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) y5[m] = z[m];
    }
    FKuramotoWave::f(0.0, z, om, kk, op);
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) ks.at(st, m) = kk[m];
  }
}

// One replay step y <- y + h * sum b_i k_i (k7 is not formed: its weight is 0).
__device__ __forceinline__ void kura_advance(double* y, double h, const double* om,
                                             const double* op, const StageSlots& ks) {
  _Pragma("unroll 1") for (int st = 0; st <= 5; ++st) {
    double z[kSlots], kk[kSlots];
    stage_input(ks, y, h, st, z);
    FKuramotoWave::f(0.0, z, om, kk, op);
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) ks.at(st, m) = kk[m];
  }
  double acc[kSlots];
  stage_sum(ks, 6, 6, acc);
  GSX_UNROLL for (int m = 0; m < kSlots; ++m) y[m] = y[m] + h * acc[m];
}

// Forward + record (gradsolve/solvers/tsit5_replay.py:100-124 per trajectory).
__global__ __launch_bounds__(kKBlock) void kura_fwd_kernel(KuraFwdArgs a) {
  __shared__ double op[2 * kOpTable];
  __shared__ double ks_all[7 * kSlots * kKBlock];
  fill_operands(op, a.K);
  const StageSlots ks(ks_all + threadIdx.x);
  const int64_t n = a.n;
  const LaneId id = lane_id(n);
  const int64_t j = id.j;
  double y[kSlots], om[kSlots], y5[kSlots];
  load_soa(a.y0, n, j, y);
  load_soa(a.omega, n, j, om);
  double t = a.t0;
  const double t1 = a.t1;
  double dt = (t1 - t) / 100.0;
  int32_t acc = 0, rej = 0, status = kReached;
  int64_t it = 0;
  bool active = id.valid && t < t1 && a.max_steps > 0;
  int64_t qn = 0;
  if (a.save_step) {
    for (; qn < a.nsave && a.ts[qn] <= a.t0; ++qn)
      if (id.valid && id.leader) a.save_step[qn * n + j] = -1;
  }
  int first = 0;  // the first trial also forms k1 = f(y0); later ones carry it (FSAL)
  while (__ballot(active) != 0ull) {
    const double h = active ? py_min(dt, t1 - t) : 0.0;
    trial_stages(ks, y, h, om, op, first, y5);
    first = 1;
    double e[kSlots];
    stage_sum(ks, 7, 7, e);
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) e[m] = h * e[m];
    const double err = kura_wrms(e, y, y5, a.rtol, a.atol);
    if (!active) continue;
    ++it;
    if (err != err) {
      rej += static_cast<int32_t>(a.max_steps - (it - 1));
      status = kExhausted;
      active = false;
      continue;
    }
    if (err <= 1.0) {
      if (a.tape && acc >= a.tape_cap) {
        status = kTapeFull;
        active = false;
        continue;
      }
      const double tn = t + h;
      if (a.states) store_row(a.states + (static_cast<int64_t>(acc) * n + j) * kKD, y);
      if (a.tape && id.leader) a.tape[static_cast<int64_t>(acc) * n + j] = h;
      if (a.save_step) {
        for (; qn < a.nsave && a.ts[qn] <= tn; ++qn) {
          if (id.leader) {
            a.save_step[qn * n + j] = acc;
            a.save_t[qn * n + j] = t;
          }
        }
      }
      t = tn;
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
        y[m] = y5[m];
        ks.at(0, m) = ks.at(6, m);  // k1 <- k7
      }
      ++acc;
    } else {
      ++rej;
    }
    const double fac = kSafety * ctrl_pow_tsit5(py_max(err, kErrFloor));
    dt = h * py_min(py_max(fac, kFacMin), kFacMax);
    if (!(t < t1)) {
      active = false;
    } else if (it >= a.max_steps) {
      status = kExhausted;
      active = false;
    }
  }
  if (!id.valid) return;
  if (status == kReached && !(t >= t1)) status = kExhausted;
  store_soa(a.y_final, n, j, y);
  if (id.leader) {
    a.n_acc[j] = acc;
    a.n_rej[j] = rej;
    a.status[j] = status;
    // Save times at or beyond the final time resolve to y_final (dense.py:318-335).
    if (a.save_step)
      for (int64_t q = a.nsave - 1; q >= 0 && a.ts[q] >= t; --q) a.save_step[q * n + j] = -2;
  }
}

// Frozen-mesh replay from (y0, omega): the states at the start of each step and y_final.
__global__ __launch_bounds__(kKBlock) void kura_replay_kernel(KuraReplayArgs a) {
  __shared__ double op[2 * kOpTable];
  __shared__ double ks_all[6 * kSlots * kKBlock];
  fill_operands(op, a.K);
  const StageSlots ks(ks_all + threadIdx.x);
  const int64_t n = a.n;
  const LaneId id = lane_id(n);
  const int64_t j = id.j;
  double y[kSlots], om[kSlots];
  load_soa(a.y0, n, j, y);
  load_soa(a.omega, n, j, om);
  const int32_t na = id.valid ? a.n_acc[j] : 0;
  for (int32_t s = 0; __ballot(s < na) != 0ull; ++s) {
    const bool act = s < na;
    const double h = act ? a.tape[static_cast<int64_t>(s) * n + j] : 0.0;
    if (act && a.states) store_row(a.states + (static_cast<int64_t>(s) * n + j) * kKD, y);
    double yn[kSlots];
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) yn[m] = y[m];
    kura_advance(yn, h, om, op, ks);
    if (act) {
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) y[m] = yn[m];
    }
  }
  if (id.valid) store_soa(a.y_final, n, j, y);
}

// ys[j][q] = one step of size ts[q] - t_s from the bracketing step's start state.
__global__ __launch_bounds__(kKBlock) void kura_saveat_kernel(KuraSaveatArgs a) {
  __shared__ double op[2 * kOpTable];
  __shared__ double ks_all[6 * kSlots * kKBlock];
  fill_operands(op, a.K);
  const StageSlots ks(ks_all + threadIdx.x);
  const int64_t n = a.n, ns = a.nsave;
  const LaneId id = lane_id(n);
  const int64_t j = id.j;
  double y0[kSlots], yf[kSlots], om[kSlots];
  load_soa(a.y0, n, j, y0);
  load_soa(a.y_final, n, j, yf);
  load_soa(a.omega, n, j, om);
  for (int64_t q = 0; q < ns; ++q) {
    const int32_t s = id.valid ? a.save_step[q * n + j] : -1;
    const bool need = id.valid && s >= 0;
    double* row = a.ys + (j * ns + q) * kKD;
    if (__ballot(need) != 0ull) {
      double y[kSlots];
      double h = 0.0;
      if (need) {
        load_row(a.states + (static_cast<int64_t>(s) * n + j) * kKD, y);
        h = a.ts[q] - a.save_t[q * n + j];
      } else {
        GSX_UNROLL for (int m = 0; m < kSlots; ++m) y[m] = y0[m];
      }
      kura_advance(y, h, om, op, ks);
      if (need) store_row(row, y);
    }
    if (id.valid && s == -1) store_row(row, y0);
    if (id.valid && s == -2) store_row(row, yf);
  }
}

// lam <- (d y_next / d y)^T lam, pb += (d y_next / d omega)^T lam at the step (y, h): the
// reverse of tsit5_step_vjp in gsx_core.h with the stage derivatives in LDS slots 0..5. The
// forward recompute stores k_1..k_5 in slots 0..4; the reverse rebuilds each stage input
// z_s from y and slots 0..s-1 (still k's), then overwrites slot s with h * zbar_s, so
// kbar_r = b_r h lam + sum_{s>r} a_sr (h zbar_s) reads the slots above it (s = 5 down to
// r + 1, the order of the generic step VJP).
__device__ __forceinline__ void kura_step_vjp(const double* y, double h, const double* om,
                                              double* lam, double* pb, const double* op,
                                              const StageSlots& ks) {
  _Pragma("unroll 1") for (int st = 0; st <= 4; ++st) {
    double z[kSlots], kk[kSlots];
    stage_input(ks, y, h, st, z);
    FKuramotoWave::f(0.0, z, om, kk, op);
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) ks.at(st, m) = kk[m];
  }
  double hl[kSlots], lnew[kSlots];
  GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
    hl[m] = h * lam[m];
    lnew[m] = lam[m];
  }
  _Pragma("unroll 1") for (int st = 5; st >= 0; --st) {
    double kb[kSlots], z[kSlots], zb[kSlots];
    const double bw = stage_w(6, st);
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) kb[m] = bw * hl[m];
    if (st < 5) {
      double part[kSlots];
      const double w5 = stage_w(5, st);
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) part[m] = w5 * ks.at(5, m);
      _Pragma("unroll 1") for (int r = 4; r > st; --r) {
        const double w = stage_w(r, st);
        GSX_UNROLL for (int m = 0; m < kSlots; ++m) part[m] = part[m] + w * ks.at(r, m);
      }
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) kb[m] = kb[m] + part[m];
    }
    stage_input(ks, y, h, st, z);
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) zb[m] = 0.0;
    FKuramotoWave::vjp(0.0, z, om, kb, zb, pb, op);
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
      lnew[m] += zb[m];
      if (st > 0) ks.at(st, m) = h * zb[m];
    }
  }
  GSX_UNROLL for (int m = 0; m < kSlots; ++m) lam[m] = lnew[m];
}

// Adjoint with saveat cotangents. One loop with a single step-VJP site: its first nsave
// iterations take, for every save time bracketed by step s, the VJP of the re-step
// (mu = (d ys[q] / d y_s)^T ysbar[q] into the mu buffer; omega_bar accumulates); the remaining
// iterations are the reverse sweep over the recorded states, adding mu at each bracketing step.
__global__ __launch_bounds__(kKBlock) void kura_adj_kernel(KuraAdjArgs a) {
  __shared__ double op[2 * kOpTable];
  __shared__ double ks_all[6 * kSlots * kKBlock];
  fill_operands(op, a.K);
  const StageSlots ks(ks_all + threadIdx.x);
  const int64_t n = a.n, ns = a.nsave;
  const LaneId id = lane_id(n);
  const int64_t j = id.j;
  double om[kSlots], pb[kSlots], lam[kSlots];
  load_soa(a.omega, n, j, om);
  GSX_UNROLL for (int m = 0; m < kSlots; ++m) pb[m] = 0.0;
  if (a.yfbar) {
    load_soa(a.yfbar, n, j, lam);
  } else {
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) lam[m] = 0.0;
  }
  const int64_t nq = a.ysbar ? ns : 0;
  int64_t qtop = ns - 1;  // highest save index bracketed by a recorded step
  if (a.ysbar && id.valid)
    while (qtop >= 0 && a.save_step[qtop * n + j] == -2) --qtop;
  const int32_t na = id.valid ? a.n_acc[j] : 0;
  for (int64_t i = 0;; ++i) {
    const bool ph1 = i < nq;  // wavefront-uniform
    if (!ph1 && __ballot(i - nq < na) == 0ull) break;
    int32_t s;
    bool act;
    if (ph1) {
      s = id.valid ? a.save_step[i * n + j] : -1;
      act = id.valid && s >= 0;
      if (id.valid && s == -2) {
        double v[kSlots];
        load_row(a.ysbar + (j * ns + i) * kKD, v);
        GSX_UNROLL for (int m = 0; m < kSlots; ++m) lam[m] += v[m];
      }
      if (__ballot(act) == 0ull) continue;
    } else {
      s = na - 1 - static_cast<int32_t>(i - nq);
      act = s >= 0;
    }
    double y[kSlots], v[kSlots];
    double h = 0.0;
    if (act) {
      load_row(a.states + (static_cast<int64_t>(s) * n + j) * kKD, y);
      h = ph1 ? a.ts[i] - a.save_t[i * n + j] : a.tape[static_cast<int64_t>(s) * n + j];
    } else {
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) y[m] = 0.0;
    }
    if (ph1) {
      if (act) {
        load_row(a.ysbar + (j * ns + i) * kKD, v);
      } else {
        GSX_UNROLL for (int m = 0; m < kSlots; ++m) v[m] = 0.0;
      }
    } else {
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) v[m] = lam[m];
    }
    kura_step_vjp(y, h, om, v, pb, op, ks);
    if (ph1) {
      if (act) store_row(a.mu + (j * ns + i) * kKD, v);
      continue;
    }
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) lam[m] = v[m];
    if (act && a.ysbar) {
      // the saves bracketed by step s, added in increasing q as the oracle does
      int64_t lo = qtop;
      while (lo >= 0 && a.save_step[lo * n + j] == s) --lo;
      for (int64_t q = lo + 1; q <= qtop; ++q) {
        double w[kSlots];
        load_row(a.mu + (j * ns + q) * kKD, w);
        GSX_UNROLL for (int m = 0; m < kSlots; ++m) lam[m] += w[m];
      }
      qtop = lo;
    }
  }
  if (!id.valid) return;
  if (a.ysbar) {
    // saves at or before t0 are y0 itself: their cotangents, summed in increasing q
    double x[kSlots];
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) x[m] = 0.0;
    for (int64_t q = 0; q < ns; ++q) {
      if (a.save_step[q * n + j] != -1) continue;
      double v[kSlots];
      load_row(a.ysbar + (j * ns + q) * kKD, v);
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) x[m] += v[m];
    }
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) lam[m] += x[m];
  }
  store_soa(a.y0bar, n, j, lam);
  store_soa(a.omega_bar, n, j, pb);
}

unsigned grid_for(int64_t n) {
  const int64_t per_block = kKBlock / 64 * kTrajPerWave;
  return static_cast<unsigned>((n + per_block - 1) / per_block);
}

}  // namespace

int kura_fwd_launch(const KuraFwdArgs& a, hipStream_t s) {
  kura_fwd_kernel<<<grid_for(a.n), kKBlock, 0, s>>>(a);
  return check_launch("kura_fwd_kernel");
}
int kura_replay_launch(const KuraReplayArgs& a, hipStream_t s) {
  kura_replay_kernel<<<grid_for(a.n), kKBlock, 0, s>>>(a);
  return check_launch("kura_replay_kernel");
}
int kura_saveat_launch(const KuraSaveatArgs& a, hipStream_t s) {
  kura_saveat_kernel<<<grid_for(a.n), kKBlock, 0, s>>>(a);
  return check_launch("kura_saveat_kernel");
}
int kura_adj_launch(const KuraAdjArgs& a, hipStream_t s) {
  kura_adj_kernel<<<grid_for(a.n), kKBlock, 0, s>>>(a);
  return check_launch("kura_adj_kernel");
}

}  // namespace gsx
