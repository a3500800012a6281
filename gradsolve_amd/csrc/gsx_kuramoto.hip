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
//
// FP32 variant (flag GSX_FP32): the field, the stages and the step VJPs run in FP32 with the
// coupling on v_mfma_f32_16x16x4f32, whose accumulator rows are blocked (D[4(l/16) + r][l%16],
// against the FP64 instruction's interleaved D[4r + l/16][l%16]), so tile T's row rho holds K
// row 8*(rho%4) + 2*(rho/4) + T. The state and the adjoint's cotangent and parameter-gradient
// accumulators are Kahan-compensated FP32 pairs (value hi + lo); time, the step size, the
// tape, the WRMS norm and the controller stay in FP64, and the HBM buffers stay FP64 (a pair
// hi + lo is exact in FP64).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "gsx_internal.h"

namespace gsx {
namespace {

typedef double kd4 __attribute__((ext_vector_type(4)));
typedef float kf4 __attribute__((ext_vector_type(4)));

constexpr int kKD = 32;            // phases per trajectory
constexpr int kSlots = 8;          // phases per lane
constexpr int kTrajPerWave = 16;
constexpr int kKBlock = 256;       // 4 wavefronts, 64 trajectories
constexpr int kOpTable = 16 * 64;  // A operands of one product direction, [T*8 + kk][lane]
constexpr double kInvD = 1.0 / 32.0;

__device__ __forceinline__ int kdim(int g, int m) { return 8 * (m & 3) + 2 * g + (m >> 2); }

// Precision policies: the MFMA, its row permutation (see the header), sin/cos, and whether the
// state is a compensated pair. LDS-qualified element types make the stage-slot and operand
// accesses compile to ds_read/ds_write (through a generic pointer they could become FLAT
// instructions).
template <class R>
struct Prec;
template <>
struct Prec<double> {
  using acc4 = kd4;
  using lds_t = __attribute__((address_space(3))) double;
  static constexpr bool kComp = false;
  __device__ static acc4 mfma(double a, double b, acc4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int krow(int T, int rho) { return 8 * (rho >> 2) + 2 * (rho & 3) + T; }
  __device__ static void sincos(double x, double* s, double* c) { gsx_sincos(x, s, c); }
  // the branch-free evaluation and the range where it equals sincos
  __device__ static void sincos_core(double x, double* s, double* c) { gsx_sincos_core(x, s, c); }
  static constexpr double kCoreRange = 1048576.0;
};
template <>
struct Prec<float> {
  using acc4 = kf4;
  using lds_t = __attribute__((address_space(3))) float;
  static constexpr bool kComp = true;
  __device__ static acc4 mfma(float a, float b, acc4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int krow(int T, int rho) { return 8 * (rho & 3) + 2 * (rho >> 2) + T; }
  __device__ static void sincos(float x, float* s, float* c) { gsx_sincosf(x, s, c); }
  __device__ static void sincos_core(float x, float* s, float* c) { gsx_sincosf_core(x, s, c); }
  static constexpr float kCoreRange = 16384.0f;
};

// Fills the block's operand tables: op[e*64 + l] for e = T*8 + kk, forward then (kTransposed)
// transposed; the kernels that evaluate only the field fill the forward table.
template <class R, bool kTransposed = true>
__device__ void fill_operands(R* op, const double* K) {
  for (int x = threadIdx.x; x < kOpTable; x += blockDim.x) {
    const int e = x >> 6, l = x & 63;
    const int T = e >> 3, kk = e & 7;
    const int row = Prec<R>::krow(T, l & 15), col = kdim(l >> 4, kk);
    op[x] = static_cast<R>(K[row * kKD + col]);
    if (kTransposed) op[kOpTable + x] = static_cast<R>(K[col * kKD + row]);
  }
  __syncthreads();
}

// Source of the MFMA A operands: the block's tables (fill_operands). fwd(T, kk) is K's row
// krow(T, lane%16), column kdim(lane/16, kk) (tile T of the product K x, k-step kk); tr(T, kk)
// is the same entry of K^T. The forward kernels read them as ordinary loads, which the compiler
// hoists into registers where they fit (measured 2-3% faster than reads kept next to their
// MFMAs); the adjoint, whose registers are full, reads them at their use (kAtUse, volatile
// loads): hoisted there, they were spilled to per-thread memory and reloaded inside the
// products.
template <class R, bool kAtUse = false>
struct LdsOps {
  using elem_t = typename std::conditional<kAtUse, volatile typename Prec<R>::lds_t,
                                           typename Prec<R>::lds_t>::type;
  const elem_t* tv;
  int l;
  __device__ explicit LdsOps(const R* op)
      : tv((const elem_t*)op), l(static_cast<int>(__lane_id())) {}
  __device__ R fwd(int T, int kk) const { return tv[(T * 8 + kk) * 64 + l]; }
  __device__ R tr(int T, int kk) const { return tv[kOpTable + (T * 8 + kk) * 64 + l]; }
};

// a = (K b1) and b = (K b2) at the lane's 8 slots, with aop(T, kk) the A operand (a forward or
// transposed source above), where prod(kk, b1, b2) produces the B operands of k-step kk right
// before that step's MFMAs: its VALU work (sin/cos of the slot) then overlaps the MFMAs of the
// k-steps before it, which run asynchronously in the matrix pipe.
template <class R, class AOp, class Prod>
__device__ __forceinline__ void coupling_by(const AOp& aop, Prod prod, R* a, R* b) {
  using P = Prec<R>;
  typename P::acc4 as0 = {R(0), R(0), R(0), R(0)}, as1 = as0, ac0 = as0, ac1 = as0;
  // the A operands of k-step kk + 1 are read before k-step kk's B operands are formed, so a
  // source read at its use (LdsOps<R, true>) has that VALU work to arrive in
  R k0n = aop(0, 0), k1n = aop(1, 0);
  GSX_UNROLL for (int kk = 0; kk < 8; ++kk) {
    const R k0 = k0n, k1 = k1n;
    if (kk + 1 < 8) {
      k0n = aop(0, kk + 1);
      k1n = aop(1, kk + 1);
    }
    R b1, b2;
    prod(kk, b1, b2);
    as0 = P::mfma(k0, b1, as0);
    ac0 = P::mfma(k0, b2, ac0);
    as1 = P::mfma(k1, b1, as1);
    ac1 = P::mfma(k1, b2, ac1);
  }
  GSX_UNROLL for (int r = 0; r < 4; ++r) {
    a[r] = as0[r];
    a[4 + r] = as1[r];
    b[r] = ac0[r];
    b[4 + r] = ac1[r];
  }
}

// a = (K s) and b = (K c) from operands already formed.
template <class R, class AOp>
__device__ __forceinline__ void coupling(const AOp& aop, const R* s, const R* c, R* a, R* b) {
  coupling_by<R>(aop, [&](int kk, R& b1, R& b2) { b1 = s[kk]; b2 = c[kk]; }, a, b);
}

// a = (K s) and b = (K c) with s, c = sin y, cos y formed slot by slot inside the product
// loop by the branch-free evaluation; when some lane of the wavefront holds a phase outside its
// range (a wavefront-uniform test, so the MFMAs stay whole-wavefront) s, c and the products are
// formed again with the range-checked sin/cos. Either way the results are those of sincos.
template <class R, class AOp>
__device__ __forceinline__ void coupling_sincos(const AOp& aop, const R* y, R* s, R* c, R* a,
                                                R* b) {
  using P = Prec<R>;
  bool out = false;
  coupling_by<R>(aop, [&](int kk, R& b1, R& b2) {
    out = out || !(std::fabs(y[kk]) < P::kCoreRange);
    P::sincos_core(y[kk], &s[kk], &c[kk]);
    b1 = s[kk];
    b2 = c[kk];
  }, a, b);
  if (__ballot(out) != 0ull) {
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) P::sincos(y[m], &s[m], &c[m]);
    coupling<R>(aop, s, c, a, b);
  }
}

// The field on the lane's 8 slots of one trajectory; every lane of the wavefront must call it
// together (the MFMA spans the wavefront). `ops` is an operand source (LdsOps).
template <class R>
struct FKuramotoWave {
  // dy = f(y); with dg, also the Jacobian's diagonal term dg_m = a_m s_m + b_m c_m
  // (sum_j K_mj cos(y_j - y_m)), which the VJP at the same point needs.
  template <class Ops>
  __device__ static void f(const R* y, const R* om, R* dy, const Ops& ops, R* dg = nullptr) {
    R s[kSlots], c[kSlots], a[kSlots], b[kSlots];
    auto kf = [&](int T, int kk) { return ops.fwd(T, kk); };
    coupling_sincos<R>(kf, y, s, c, a, b);
    GSX_UNROLL for (int m = 0; m < kSlots; ++m)
      dy[m] = om[m] + R(kInvD) * (a[m] * c[m] - b[m] * s[m]);
    if (dg) {
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) dg[m] = a[m] * s[m] + b[m] * c[m];
    }
  }

  // yb += J^T v, pb += v: (J^T v)_m = (1/32) [c_m (K^T (v c))_m + s_m (K^T (v s))_m
  //                                            - v_m (a_m s_m + b_m c_m)].
  // dg_in: the diagonal term from f at the same y (skips the forward coupling products; sin/cos
  // of each slot are then formed inside the transposed products, as coupling_sincos does).
  template <class Ops>
  __device__ static void vjp(const R* y, const R* v, R* yb, R* pb, const Ops& ops,
                             const R* dg_in = nullptr) {
    using P = Prec<R>;
    auto kf = [&](int T, int kk) { return ops.fwd(T, kk); };
    auto kt = [&](int T, int kk) { return ops.tr(T, kk); };
    R s[kSlots], c[kSlots], dg[kSlots], vc[kSlots], vs[kSlots], u[kSlots], w[kSlots];
    if (dg_in) {
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) dg[m] = dg_in[m];
      bool out = false;
      coupling_by<R>(kt, [&](int kk, R& b1, R& b2) {
        out = out || !(std::fabs(y[kk]) < P::kCoreRange);
        P::sincos_core(y[kk], &s[kk], &c[kk]);
        vc[kk] = v[kk] * c[kk];
        vs[kk] = v[kk] * s[kk];
        b1 = vc[kk];
        b2 = vs[kk];
      }, u, w);
      if (__ballot(out) != 0ull) {
        GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
          P::sincos(y[m], &s[m], &c[m]);
          vc[m] = v[m] * c[m];
          vs[m] = v[m] * s[m];
        }
        coupling<R>(kt, vc, vs, u, w);
      }
    } else {
      R a[kSlots], b[kSlots];
      coupling_sincos<R>(kf, y, s, c, a, b);
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
        dg[m] = a[m] * s[m] + b[m] * c[m];
        vc[m] = v[m] * c[m];
        vs[m] = v[m] * s[m];
      }
      coupling<R>(kt, vc, vs, u, w);
    }
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
      yb[m] += R(kInvD) * (c[m] * u[m] + s[m] * w[m] - v[m] * dg[m]);
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

// A value of the state or of an accumulator: hi alone in FP64, the compensated pair hi + lo
// in FP32 (lo stays 0 in FP64).
template <class R>
struct Vec {
  R hi[kSlots], lo[kSlots];
  __device__ __forceinline__ void set(const double* v) {
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
      hi[m] = static_cast<R>(v[m]);
      lo[m] = Prec<R>::kComp ? static_cast<R>(v[m] - static_cast<double>(hi[m])) : R(0);
    }
  }
  __device__ __forceinline__ void zero() {
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) hi[m] = lo[m] = R(0);
  }
  __device__ __forceinline__ void get(double* v) const {
    GSX_UNROLL for (int m = 0; m < kSlots; ++m)
      v[m] = Prec<R>::kComp ? static_cast<double>(hi[m]) + static_cast<double>(lo[m])
                            : static_cast<double>(hi[m]);
  }
  // += d (R increments): Kahan in FP32, plain in FP64
  __device__ __forceinline__ void add(const R* d) {
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
      if constexpr (Prec<R>::kComp) {
        kahan_add(hi[m], lo[m], d[m]);
      } else {
        hi[m] += d[m];
      }
    }
  }
  // += v (FP64 values): both halves of the value in FP32
  __device__ __forceinline__ void add_f64(const double* v) {
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
      if constexpr (Prec<R>::kComp) {
        const float vh = static_cast<float>(v[m]);
        kahan_add(hi[m], lo[m], vh);
        kahan_add(hi[m], lo[m], static_cast<float>(v[m] - static_cast<double>(vh)));
      } else {
        hi[m] += v[m];
      }
    }
  }
};

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
// code, far past the instruction cache. Each thread reads back only its own column; the
// accesses are ordinary LDS loads and stores (address space 3), which the compiler batches
// (ds_read2, one wait per group) -- volatile accesses each waited for their own result.
template <class R>
struct StageSlots {
  using lds_t = typename Prec<R>::lds_t;
  lds_t* base;
  __device__ __forceinline__ explicit StageSlots(R* column) : base((lds_t*)column) {}
  __device__ __forceinline__ lds_t& at(int s, int m) const {
    return base[(s * kSlots + m) * kKBlock];
  }
};

// Row st of the stage weights: rows 1..5 the inputs of k2..k6 (a_{st+1, 1..st}), row 6 the
// solution weights b_1..b_6, row 7 the error weights e_1..e_7 (b_i - bhat_i).
template <class R>
__device__ __forceinline__ R stage_w(int row, int r) {
  using namespace tab;
  constexpr R W[8][7] = {{0, 0, 0, 0, 0, 0, 0},
                         {R(a21), 0, 0, 0, 0, 0, 0},
                         {R(a31), R(a32), 0, 0, 0, 0, 0},
                         {R(a41), R(a42), R(a43), 0, 0, 0, 0},
                         {R(a51), R(a52), R(a53), R(a54), 0, 0, 0},
                         {R(a61), R(a62), R(a63), R(a64), R(a65), 0, 0},
                         {R(b1), R(b2), R(b3), R(b4), R(b5), R(b6), 0},
                         {R(e1), R(e2), R(e3), R(e4), R(e5), R(e6), R(e7)}};
  return W[row][r];
}

// Input of stage S (1..5): y + h * sum_{r<S} a_{S+1,r+1} k_r over slots 0..S-1, summed left to
// right (the reference's order). S is a compile-time constant, so the slot reads issue
// together and the weights are immediates.
template <class R, int S>
__device__ __forceinline__ void stage_input_ct(const StageSlots<R>& ks, const R* y, R h, R* z) {
  GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
    R acc = stage_w<R>(S, 0) * ks.at(0, m);
    GSX_UNROLL for (int r = 1; r < S; ++r) acc = acc + stage_w<R>(S, r) * ks.at(r, m);
    z[m] = y[m] + h * acc;
  }
}

// Input of stage st (0: y itself; 1..5: stage_input_ct), dispatched on the wavefront-uniform
// stage index. In FP32, y is the compensated state's leading half.
template <class R>
__device__ __forceinline__ void stage_input(const StageSlots<R>& ks, const R* y, R h, int st,
                                            R* z) {
  switch (st) {
    case 1: stage_input_ct<R, 1>(ks, y, h, z); break;
    case 2: stage_input_ct<R, 2>(ks, y, h, z); break;
    case 3: stage_input_ct<R, 3>(ks, y, h, z); break;
    case 4: stage_input_ct<R, 4>(ks, y, h, z); break;
    case 5: stage_input_ct<R, 5>(ks, y, h, z); break;
    default: GSX_UNROLL for (int m = 0; m < kSlots; ++m) z[m] = y[m];
  }
}

// The stages of one step in four LDS slots per lane (the forward, replay and saveat kernels;
// with a slot per stage a 256-lane block took 128 KiB of LDS: one block, so one wavefront, per
// SIMD). S0 holds k1 for the whole step (the FSAL stage: a rejected trial still needs it); S1..S3
// take k2..k4. Once k4 is formed, its stage input z5 = y + h P5 goes to registers and S1..S3
// become the running sums P6 = sum a6r k_r, Y = sum b_r k_r and (kErr) E = sum e_r k_r over
// r <= 4; k5, k6 and k7 are added to them as they are formed. Every sum keeps the
// left-to-right order of stage_input (the reference's order), so the results are those of a
// slot per stage, bit for bit. first = 0 forms k1 = f(y) into S0; kErr also forms k7 = f(y5)
// into S1 (free once z6 is formed; the caller copies it to S0 on acceptance) and leaves E in
// S3. y5 and, in FP32, the increment inc = h Y are formed as the stage sums form them. The
// stage loop is not unrolled (one copy of the field); its branches are on the
// wavefront-uniform stage index.
template <class R, bool kErr, class Ops>
__device__ __forceinline__ void compact_stages(const StageSlots<R>& ks, const R* y, R h,
                                               const R* om, const Ops& op, int first, R* y5,
                                               R* inc) {
  constexpr int kLast = kErr ? 6 : 5;
  _Pragma("unroll 1") for (int st = first; st <= kLast; ++st) {
    R z[kSlots], kk[kSlots];
    if (st <= 3) {
      stage_input<R>(ks, y, h, st, z);
    } else if (st == 4) {
      // k1..k4 -> z5 (registers), P6 -> S1, Y -> S2, E -> S3
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
        const R k1 = ks.at(0, m), k2 = ks.at(1, m), k3 = ks.at(2, m), k4 = ks.at(3, m);
        R p5 = stage_w<R>(4, 0) * k1;
        p5 = p5 + stage_w<R>(4, 1) * k2;
        p5 = p5 + stage_w<R>(4, 2) * k3;
        p5 = p5 + stage_w<R>(4, 3) * k4;
        R p6 = stage_w<R>(5, 0) * k1;
        p6 = p6 + stage_w<R>(5, 1) * k2;
        p6 = p6 + stage_w<R>(5, 2) * k3;
        p6 = p6 + stage_w<R>(5, 3) * k4;
        R yb = stage_w<R>(6, 0) * k1;
        yb = yb + stage_w<R>(6, 1) * k2;
        yb = yb + stage_w<R>(6, 2) * k3;
        yb = yb + stage_w<R>(6, 3) * k4;
        z[m] = y[m] + h * p5;
        ks.at(1, m) = p6;
        ks.at(2, m) = yb;
        if (kErr) {
          R eb = stage_w<R>(7, 0) * k1;
          eb = eb + stage_w<R>(7, 1) * k2;
          eb = eb + stage_w<R>(7, 2) * k3;
          eb = eb + stage_w<R>(7, 3) * k4;
          ks.at(3, m) = eb;
        }
      }
    } else if (st == 5) {
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) z[m] = y[m] + h * ks.at(1, m);
    } else {
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
        const R acc = ks.at(2, m);
        if constexpr (Prec<R>::kComp) {
          inc[m] = h * acc;
          y5[m] = y[m] + inc[m];
        } else {
          y5[m] = y[m] + h * acc;
        }
        z[m] = y5[m];
      }
    }
    FKuramotoWave<R>::f(z, om, kk, op);
    if (st <= 3) {
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) ks.at(st, m) = kk[m];
    } else if (st == 4) {
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
        ks.at(1, m) = ks.at(1, m) + stage_w<R>(5, 4) * kk[m];
        ks.at(2, m) = ks.at(2, m) + stage_w<R>(6, 4) * kk[m];
        if (kErr) ks.at(3, m) = ks.at(3, m) + stage_w<R>(7, 4) * kk[m];
      }
    } else if (st == 5) {
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
        ks.at(2, m) = ks.at(2, m) + stage_w<R>(6, 5) * kk[m];
        if (kErr) ks.at(3, m) = ks.at(3, m) + stage_w<R>(7, 5) * kk[m];
      }
    } else {
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
        ks.at(3, m) = ks.at(3, m) + stage_w<R>(7, 6) * kk[m];
        ks.at(1, m) = kk[m];
      }
    }
  }
  if constexpr (!kErr) {
    // the step's result from Y (S2)
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
      const R acc = ks.at(2, m);
      if constexpr (Prec<R>::kComp) {
        inc[m] = h * acc;
        y5[m] = y[m] + inc[m];
      } else {
        y5[m] = y[m] + h * acc;
      }
    }
  }
}

// One replay step of the state (k7 is not formed: its weight is 0).
template <class R, class Ops>
__device__ __forceinline__ void kura_advance(Vec<R>& y, R h, const R* om, const Ops& op,
                                              const StageSlots<R>& ks) {
  R y5[kSlots], inc[kSlots];
  compact_stages<R, false>(ks, y.hi, h, om, op, 0, y5, inc);
  if constexpr (Prec<R>::kComp) {
    y.add(inc);
  } else {
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) y.hi[m] = y5[m];
  }
}

// Forward + record (gradsolve/solvers/tsit5_replay.py:100-124 per trajectory).
template <class R>
__global__ __launch_bounds__(kKBlock) __attribute__((amdgpu_waves_per_eu(2)))
void kura_fwd_kernel(KuraFwdArgs a) {
  __shared__ R op[kOpTable];
  __shared__ R ks_all[4 * kSlots * kKBlock];
  fill_operands<R, false>(op, a.K);
  const LdsOps<R> ops(op);
  const StageSlots<R> ks(ks_all + threadIdx.x);
  const int64_t n = a.n;
  const LaneId id = lane_id(n);
  const int64_t j = id.j;
  Vec<R> y;
  R om[kSlots], y5[kSlots], inc[kSlots];
  {
    double v[kSlots];
    load_soa(a.y0, n, j, v);
    y.set(v);
    load_soa(a.omega, n, j, v);
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) om[m] = static_cast<R>(v[m]);
  }
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
    const R hr = static_cast<R>(h);
    compact_stages<R, true>(ks, y.hi, hr, om, ops, first, y5, inc);
    first = 1;
    double e[kSlots], yd[kSlots], y5d[kSlots];
    {
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
        e[m] = static_cast<double>(hr * ks.at(3, m));
        y5d[m] = static_cast<double>(y5[m]);
      }
      y.get(yd);
    }
    const double err = kura_wrms(e, yd, y5d, a.rtol, a.atol);
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
      if (a.states) store_row(a.states + (static_cast<int64_t>(acc) * n + j) * kKD, yd);
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
      if constexpr (Prec<R>::kComp) {
        y.add(inc);
      } else {
        GSX_UNROLL for (int m = 0; m < kSlots; ++m) y.hi[m] = y5[m];
      }
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) ks.at(0, m) = ks.at(1, m);  // k1 <- k7
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
  {
    double v[kSlots];
    y.get(v);
    store_soa(a.y_final, n, j, v);
  }
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
template <class R>
__global__ __launch_bounds__(kKBlock) __attribute__((amdgpu_waves_per_eu(2)))
void kura_replay_kernel(KuraReplayArgs a) {
  __shared__ R op[kOpTable];
  __shared__ R ks_all[4 * kSlots * kKBlock];
  fill_operands<R, false>(op, a.K);
  const LdsOps<R> ops(op);
  const StageSlots<R> ks(ks_all + threadIdx.x);
  const int64_t n = a.n;
  const LaneId id = lane_id(n);
  const int64_t j = id.j;
  Vec<R> y;
  R om[kSlots];
  {
    double v[kSlots];
    load_soa(a.y0, n, j, v);
    y.set(v);
    load_soa(a.omega, n, j, v);
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) om[m] = static_cast<R>(v[m]);
  }
  const int32_t na = id.valid ? a.n_acc[j] : 0;
  for (int32_t s = 0; __ballot(s < na) != 0ull; ++s) {
    const bool act = s < na;
    const double h = act ? a.tape[static_cast<int64_t>(s) * n + j] : 0.0;
    if (act && a.states) {
      double v[kSlots];
      y.get(v);
      store_row(a.states + (static_cast<int64_t>(s) * n + j) * kKD, v);
    }
    Vec<R> yn = y;
    kura_advance<R>(yn, static_cast<R>(h), om, ops, ks);
    if (act) y = yn;
  }
  if (id.valid) {
    double v[kSlots];
    y.get(v);
    store_soa(a.y_final, n, j, v);
  }
}

// ys[j][q] = one step of size ts[q] - t_s from the bracketing step's start state.
template <class R>
__global__ __launch_bounds__(kKBlock) __attribute__((amdgpu_waves_per_eu(2)))
void kura_saveat_kernel(KuraSaveatArgs a) {
  __shared__ R op[kOpTable];
  __shared__ R ks_all[4 * kSlots * kKBlock];
  fill_operands<R, false>(op, a.K);
  const LdsOps<R> ops(op);
  const StageSlots<R> ks(ks_all + threadIdx.x);
  const int64_t n = a.n, ns = a.nsave;
  const LaneId id = lane_id(n);
  const int64_t j = id.j;
  double y0[kSlots], yf[kSlots];
  R om[kSlots];
  load_soa(a.y0, n, j, y0);
  load_soa(a.y_final, n, j, yf);
  {
    double v[kSlots];
    load_soa(a.omega, n, j, v);
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) om[m] = static_cast<R>(v[m]);
  }
  for (int64_t q = 0; q < ns; ++q) {
    const int32_t s = id.valid ? a.save_step[q * n + j] : -1;
    const bool need = id.valid && s >= 0;
    double* row = a.ys + (j * ns + q) * kKD;
    if (__ballot(need) != 0ull) {
      double v[kSlots];
      double h = 0.0;
      if (need) {
        load_row(a.states + (static_cast<int64_t>(s) * n + j) * kKD, v);
        h = a.ts[q] - a.save_t[q * n + j];
      } else {
        GSX_UNROLL for (int m = 0; m < kSlots; ++m) v[m] = y0[m];
      }
      Vec<R> y;
      y.set(v);
      kura_advance<R>(y, static_cast<R>(h), om, ops, ks);
      y.get(v);
      if (need) store_row(row, v);
    }
    if (id.valid && s == -1) store_row(row, y0);
    if (id.valid && s == -2) store_row(row, yf);
  }
}

// The later stages' share of kbar_St (St <= 4) in the step VJP below:
// sum_{s=6..St+2} a_{s,St+1} (h zbar_s) with h zbar_s in slot s - 2, then a_{St+2,St+1}
// (h zbar_{St+1}) from the carried registers, summed in that order. St is a compile-time
// constant (batched slot reads, immediate weights).
template <class R, int St>
__device__ __forceinline__ void stage_cot_ct(const StageSlots<R>& ks, const R* carry, R* part) {
  GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
    R acc = stage_w<R>(5, St) * (St == 4 ? carry[m] : ks.at(3, m));
    GSX_UNROLL for (int r = 4; r > St + 1; --r) acc = acc + stage_w<R>(r, St) * ks.at(r - 2, m);
    if (St < 4) acc = acc + stage_w<R>(St + 1, St) * carry[m];
    part[m] = acc;
  }
}

// lam <- (d y_next / d y)^T lam, pb += (d y_next / d omega)^T lam at the step (y, h): the
// reverse of tsit5_step_vjp in gsx_core.h with the stage derivatives in four LDS slots. The
// forward recompute stores k_1..k_4 in slots 0..3; k_5 only enters the stage-6 input
// z_6 = y + h sum_{r<=5} a_6r k_r, which goes to per-thread memory beside the diagonal terms.
// The reverse rebuilds each stage input z_s (s <= 5) from y and slots 0..s-2, after which
// k_{s-1} in slot s - 2 is dead and takes h zbar_{s+1}, carried in registers from the VJP
// before (not live across the coupling products). So kbar_r = b_r h lam + sum_{s>r} a_sr
// (h zbar_s) reads h zbar_s from slot s - 2 and the newest, s = r + 1, from the carried
// registers; the terms are summed from s = 6 down to r + 1, the order of the generic step
// VJP. In FP32 the step's contributions to lam and pb are summed separately and added to the
// compensated pairs once.
template <class R, class Ops>
__device__ __forceinline__ void kura_step_vjp(const R* y, R h, const R* om, Vec<R>& lam,
                                              Vec<R>& pb, const Ops& op,
                                              const StageSlots<R>& ks) {
  // The recompute keeps each stage's diagonal term dg (per-thread memory: the stage index is
  // dynamic), so the VJPs of stages 1..5 run only the transposed products: 384 instead of 544
  // MFMAs per step. dgs[5] holds z_6.
  R dgs[6 * kSlots];
  _Pragma("unroll 1") for (int st = 0; st <= 4; ++st) {
    R z[kSlots], kk[kSlots];
    stage_input<R>(ks, y, h, st, z);
    FKuramotoWave<R>::f(z, om, kk, op, dgs + st * kSlots);
    if (st < 4) {
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) ks.at(st, m) = kk[m];
    } else {
      // z_6 in stage_input's left-to-right order
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
        R acc = stage_w<R>(5, 0) * ks.at(0, m);
        acc = acc + stage_w<R>(5, 1) * ks.at(1, m);
        acc = acc + stage_w<R>(5, 2) * ks.at(2, m);
        acc = acc + stage_w<R>(5, 3) * ks.at(3, m);
        acc = acc + stage_w<R>(5, 4) * kk[m];
        dgs[5 * kSlots + m] = y[m] + h * acc;
      }
    }
  }
  constexpr bool comp = Prec<R>::kComp;
  R hl[kSlots], dl[kSlots], dp[kSlots], carry[kSlots];
  GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
    hl[m] = h * lam.hi[m];
    dl[m] = comp ? R(0) : lam.hi[m];
    dp[m] = R(0);
    carry[m] = R(0);
  }
  _Pragma("unroll 1") for (int st = 5; st >= 0; --st) {
    R kb[kSlots], z[kSlots], zb[kSlots];
    if (st == 5) {
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) z[m] = dgs[5 * kSlots + m];
    } else {
      stage_input<R>(ks, y, h, st, z);
    }
    const R bw = stage_w<R>(6, st);
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) kb[m] = bw * hl[m];
    if (st < 5) {
      R part[kSlots];
      switch (st) {
        case 0: stage_cot_ct<R, 0>(ks, carry, part); break;
        case 1: stage_cot_ct<R, 1>(ks, carry, part); break;
        case 2: stage_cot_ct<R, 2>(ks, carry, part); break;
        case 3: stage_cot_ct<R, 3>(ks, carry, part); break;
        default: stage_cot_ct<R, 4>(ks, carry, part);
      }
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) kb[m] = kb[m] + part[m];
      // z_st is formed, so k_{st-1} in slot st - 1 is dead: it takes h zbar_{st+1}
      if (st >= 1) {
        GSX_UNROLL for (int m = 0; m < kSlots; ++m) ks.at(st - 1, m) = carry[m];
      }
    }
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) zb[m] = R(0);
    FKuramotoWave<R>::vjp(z, kb, zb, comp ? dp : pb.hi, op, st < 5 ? dgs + st * kSlots : nullptr);
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) {
      dl[m] += zb[m];
      carry[m] = h * zb[m];
    }
  }
  if constexpr (comp) {
    lam.add(dl);
    pb.add(dp);
  } else {
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) lam.hi[m] = dl[m];
  }
}

// Adjoint with saveat cotangents. One loop with a single step-VJP site: its first nsave
// iterations take, for every save time bracketed by step s, the VJP of the re-step
// (mu = (d ys[q] / d y_s)^T ysbar[q] into the mu buffer; omega_bar accumulates); the remaining
// iterations are the reverse sweep over the recorded states, adding mu at each bracketing step.
// Its LDS holds the four stage slots and both operand tables: 80 KiB per 256-lane block in
// FP64, so two blocks, two wavefronts per SIMD, fit.
template <class R>
__global__ __launch_bounds__(kKBlock) __attribute__((amdgpu_waves_per_eu(2)))
void kura_adj_kernel(KuraAdjArgs a) {
  __shared__ R ks_all[4 * kSlots * kKBlock];
  __shared__ R op_tab[2 * kOpTable];
  fill_operands<R>(op_tab, a.K);
  const LdsOps<R, true> op(op_tab);
  const StageSlots<R> ks(ks_all + threadIdx.x);
  const int64_t n = a.n, ns = a.nsave;
  const LaneId id = lane_id(n);
  const int64_t j = id.j;
  R om[kSlots];
  Vec<R> pb, lam;
  {
    double v[kSlots];
    load_soa(a.omega, n, j, v);
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) om[m] = static_cast<R>(v[m]);
  }
  pb.zero();
  if (a.yfbar) {
    double v[kSlots];
    load_soa(a.yfbar, n, j, v);
    lam.set(v);
  } else {
    lam.zero();
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
        lam.add_f64(v);
      }
      if (__ballot(act) == 0ull) continue;
    } else {
      s = na - 1 - static_cast<int32_t>(i - nq);
      act = s >= 0;
    }
    R y[kSlots];
    Vec<R> v;
    double h = 0.0;
    if (act) {
      double st[kSlots];
      load_row(a.states + (static_cast<int64_t>(s) * n + j) * kKD, st);
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) y[m] = static_cast<R>(st[m]);
      h = ph1 ? a.ts[i] - a.save_t[i * n + j] : a.tape[static_cast<int64_t>(s) * n + j];
    } else {
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) y[m] = R(0);
    }
    if (ph1) {
      if (act) {
        double w[kSlots];
        load_row(a.ysbar + (j * ns + i) * kKD, w);
        v.set(w);
      } else {
        v.zero();
      }
    } else {
      v = lam;
    }
    // The reverse sweep reads the recorded states backwards: one word per 64 bytes of the next
    // step's state rows and tape entries is read now (volatile, so it stays here), which brings
    // those lines into the cache while this step is differentiated; the next iteration's loads
    // then do not wait on HBM. The words are only kept alive, never used.
    uint32_t pf_row = 0, pf_h = 0;
    if (!ph1 && s >= 1) {
      const int64_t sn = static_cast<int64_t>(s - 1) * n + j;
      const int g = static_cast<int>(__lane_id()) >> 4;
      pf_row = *reinterpret_cast<const volatile uint32_t*>(a.states + sn * kKD + 8 * g);
      pf_h = *reinterpret_cast<const volatile uint32_t*>(a.tape + sn);
    }
    kura_step_vjp<R>(y, static_cast<R>(h), om, v, pb, op, ks);
    asm volatile("" ::"v"(pf_row), "v"(pf_h));
    if (ph1) {
      if (act) {
        double w[kSlots];
        v.get(w);
        store_row(a.mu + (j * ns + i) * kKD, w);
      }
      continue;
    }
    lam = v;
    if (act && a.ysbar) {
      // the saves bracketed by step s, added in increasing q as the oracle does
      int64_t lo = qtop;
      while (lo >= 0 && a.save_step[lo * n + j] == s) --lo;
      for (int64_t q = lo + 1; q <= qtop; ++q) {
        double w[kSlots];
        load_row(a.mu + (j * ns + q) * kKD, w);
        lam.add_f64(w);
      }
      qtop = lo;
    }
  }
  if (!id.valid) return;
  double out[kSlots];
  lam.get(out);
  if (a.ysbar) {
    // saves at or before t0 are y0 itself: their cotangents, summed in increasing q
    double x[kSlots];
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) x[m] = 0.0;
    for (int64_t q = 0; q < ns; ++q) {
      if (a.save_step[q * n + j] != -1) continue;
      double w[kSlots];
      load_row(a.ysbar + (j * ns + q) * kKD, w);
      GSX_UNROLL for (int m = 0; m < kSlots; ++m) x[m] += w[m];
    }
    GSX_UNROLL for (int m = 0; m < kSlots; ++m) out[m] += x[m];
  }
  store_soa(a.y0bar, n, j, out);
  pb.get(out);
  store_soa(a.omega_bar, n, j, out);
}

unsigned grid_for(int64_t n) {
  const int64_t per_block = kKBlock / 64 * kTrajPerWave;
  return static_cast<unsigned>((n + per_block - 1) / per_block);
}

}  // namespace

int kura_fwd_launch(const KuraFwdArgs& a, bool fp32, hipStream_t s) {
  if (fp32) {
    kura_fwd_kernel<float><<<grid_for(a.n), kKBlock, 0, s>>>(a);
  } else {
    kura_fwd_kernel<double><<<grid_for(a.n), kKBlock, 0, s>>>(a);
  }
  return check_launch("kura_fwd_kernel");
}
int kura_replay_launch(const KuraReplayArgs& a, bool fp32, hipStream_t s) {
  if (fp32) {
    kura_replay_kernel<float><<<grid_for(a.n), kKBlock, 0, s>>>(a);
  } else {
    kura_replay_kernel<double><<<grid_for(a.n), kKBlock, 0, s>>>(a);
  }
  return check_launch("kura_replay_kernel");
}
int kura_saveat_launch(const KuraSaveatArgs& a, bool fp32, hipStream_t s) {
  if (fp32) {
    kura_saveat_kernel<float><<<grid_for(a.n), kKBlock, 0, s>>>(a);
  } else {
    kura_saveat_kernel<double><<<grid_for(a.n), kKBlock, 0, s>>>(a);
  }
  return check_launch("kura_saveat_kernel");
}
int kura_adj_launch(const KuraAdjArgs& a, bool fp32, hipStream_t s) {
  if (fp32) {
    kura_adj_kernel<float><<<grid_for(a.n), kKBlock, 0, s>>>(a);
  } else {
    kura_adj_kernel<double><<<grid_for(a.n), kKBlock, 0, s>>>(a);
  }
  return check_launch("kura_adj_kernel");
}

}  // namespace gsx
