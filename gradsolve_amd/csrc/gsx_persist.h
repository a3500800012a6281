// gsx_persist.h -- persistent, work-queue forms of the Tsit5 forward+record and adjoint for
// ensembles whose trajectories need very different step counts (a per-trajectory t1, wide
// parameter ranges). Compiled in GSX_NS.
//
// Forward. A fixed grid of L lanes loops until the queue is empty. Every iteration each busy
// lane takes one trial step of its trajectory; lanes that have finished (or not started) take
// the next trajectory indices from their wavefront's claimed range of the queue. A wavefront
// claims kQueueChunk indices with one atomic on the global counter when its range runs out,
// so the counter sees one atomic per kQueueChunk trajectories (one device-scope atomic per
// refill put all 3072 wavefronts in line behind one address). A wavefront's lanes stay busy
// even when their trajectories' lengths differ by 20x.
//
// Record layout. Each lane appends to its own streams:
//   rec   [rec_cap][L][R]          segment records: record r of lane l holds, at
//                                  ((r * L) + l) * R, the state at the start of one
//                                  checkpoint segment (DCK doubles; in FP32 the pairs
//                                  (s, lo), each in the bytes of one double) followed by the
//                                  step sizes of the segment's C accepted steps. R is DCK + C
//                                  rounded up to 8 doubles, so a record is whole 64-byte
//                                  sectors and the adjoint reads each segment's checkpoint
//                                  and steps with one contiguous access.
//   order [traj_cap][L]            the trajectories a lane processed, in order
// and per trajectory j: lane_of[j], rec_start[j] (its first record in that lane's stream),
// n_acc[j], n_rej[j], status[j]. Trajectory j's accepted step s is slot DCK + s % C of record
// rec_start[j] + s / C.
//
// Why records. Lanes advance at their own pace, so the lanes of a wavefront write and read
// different positions of their streams: every per-lane access is a separate memory
// transaction. A 24-byte checkpoint or an 8-byte step size each touching its own line cost
// 3-6x their size in HBM traffic (2^22 uneven Lorenz, C = 2: 46 GB written by the forward
// and 49 GB fetched by the adjoint for 15 GB of data, the adjoint at 48% VALU busy). A record
// is written within C accepted steps, while its line is still in L2, and read whole.
//
// Writing records. For records of up to kStageMaxLen doubles the forward assembles each
// lane's open record in LDS and stores it whole (16-byte stores) once its C steps are in or its
// trajectory ends; a record written piecemeal over C steps is evicted from L2 half-written, and
// partial-sector writes to ECC-protected HBM cost a read-modify-write each (2^22 uneven Lorenz,
// C = 2: forward at 37% VALU busy). Longer records (C = 16, 32) are written in place: their
// LDS would not fit four 256-lane blocks per CU.
//
// Adjoint. Lane l walks its own trajectory list backwards, one record (checkpoint segment:
// re-integrate into LDS, then reverse; at C = 2 the pair adjoint) per iteration, so
// per-iteration work is uniform across lanes.
#pragma once

#include "gsx_core.h"
#include "gsx_f32.h"

namespace gsx {

constexpr unsigned long long kQueueChunk = 64;  // queue indices a wavefront claims per atomic
constexpr int kStageMaxLen = 16;                // longest record assembled in LDS (doubles)

// Doubles per segment record: the checkpoint rows and C step sizes, in whole 64-byte sectors.
GSX_HD constexpr int persist_record_len(int dck, int C) { return (dck + C + 7) / 8 * 8; }

// First double of record r of lane `lane`.
GSX_HD int64_t persist_record(int64_t r, int64_t lane, int64_t L, int R) {
  return (r * L + lane) * R;
}

}  // namespace gsx

namespace GSX_NS {

using ::gsx::kQueueChunk;
using ::gsx::persist_record;
using ::gsx::PersistAdjArgs;
using ::gsx::PersistFwdArgs;

// Per-lane integration state of the FP64 and FP32 variants.
template <class F>
struct StateF64 {
  static constexpr int D = F::D, P = F::P;
  double y[D], y5[D], p[P];
  double k[7][D];
  GSX_HD void load(const FwdArgs& a, int64_t j) {
    GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = a.y0[d * a.n + j];
    GSX_UNROLL for (int q = 0; q < P; ++q) p[q] = a.p[q * a.n + j];
    F::f(0.0, y, p, k[0], a.aux);
  }
  GSX_HD double trial(double h, double rtol, double atol, const double* aux) {
    using namespace tab;
    tsit5_stages<F, double>(0.0, y, h, p, k, y5, aux);
    F::f(0.0, y5, p, k[6], aux);
    double e[D];
    GSX_UNROLL for (int i = 0; i < D; ++i)
      e[i] = h * (e1 * k[0][i] + e2 * k[1][i] + e3 * k[2][i] + e4 * k[3][i] + e5 * k[4][i] +
                  e6 * k[5][i] + e7 * k[6][i]);
    return tsit5_err<D>(e, y, y5, rtol, atol);
  }
  GSX_HD void accept() {
    GSX_UNROLL for (int d = 0; d < D; ++d) {
      y[d] = y5[d];
      k[0][d] = k[6][d];
    }
  }
  // The checkpoint rows at ck[0], ck[stride], ...
  GSX_HD void store_ckpt(double* ck, int64_t stride) const {
    GSX_UNROLL for (int d = 0; d < D; ++d) ck[d * stride] = y[d];
  }
  GSX_HD void write_final(const FwdArgs& a, int64_t j) const {
    GSX_UNROLL for (int d = 0; d < D; ++d) a.y_final[d * a.n + j] = y[d];
  }
};

template <class F>
struct StateF32 {
  static constexpr int D = F::D, P = F::P;
  float y[D], lo[D], y5[D], inc[D], p[P];
  float k[7][D];
  GSX_HD void load(const FwdArgs& a, int64_t j) {
    GSX_UNROLL for (int d = 0; d < D; ++d) {
      const double v = a.y0[d * a.n + j];
      y[d] = static_cast<float>(v);
      lo[d] = static_cast<float>(v - static_cast<double>(y[d]));
    }
    GSX_UNROLL for (int q = 0; q < P; ++q) p[q] = static_cast<float>(a.p[q * a.n + j]);
    F::f(0.0f, y, p, k[0], a.aux);
  }
  GSX_HD double trial(double h, double rtol, double atol, const double* aux) {
    using namespace tab;
    const float hf = static_cast<float>(h);
    tsit5_stages<F, float>(0.0f, y, hf, p, k, y5, aux, inc);
    F::f(0.0f, y5, p, k[6], aux);
    double e[D], yd[D], y5d[D];
    GSX_UNROLL for (int i = 0; i < D; ++i) {
      e[i] = static_cast<double>(hf * (float(e1) * k[0][i] + float(e2) * k[1][i] +
                                       float(e3) * k[2][i] + float(e4) * k[3][i] +
                                       float(e5) * k[4][i] + float(e6) * k[5][i] +
                                       float(e7) * k[6][i]));
      yd[i] = static_cast<double>(y[i]) + static_cast<double>(lo[i]);
      y5d[i] = static_cast<double>(y5[i]);
    }
    return tsit5_err<D>(e, yd, y5d, rtol, atol);
  }
  GSX_HD void accept() {
    GSX_UNROLL for (int d = 0; d < D; ++d) {
      kahan_add(y[d], lo[d], inc[d]);
      k[0][d] = k[6][d];
    }
  }
  // An FP32 checkpoint row pair (s, lo) occupies the bytes of one FP64 row.
  GSX_HD void store_ckpt(double* ck, int64_t stride) const {
    GSX_UNROLL for (int d = 0; d < D; ++d) {
      float* cell = reinterpret_cast<float*>(ck + d * stride);
      cell[0] = y[d];
      cell[1] = lo[d];
    }
  }
  GSX_HD void write_final(const FwdArgs& a, int64_t j) const {
    GSX_UNROLL for (int d = 0; d < D; ++d)
      a.y_final[d * a.n + j] = static_cast<double>(y[d]) + static_cast<double>(lo[d]);
  }
};

}  // namespace GSX_NS
