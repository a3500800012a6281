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
//   tape  [lane_cap/16][L][16]     accepted step sizes; a lane's steps 16k..16k+15 are one
//                                  128-byte line, so lanes whose positions have drifted apart
//                                  still write and read whole lines
//   ckpt  [ckpt_cap][L][DCK]       trajectory start states and every C-th accepted state;
//                                  one checkpoint's rows are contiguous
//   order [traj_cap][L]            the trajectories a lane processed, in order
// and per trajectory j: lane[j], tape_start[j], ckpt_start[j] (positions in that lane's
// streams), n_acc[j], n_rej[j], status[j].
//
// Adjoint. Lane l walks its own trajectory list backwards, one checkpoint segment (C accepted
// steps: re-integrate into LDS, then reverse) per iteration, so per-iteration work is uniform
// across lanes.
#pragma once

#include "gsx_core.h"
#include "gsx_f32.h"

namespace gsx {

constexpr int kTapeBlock = 16;                 // tape entries of one lane per 128-byte line
constexpr unsigned long long kQueueChunk = 64;  // queue indices a wavefront claims per atomic

// Position pos of lane `lane`'s tape stream in the [lane_cap/16][L][16] array.
GSX_HD int64_t persist_tape_index(int64_t pos, int64_t lane, int64_t L) {
  return ((pos >> 4) * L + lane) * kTapeBlock + (pos & (kTapeBlock - 1));
}

}  // namespace gsx

namespace GSX_NS {

using ::gsx::kQueueChunk;
using ::gsx::persist_tape_index;
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
  GSX_HD void store_ckpt(double* ck, int64_t row0, int64_t L) const {
    GSX_UNROLL for (int d = 0; d < D; ++d) ck[(row0 + d) * L] = y[d];
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
  GSX_HD void store_ckpt(double* ck, int64_t row0, int64_t L) const {
    GSX_UNROLL for (int d = 0; d < D; ++d) {
      float* cell = reinterpret_cast<float*>(ck + (row0 + d) * L);
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
