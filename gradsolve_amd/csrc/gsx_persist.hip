// gsx_persist.hip -- persistent work-queue Tsit5 forward+record and its adjoint (gfx950).
// Layout and scheduling are described in gsx_persist.h.
#include <hip/hip_runtime.h>

#include "gsx_internal.h"
#include "gsx_persist.h"

namespace GSX_NS {

using ::gsx::check_launch;
using ::gsx::kMaxCkptInterval;
using ::gsx::set_error;

namespace {

constexpr int kPBlock = 256;
constexpr int kPAdjBlock = 64;

__global__ void queue_reset_kernel(unsigned long long* q) { *q = 0ull; }

// Held to 128 VGPRs (four waves per SIMD): the Lorenz FP64 lane otherwise takes 129 and would
// cap the persistent grid at three waves.
// kStage: the open record is assembled in LDS (dynamic shared memory of rec_len * kPBlock
// doubles, element i of this lane's record at [i * kPBlock + threadIdx.x]) and stored whole.
template <class F, class State, bool kStage>
__global__ __launch_bounds__(kPBlock) __attribute__((amdgpu_waves_per_eu(4)))
void persist_fwd_kernel(PersistFwdArgs pa) {
  static_assert(F::autonomous, "the persistent engine is built for autonomous fields");
  constexpr int DCK = ckpt_rows<F>();
  extern __shared__ double stage_all[];
  double* const stg = stage_all + threadIdx.x;
  const FwdArgs& a = pa.f;
  const int64_t L = pa.lanes;
  const int R = pa.rec_len;
  const int64_t lane = static_cast<int64_t>(blockIdx.x) * kPBlock + threadIdx.x;
  const int wl = static_cast<int>(__lane_id());
  const int32_t C = a.ckpt_interval;
  const unsigned long long n = static_cast<unsigned long long>(a.n);
  State st;
  bool active = false, done = false;
  int64_t j = 0;
  double t = 0.0, t1 = 0.0, dt = 0.0;
  int32_t acc = 0, rej = 0;
  int64_t it = 0;
  int32_t rcursor = 0, tcount = 0;
  int32_t seg_left = 0;   // free step slots of the open record
  bool rec_open = false;
  double* rp = nullptr;   // the open record in HBM
  double* hp = nullptr;   // its next step slot (in LDS when staged)
  // The wavefront's claimed queue range [cbase, cend) and whether the queue has run dry
  // (wavefront-uniform).
  unsigned long long cbase = 0ull, cend = 0ull;
  bool drained = false;

  // Stores the staged record whole (a partial one when its trajectory ended early).
  auto close_record = [&]() {
    if constexpr (kStage) {
      for (int i = 0; i < R; i += 2)
        *reinterpret_cast<double2*>(rp + i) = make_double2(stg[i * kPBlock],
                                                           stg[(i + 1) * kPBlock]);
    }
    rec_open = false;
  };
  auto finish = [&](int32_t status) {
    if (rec_open) close_record();
    st.write_final(a, j);
    a.n_acc[j] = acc;
    a.n_rej[j] = rej;
    a.status[j] = status;
    active = false;
  };
  // Opens the lane's next record at the current state (false: the stream is full).
  auto open_record = [&]() -> bool {
    if (rcursor >= pa.rec_cap) return false;
    rp = pa.rec + persist_record(rcursor, lane, L, R);
    if constexpr (kStage) {
      st.store_ckpt(stg, kPBlock);
      hp = stg + DCK * kPBlock;
    } else {
      st.store_ckpt(rp, 1);
      hp = rp + DCK;
    }
    ++rcursor;
    seg_left = C;
    rec_open = true;
    return true;
  };

  while (true) {
    // Refill the idle lanes from the claimed range; claim kQueueChunk more indices (one
    // atomic for the wavefront) when the range holds fewer than the idle lanes.
    const bool need = !active && !done;
    const unsigned long long mask = __ballot(need);
    if (mask) {
      const unsigned long long cnt = static_cast<unsigned long long>(__popcll(mask));
      const unsigned long long avail = cend - cbase;
      unsigned long long nb = 0ull, ne = 0ull;
      if (avail < cnt && !drained) {
        const int leader = __ffsll(static_cast<long long>(mask)) - 1;
        if (wl == leader) nb = atomicAdd(pa.queue, kQueueChunk);
        nb = __shfl(nb, leader);
        ne = nb < n ? (nb + kQueueChunk < n ? nb + kQueueChunk : n) : nb;
        drained = (ne == nb);
      }
      if (need) {
        const unsigned long long r =
            static_cast<unsigned long long>(__popcll(mask & ((1ull << wl) - 1ull)));
        bool got = true;
        unsigned long long jj = 0ull;
        if (r < avail) {
          jj = cbase + r;
        } else if (r - avail < ne - nb) {
          jj = nb + (r - avail);
        } else {
          got = false;
        }
        if (got && pa.queue_order) {
          // A planned order (e.g. longest first); an entry outside [0, n) hands out no work.
          const int32_t jo = pa.queue_order[jj];
          got = jo >= 0 && static_cast<unsigned long long>(jo) < n;
          jj = static_cast<unsigned long long>(jo);
        }
        if (got) {
          j = static_cast<int64_t>(jj);
          st.load(a, j);
          t = a.t0;
          t1 = a.t1_per ? a.t1_per[j] : a.t1;
          dt = (t1 - t) / 100.0;
          acc = 0;
          rej = 0;
          it = 0;
          active = true;
          if (tcount >= pa.traj_cap || rcursor >= pa.rec_cap) {
            finish(kTapeFull);
          } else {
            pa.order[static_cast<int64_t>(tcount) * L + lane] = static_cast<int32_t>(j);
            ++tcount;
            pa.lane_of[j] = static_cast<int32_t>(lane);
            pa.rec_start[j] = rcursor;
            if (t < t1) {
              open_record();  // room checked above
            } else {
              finish(kReached);
            }
          }
        } else {
          done = true;
        }
      }
      if (cnt <= avail) {
        cbase += cnt;
      } else if (ne > nb) {
        const unsigned long long used = cnt - avail;
        cbase = nb + (used < ne - nb ? used : ne - nb);
        cend = ne;
      } else {
        cbase = cend;
      }
    }
    if (__ballot(!done || active) == 0ull) break;
    if (!active) continue;

    // One trial step (gradsolve/solvers/tsit5_replay.py:107-124).
    const double h = nn_min(dt, t1 - t);
    const double err = st.trial(h, a.rtol, a.atol, a.aux);
    ++it;
    if (err != err) {
      rej += static_cast<int32_t>(a.max_steps - (it - 1));
      finish(kExhausted);
      continue;
    }
    if (err <= 1.0) {
      t = t + h;
      st.accept();
      *hp = h;
      hp += kStage ? kPBlock : 1;
      ++acc;
      if (--seg_left == 0) {
        close_record();
        if (t < t1 && !open_record()) {
          finish(kTapeFull);
          continue;
        }
      }
    } else {
      ++rej;
    }
    dt = tsit5_next_dt(h, err);
    if (!(t < t1)) {
      finish(kReached);
    } else if (it >= a.max_steps) {
      finish(kExhausted);
    }
  }
  pa.lane_count[lane] = tcount;
}

// Waves per SIMD the adjoint asks for: four, except where its segment states in LDS
// (CMAX x 3 doubles x 64 lanes per block) leave room for fewer.
constexpr int persist_adj_waves(int cmax) { return cmax <= 8 ? 4 : cmax == 16 ? 2 : 1; }

// One lane walks its trajectories backwards, one segment record per iteration.
template <class F, int CMAX, bool FP32>
__global__ __launch_bounds__(kPAdjBlock)
__attribute__((amdgpu_waves_per_eu(persist_adj_waves(CMAX))))
void persist_adj_kernel(PersistAdjArgs pa) {
  static_assert(F::autonomous, "the persistent engine is built for autonomous fields");
  constexpr int D = F::D, P = F::P, DCK = ckpt_rows<F>();
  // segment states for re-integration (the pair path at C = 2 keeps none)
  constexpr int kSegRows = (CMAX == 2 && !FP32) ? 1 : CMAX;
  __shared__ double seg_all[kSegRows * DCK * kPAdjBlock];
  const AdjArgs& a = pa.a;
  const int64_t L = pa.lanes;
  const int R = pa.rec_len;
  const int64_t lane = static_cast<int64_t>(blockIdx.x) * kPAdjBlock + threadIdx.x;
  if (lane >= L) return;
  double* seg = seg_all + threadIdx.x;
  const int64_t n = a.n;
  const int32_t C = a.ckpt_interval;
  int32_t k = pa.lane_count[lane] - 1;
  bool have = false;
  int64_t j = 0;
  int32_t na = 0, m = 0, rstart = 0;
  double lam_d[D], pb_d[P], p_d[P];
  float lam_f[D], lam_lo[D], pb_f[P], pb_lo[P], p_f[P];
  while (true) {
    if (!have) {
      if (k < 0) break;
      j = pa.order[static_cast<int64_t>(k) * L + lane];
      --k;
      na = a.n_acc[j];
      rstart = pa.rec_start[j];
      m = (na + C - 1) / C - 1;
      GSX_UNROLL for (int d = 0; d < D; ++d) {
        const double v = a.ybar[d * n + j];
        lam_d[d] = v;
        lam_f[d] = static_cast<float>(v);
        lam_lo[d] = static_cast<float>(v - static_cast<double>(lam_f[d]));
      }
      GSX_UNROLL for (int q = 0; q < P; ++q) {
        p_d[q] = a.p[q * n + j];
        p_f[q] = static_cast<float>(p_d[q]);
        pb_d[q] = 0.0;
        pb_f[q] = 0.0f;
        pb_lo[q] = 0.0f;
      }
      have = true;
    }
    if (m >= 0) {
      const int32_t s0 = m * C;
      const int32_t s1 = (s0 + C < na) ? s0 + C : na;
      const double* rp = pa.rec + persist_record(rstart + m, lane, L, R);
      const double* hp = rp + DCK;  // the segment's step sizes
      if constexpr (FP32) {
        float y[D], lo[D];
        GSX_UNROLL for (int d = 0; d < D; ++d) {
          const float* cell = reinterpret_cast<const float*>(rp + d);
          y[d] = cell[0];
          lo[d] = cell[1];
        }
        for (int32_t s = s0; s < s1; ++s) {
          double* slot = seg + static_cast<int64_t>(s - s0) * D * kPAdjBlock;
          GSX_UNROLL for (int d = 0; d < D; ++d) {
            float* cell = reinterpret_cast<float*>(slot + d * kPAdjBlock);
            cell[0] = y[d];
            cell[1] = lo[d];
          }
          if (s + 1 < s1) tsit5_advance_f32<F>(y, lo, static_cast<float>(hp[s - s0]), p_f, a.aux);
        }
        for (int32_t s = s1 - 1; s >= s0; --s) {
          const double* slot = seg + static_cast<int64_t>(s - s0) * D * kPAdjBlock;
          GSX_UNROLL for (int d = 0; d < D; ++d)
            y[d] = reinterpret_cast<const float*>(slot + d * kPAdjBlock)[0];
          tsit5_step_vjp_f32<F>(y, static_cast<float>(hp[s - s0]), p_f, lam_f, lam_lo, pb_f,
                                pb_lo, a.aux);
        }
      } else if constexpr (CMAX == 2) {
        // Checkpoint interval 2: one advance from the checkpoint forms step s0's stage inputs
        // and y_{s0+1}; step s0+1's own stage inputs serve its VJP, step s0's reuse the kept
        // ones, so every step's stages are formed once and nothing goes through LDS (the
        // pair adjoint of tsit5_adj_pair_segment; a missing second step runs with h = 0,
        // which leaves lam and pb unchanged).
        double y[D];
        GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = rp[d];
        const double h0 = hp[0];
        const double h1 = (s0 + 1 < na) ? hp[1] : 0.0;
        double z0[6][D], y1[D];
        tsit5_stage_inputs<F>(0.0, y, h0, p_d, z0, y1, a.aux);
        {
          double z1[6][D];
          tsit5_stage_inputs<F>(0.0, y1, h1, p_d, z1, nullptr, a.aux);
          tsit5_step_vjp_z<F>(0.0, z1, h1, p_d, lam_d, pb_d, a.aux);
        }
        tsit5_step_vjp_z<F>(0.0, z0, h0, p_d, lam_d, pb_d, a.aux);
        (void)s1;
        (void)seg;
      } else {
        double y[D];
        GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = rp[d];
        for (int32_t s = s0; s < s1; ++s) {
          double* slot = seg + static_cast<int64_t>(s - s0) * D * kPAdjBlock;
          GSX_UNROLL for (int d = 0; d < D; ++d) slot[d * kPAdjBlock] = y[d];
          if (s + 1 < s1) tsit5_advance<F>(0.0, y, hp[s - s0], p_d, a.aux);
        }
        for (int32_t s = s1 - 1; s >= s0; --s) {
          const double* slot = seg + static_cast<int64_t>(s - s0) * D * kPAdjBlock;
          GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = slot[d * kPAdjBlock];
          tsit5_step_vjp<F>(0.0, y, hp[s - s0], p_d, lam_d, pb_d, a.aux);
        }
      }
      --m;
    }
    if (m < 0) {
      GSX_UNROLL for (int d = 0; d < D; ++d)
        a.y0bar[d * n + j] = FP32 ? static_cast<double>(lam_f[d]) + static_cast<double>(lam_lo[d])
                                  : lam_d[d];
      GSX_UNROLL for (int q = 0; q < P; ++q)
        a.pbar[q * n + j] = FP32 ? static_cast<double>(pb_f[q]) + static_cast<double>(pb_lo[q])
                                 : pb_d[q];
      have = false;
    }
  }
}

template <class F, class State>
int launch_pfwd(const PersistFwdArgs& pa, hipStream_t s) {
  queue_reset_kernel<<<1, 1, 0, s>>>(pa.queue);
  int rc = check_launch("queue_reset_kernel");
  if (rc) return rc;
  const unsigned grid = static_cast<unsigned>(pa.lanes / kPBlock);
  if (pa.rec_len <= ::gsx::kStageMaxLen) {
    const size_t lds = static_cast<size_t>(pa.rec_len) * kPBlock * sizeof(double);
    persist_fwd_kernel<F, State, true><<<grid, kPBlock, lds, s>>>(pa);
  } else {
    persist_fwd_kernel<F, State, false><<<grid, kPBlock, 0, s>>>(pa);
  }
  return check_launch("persist_fwd_kernel");
}

template <class F, int CMAX, bool FP32>
int launch_padj_c(const PersistAdjArgs& pa, hipStream_t s) {
  const unsigned grid = static_cast<unsigned>((pa.lanes + kPAdjBlock - 1) / kPAdjBlock);
  persist_adj_kernel<F, CMAX, FP32><<<grid, kPAdjBlock, 0, s>>>(pa);
  return check_launch("persist_adj_kernel");
}

template <class F, bool FP32>
int launch_padj(const PersistAdjArgs& pa, hipStream_t s) {
  const int C = pa.a.ckpt_interval;
  if (C <= 1) return launch_padj_c<F, 1, FP32>(pa, s);
  if (C == 2 && !FP32) return launch_padj_c<F, 2, FP32>(pa, s);
  if (C <= 4) return launch_padj_c<F, 4, FP32>(pa, s);
  if (C <= 8) return launch_padj_c<F, 8, FP32>(pa, s);
  if (C <= 16) return launch_padj_c<F, 16, FP32>(pa, s);
  if (C <= 32) return launch_padj_c<F, 32, FP32>(pa, s);
  return set_error(GSX_E_INVALID, "ckpt_interval %d exceeds the adjoint's maximum of %d", C,
                   kMaxCkptInterval);
}

}  // namespace

int persist_fwd_dispatch(int field, int D, bool fp32, const PersistFwdArgs& pa,
                         hipStream_t s) {
  switch (field) {
    case GSX_FIELD_LORENZ:
      return fp32 ? launch_pfwd<FLorenz, StateF32<FLorenz>>(pa, s)
                  : launch_pfwd<FLorenz, StateF64<FLorenz>>(pa, s);
    case GSX_FIELD_LORENZ3:
      return fp32 ? launch_pfwd<FLorenz3, StateF32<FLorenz3>>(pa, s)
                  : launch_pfwd<FLorenz3, StateF64<FLorenz3>>(pa, s);
    default: break;
  }
  return set_error(GSX_E_UNSUPPORTED, "persistent tsit5 forward: field %d with D=%d is not "
                   "built", field, D);
}

int persist_adj_dispatch(int field, int D, bool fp32, const PersistAdjArgs& pa,
                         hipStream_t s) {
  switch (field) {
    case GSX_FIELD_LORENZ:
      return fp32 ? launch_padj<FLorenz, true>(pa, s) : launch_padj<FLorenz, false>(pa, s);
    case GSX_FIELD_LORENZ3:
      return fp32 ? launch_padj<FLorenz3, true>(pa, s) : launch_padj<FLorenz3, false>(pa, s);
    default: break;
  }
  return set_error(GSX_E_UNSUPPORTED, "persistent tsit5 adjoint: field %d with D=%d is not "
                   "built", field, D);
}

}  // namespace GSX_NS
