// gsx_args.h -- argument blocks and status codes shared by every kernel variant.
#pragma once

#include "gsx_platform.h"

namespace gsx {

// Per-trajectory status codes (gradsolve/solvers/record_jax.py:71).
enum : int32_t { kReached = 0, kExhausted = 1, kDegenerate = 2, kTapeFull = 3 };

struct FwdArgs {
  int64_t n;
  const double* y0;
  const double* p;
  const double* aux;
  double t0, t1;
  const double* t1_per;  // [n] per-trajectory end time or nullptr
  double rtol, atol;
  int64_t max_steps;
  double* y_final;
  int32_t* n_acc;
  int32_t* n_rej;
  int32_t* status;
  double* tape;  // [cap][n] or nullptr
  int64_t tape_cap;
  double* ckpt;  // [slots][DCK][n] or nullptr
  int32_t ckpt_interval;
};

struct ReplayArgs {
  int64_t n;
  const double* y0;
  const double* p;
  const double* aux;
  double t0;
  const double* tape;
  const int32_t* n_acc;
  double* y_final;
  double* ckpt;  // [slots][DCK][n] or nullptr
  int32_t ckpt_interval;
};

struct AdjArgs {
  int64_t n;
  const double* y0;
  const double* p;
  const double* aux;
  double t0;
  const double* tape;
  int64_t tape_cap;
  const int32_t* n_acc;
  const double* ckpt;
  int32_t ckpt_interval;
  const double* ybar;
  double* y0bar;
  double* pbar;
  // Save-time cotangents (saveat gradients; tsit5_saveat_vjp_lane prepares them). With
  // nsave > 0, mu[q] (a cotangent on the start state of the step bracketing save q) is added
  // to lam when the reverse sweep reaches that step, save_step[q] == -1 adds it at y0, and
  // pbar starts from pbar_in. ybar then already holds the final-state saves.
  int64_t nsave = 0;
  const int32_t* save_step = nullptr;  // [nsave][n]
  const double* mu = nullptr;          // [nsave][D][n]
  const double* pbar_in = nullptr;     // [P][n] or nullptr (start from zero)
  // Dense-output cotangents (the Rodas5P continuous extension): with ysbar set, save q
  // bracketed by step s = save_step[q] contributes at that step through its interpolant at
  // theta = (ts[q] - save_t[q]) / h_s; ysbar [nsave][D][n].
  const double* ts = nullptr;
  const double* save_t = nullptr;
  const double* ysbar = nullptr;
};

// The wide Tsit5 lanes (gsx_wide.h): the register lanes' argument blocks plus a workspace.
// Workspace rows, D doubles each per trajectory, [row][D][n]:
//   forward / replay: y, y5 (exchanged on accept), z (stage input), k1..k7 (k1 and k7
//   exchanged on accept: FSAL without a copy)
//   adjoint: for each of the segment's C steps its start state and stage inputs z2..z6
//   (6 rows), then k1..k6, kbar1..kbar6, lam, zbar
constexpr int kWideFwdRows = 10;
GSX_HD constexpr int wide_adj_rows(int C) { return 6 * C + 14; }

struct WideFwdArgs {
  FwdArgs f;
  double* ws;  // [kWideFwdRows][D][n]
};
struct WideReplayArgs {
  ReplayArgs r;
  double* ws;  // [kWideFwdRows][D][n]
};
struct WideAdjArgs {
  AdjArgs a;
  double* ws;  // [wide_adj_rows(C)][D][n]
};

// Dense output on a recorded mesh (gradsolve/solvers/dense.py:106-163): for every save time
// ts[q] (ascending, shared by the ensemble), the accepted step s that brackets it
// (t_s < ts[q] <= t_s + h_s), its start time and state, and the saved state, one Tsit5 step of
// size ts[q] - t_s from there. Saves at or before t0 are y0 (save_step -1); saves at or after
// the final time are y_final (save_step n_acc).
struct SaveArgs {
  int64_t n;
  const double* y0;
  const double* p;
  const double* aux;
  double t0;
  const double* tape;     // [cap][n]
  const int32_t* n_acc;
  const double* ts;       // [nsave]
  int64_t nsave;
  double* ys;             // [nsave][D][n]
  int32_t* save_step;     // [nsave][n]
  double* save_t;         // [nsave][n]
  double* states;         // [nsave][D][n] bracket start states, or nullptr
  // [D][n] final states to copy into the final-time saves (the forward's own y_final, so that
  // those saves equal it bit for bit), or nullptr to take the replay's
  const double* y_final = nullptr;
};

// The re-steps' VJPs for save cotangents ysbar: mu (see AdjArgs), the final-state cotangent
// with the final-time saves folded in (ybar_eff), and the re-steps' parameter gradient.
struct SaveVjpArgs {
  int64_t n;
  const double* p;
  const double* aux;
  const int32_t* n_acc;
  const double* ts;
  int64_t nsave;
  const int32_t* save_step;
  const double* save_t;
  const double* states;
  const double* ysbar;    // [nsave][D][n]
  const double* ybar;     // [D][n] final-state cotangent or nullptr
  double* mu;             // [nsave][D][n]
  double* ybar_eff;       // [D][n]
  double* pbar;           // [P][n]
};

// Persistent work-queue forms (gsx_persist.h describes the layout).
struct PersistFwdArgs {
  FwdArgs f;           // n, y0, p, aux, t0, t1, t1_per, rtol, atol, max_steps, y_final,
                       // n_acc, n_rej, status, ckpt_interval (C); f.tape / f.ckpt unused
  int64_t lanes;       // L = grid * block
  double* rec;         // [rec_cap][L][rec_len] segment records
  int64_t rec_cap;     // records per lane
  int32_t rec_len;     // doubles per record
  int64_t traj_cap;    // trajectories per lane
  int32_t* order;      // [traj_cap][L]
  int32_t* lane_count; // [L] trajectories processed by each lane
  int32_t* lane_of;    // [n]
  int32_t* rec_start;  // [n] record of the trajectory's first segment, in its lane's stream
  unsigned long long* queue;  // one counter, zero at launch
  const int32_t* queue_order = nullptr;  // [n] order the queue hands trajectories out, or null
};

struct PersistAdjArgs {
  AdjArgs a;           // n, p, aux, n_acc, ckpt_interval, ybar, y0bar, pbar
  int64_t lanes;
  const double* rec;   // the forward's segment records
  int32_t rec_len;
  const int32_t* order;
  const int32_t* lane_count;
  const int32_t* rec_start;
};


// Kuramoto ensemble (gsx_kuramoto.hip describes the layouts). D = P = 32.
struct KuraFwdArgs {
  int64_t n;
  const double* K;       // [32][32] row-major, shared
  const double* y0;      // [32][n]
  const double* omega;   // [32][n]
  double t0, t1, rtol, atol;
  int64_t max_steps;
  const double* ts;      // [nsave] sorted save times, or nullptr
  int64_t nsave;
  double* y_final;       // [32][n]
  int32_t* n_acc;
  int32_t* n_rej;
  int32_t* status;
  double* tape;          // [tape_cap][n] or nullptr (forward only)
  int64_t tape_cap;
  double* states;        // [tape_cap][n][32] or nullptr
  int32_t* save_step;    // [nsave][n] or nullptr
  double* save_t;        // [nsave][n]
};

struct KuraReplayArgs {
  int64_t n;
  const double* K;
  const double* y0;
  const double* omega;
  const double* tape;
  const int32_t* n_acc;
  double* y_final;
  double* states;
};

struct KuraSaveatArgs {
  int64_t n;
  const double* K;
  const double* y0;
  const double* omega;
  const double* y_final;
  const double* states;
  const double* ts;
  int64_t nsave;
  const int32_t* save_step;
  const double* save_t;
  double* ys;            // [n][nsave][32]
};

struct KuraAdjArgs {
  int64_t n;
  const double* K;
  const double* omega;
  const double* tape;
  const int32_t* n_acc;
  const double* states;
  const double* ts;
  int64_t nsave;
  const int32_t* save_step;
  const double* save_t;
  const double* ysbar;   // [n][nsave][32] or nullptr
  const double* yfbar;   // [32][n] or nullptr
  double* mu;            // [n][nsave][32] scratch (when ysbar is given)
  double* y0bar;         // [32][n]
  double* omega_bar;     // [32][n]
};


// The naive reverse sweep the benchmark compares against (gsx_plain.hip).
struct PlainArgs {
  int64_t n;
  const double* aux;
  const double* y0;
  const double* p;
  double t0, t1, rtol, atol;
  int64_t max_steps;
  double* y_final;
  int32_t* n_acc;
  int32_t* n_rej;
  int32_t* status;
  double* tape;          // [cap][n]
  double* stages;        // [cap][6][D][n]
  int64_t cap;
  const double* ybar;    // adjoint: [D][n]
  double* y0bar;
  double* pbar;
};

}  // namespace gsx
