// gsx_args.h -- argument blocks and status codes shared by every kernel variant.
#pragma once

#include <cstdint>

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
};

}  // namespace gsx
