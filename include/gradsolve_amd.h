/*
 * gradsolve_amd.h -- C ABI of the MI355X ensemble ODE record-and-replay library
 * (libgradsolve_amd.so, gfx950).
 *
 * Every entry point takes plain device pointers, sizes and a HIP stream handle passed as
 * `void*` (a hipStream_t; NULL = the legacy default stream). Calls are stream-ordered and
 * asynchronous: they enqueue kernels and return; the caller synchronizes. No entry point
 * allocates device memory: the caller owns every buffer, sized with the query functions.
 *
 * Device layout (structure of arrays; j = trajectory index, n = ensemble size):
 *   y0, y_final, ybar, y0bar : double[D][n]
 *   p, pbar                  : double[P][n]
 *   t1_per                   : double[n]                 (optional per-trajectory end time)
 *   n_acc, n_rej, status     : int32[n]
 *   tape                     : double[tape_cap][n]       accepted step s of trajectory j
 *                                                         at tape[s*n + j]
 *   ckpt                     : double[gsx_ckpt_slots(..)][gsx_ckpt_rows(field)][n]
 *
 * Status codes per trajectory (status[j]): 0 reached t1, 1 exhausted max_steps,
 * 2 step-size underflow / degenerate step, 3 tape capacity exceeded.
 *
 * Return codes: GSX_OK (0) or a negative GSX_E* value; gsx_last_error() gives the text.
 */
#ifndef GRADSOLVE_AMD_H
#define GRADSOLVE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum gsx_status {
  GSX_OK = 0,
  GSX_E_INVALID = -1,     /* bad argument: unknown field, dimension mismatch, null pointer */
  GSX_E_HIP = -2,         /* a HIP runtime call or kernel launch failed */
  GSX_E_UNSUPPORTED = -3  /* valid request this build does not implement */
};

/* Built-in fields (the registered fields of gradsolve's fused kernels). */
enum gsx_field {
  GSX_FIELD_LORENZ = 1,       /* D=3, P=1 (rho); sigma=10, beta=8/3 */
  GSX_FIELD_LINEAR_DECAY = 2, /* D in {1,2,3}, P=1: f = -p0*y */
  GSX_FIELD_ROBERTSON = 3,    /* D=3, P=3 (k1, k2, k3); stiff */
  GSX_FIELD_KURAMOTO = 4,     /* D=32, P=32 (omega); aux = K[32*32] row-major */
  GSX_FIELD_VDP = 5,          /* D=2, P=1 (mu) */
  GSX_FIELD_LORENZ3 = 6,      /* D=3, P=3 (sigma, beta, rho) */
  GSX_FIELD_FORCED_STIFF = 8  /* D=2, P=3 (a, b, c); stiff, reads t:
                                 f0 = -a (y0 - y1^2) + 0.4 sin 3t,
                                 f1 = y0 - y1 - b y1^3 + 0.2 cos t + c sin(t) y0 */
};

/* Fields registered at run time (gsx_field_register) take ids from GSX_FIELD_RUNTIME_BASE. */
enum { GSX_FIELD_RUNTIME_BASE = 1000 };

/* gsx_field_register flags: a stiff field (with its analytic Jacobian; served by the
 * Rosenbrock23 and Rodas5P entry points) and a field whose right-hand side reads t
 * (checkpoints then carry t in an extra row, see gsx_ckpt_rows; a stiff one also supplies
 * df/dt for Rodas5P). */
enum gsx_field_flags { GSX_FIELD_STIFF = 1, GSX_FIELD_TIME_DEPENDENT = 2 };

/* Arithmetic flags. Default (0): the kernels contract a*b + c into fused multiply-adds.
 * GSX_STRICT: contraction off, so every operation rounds once in the reference loop's
 * textual order (gradsolve/solvers/tsit5_replay.py:100-124); the recorded tape then matches
 * an IEEE-sequential host loop up to the last-ulp differences of the device pow.
 * GSX_FP32 (Tsit5 and Rosenbrock23, autonomous fields): stages, the field and its adjoint
 * in FP32 with a Kahan-compensated state and compensated adjoint accumulators; time, step
 * size, error norm, controller and tape stay FP64, and on Rosenbrock23 so do W = I - d h J, its
 * LU and the triangular solves; checkpoints hold the compensated pair in the bytes of an FP64
 * checkpoint. The two flags are exclusive. */
enum gsx_flags { GSX_STRICT = 1, GSX_FP32 = 2 };

/* Registers a field given as HIP C++ source and returns its id in *field_id. The id is
 * accepted wherever a built-in field id is: gsx_tsit5_fwd / _replay / _adjoint in every
 * arithmetic variant (GSX_FP32 for autonomous fields), gsx_rb23_* for stiff fields, and the
 * layout queries. The source defines `struct Field` (the library places it in a namespace of
 * its own, where the gsx_core.h helpers are visible) with the members of the built-in fields:
 *   static constexpr int D, P;  static constexpr bool autonomous;
 *   template <class T, class YP, class DP> GSX_HD static void f(T t, YP y, const T* p,
 *                                                             DP dy, const double* aux);
 *   template <class T, class YP, class VP, class YBP> GSX_HD static void vjp(
 *       T t, YP y, const T* p, VP v, YBP yb, T* pb, const double* aux);
 * (vjp adds v^T df/dy to yb and v^T df/dp to pb; y, dy, v and yb are indexed with [i]: arrays
 * in the register lanes, strided workspace views in the wide lane), and for stiff fields
 *   template <class T> GSX_HD static void jac(T t, const T* y, const T* p, T* J,
 *                                             const double* aux);       (J row-major D x D)
 *   template <class T> GSX_HD static void jac_vjp(const T* y, const T* p, const T* Jb,
 *                                                 T* yb, T* pb, const double* aux);
 * D, P and autonomy must agree with the arguments and flags. D is at most 4096; a field wider
 * than 64 runs on the wide Tsit5 lane only (gsx_tsit5_wide_*), and a stiff field is at most
 * 12 wide (each lane of the stiff lanes factorises a D x D W in registers; the reference
 * refuses its fused stiff kernel above the same 12, gradsolve/warp/warp_rosenbrock.py:67,
 * 324-340). The kernels are compiled for gfx950 with hiprtc, one per program: the recording
 * forward (Tsit5, the wide Tsit5 lane above D = 64, or Rosenbrock23 for a stiff field) now, so that a compile error returns GSX_E_INVALID with the compiler's log in
 * gsx_last_error(); every other kernel (replay, adjoints, strict and FP32 variants) on its
 * first launch. A segment adjoint holds C checkpoint rows per lane in LDS, so for wide fields
 * checkpoint intervals whose kernel would need more than 64 KiB per wavefront are rejected
 * (GSX_E_INVALID); C = 2 (the pair adjoint) needs none. Replaces gradsolve's register_jax_field and the
 * fused-kernel builds it triggers (gradsolve/warp/jax_field.py:516-564 registering into
 * gradsolve/warp/_warp_kernel.py and gradsolve/warp/warp_rosenbrock.py). Thread-safe. */
int gsx_field_register(const char* source, int D, int P, int flags, int* field_id);

/* Unloads a registered field's kernels from every device they were loaded on; the id is not
 * reused and later calls with it fail with GSX_E_INVALID. */
int gsx_field_release(int field_id);

/* Library version (major*10000 + minor*100 + patch). */
int gsx_version(void);

/* Text of the last error on the calling thread ("" if none). */
const char* gsx_last_error(void);

/* The Tsit5 tableau compiled into the kernels: c2..c6 (5), a21..a65 (15, row order),
 * b1..b6 (6), e1..e7 (7). Host memory. No reference counterpart; used to check the
 * kernels' constants against gradsolve/solvers/tsit5_step.py:31-77. */
int gsx_tsit5_tableau(double* c, double* a, double* b, double* e);

/* Rows per checkpoint slot for `field` (D, plus one for t when the field is
 * non-autonomous), and the number of checkpoint slots needed for a tape of `tape_cap`
 * accepted steps checkpointed every `ckpt_interval` accepted steps. */
int gsx_ckpt_rows(int field, int D);
int64_t gsx_ckpt_slots(int64_t tape_cap, int32_t ckpt_interval);

/* Fused adaptive Tsit5 forward + accepted-h record, one trajectory per lane.
 * Replaces the XLA FFI handler Tsit5Fwd (gradsolve/cuda/kernels/tsit5_template.cu.in:127-178:
 * y0[D,n], rho[n] -> yf[D,n], nsteps[n]; attrs t1, rtol, atol, max_steps) and the Warp
 * forward+record kernel _tsit5_adaptive_general
 * (gradsolve/warp/_warp_kernel.py:325-399, driven by warp_ode._launch, warp_ode.py:215-327).
 * tape may be NULL (forward only); ckpt may be NULL (no adjoint will follow).
 * t1_per may be NULL (every trajectory ends at t1). dt0 = (t1 - t0)/100 per trajectory.
 * Without a tape, n_acc, n_rej and status may each be NULL, so Tsit5Fwd binds 1:1 as
 * (y0, p) -> (y_final, n_acc = nsteps); with a tape, n_acc and status are required (the
 * tape's lengths and its "tape full" status). */
int gsx_tsit5_fwd(int field, int64_t n, int D, int P, const double* aux,
                  const double* y0, const double* p, double t0, double t1,
                  const double* t1_per, double rtol, double atol, int64_t max_steps,
                  double* y_final, int32_t* n_acc, int32_t* n_rej, int32_t* status,
                  double* tape, int64_t tape_cap, double* ckpt, int32_t ckpt_interval,
                  int flags, void* stream);

/* Frozen-mesh replay from (possibly new) y0, p over a recorded tape: the forward of
 * replay_solve_jax (gradsolve/warp/warp_replay.py:30-55, one tsit5_step per recorded h,
 * gradsolve/solvers/tsit5_step.py:80-102), writing the checkpoints gsx_tsit5_adjoint reads.
 * The tape holds the first n_acc[j] accepted steps of trajectory j at tape[s*n + j]. */
int gsx_tsit5_replay(int field, int64_t n, int D, int P, const double* aux,
                     const double* y0, const double* p, double t0,
                     const double* tape, const int32_t* n_acc,
                     double* y_final, double* ckpt, int32_t ckpt_interval, int flags,
                     void* stream);

/* Reverse sweep: the discrete adjoint of the frozen-h Tsit5 replay, walking the tape
 * backwards from the checkpoints written by gsx_tsit5_fwd (same tape_cap and
 * ckpt_interval). ybar = dL/dy_final; outputs dL/dy0 and dL/dp.
 * Replaces jax.grad through replay_solve_jax (gradsolve/warp/warp_replay.py:30-55). */
int gsx_tsit5_adjoint(int field, int64_t n, int D, int P, const double* aux,
                      const double* y0, const double* p, double t0,
                      const double* tape, int64_t tape_cap, const int32_t* n_acc,
                      const double* ckpt, int32_t ckpt_interval,
                      const double* ybar, double* y0bar, double* pbar, int flags,
                      void* stream);

/* Dense output on a recorded mesh (saveat) for any Tsit5 field: replays the tape and, for each
 * save time ts[q] (ascending, shared by the ensemble), writes ys[q][d][j]: y0 for ts[q] <= t0,
 * y_final for ts[q] at or after the final time, and otherwise one Tsit5 step of size
 * ts[q] - t_s from the start (t_s, y_s) of the accepted step with t_s < ts[q] <= t_s + h_s.
 * The bracket is returned for the gradient: save_step[q][j] (-1 at y0, n_acc at y_final),
 * save_t[q][j] = t_s, and (states non-NULL) states[q][d][j] = y_s. y_final (optional) is
 * copied into the final-time saves, so that they equal the forward's final state bit for bit
 * (NULL: the replay's). flags: 0 or GSX_STRICT. Replaces replay_solve_saveat_jax
 * (gradsolve/warp/warp_replay.py:58-75 -> gradsolve/solvers/dense.py:106-163) for the Tsit5
 * lanes. */
int gsx_tsit5_saveat(int field, int64_t n, int D, int P, const double* aux,
                     const double* y0, const double* p, double t0,
                     const double* tape, const int32_t* n_acc, const double* y_final,
                     const double* ts, int64_t nsave, double* ys, int32_t* save_step,
                     double* save_t, double* states, int flags, void* stream);

/* Gradient of a saveat series (and optionally of the final state): the VJPs of the re-steps
 * (mu, a caller scratch buffer [nsave][D][n]) followed by the adjoint sweep of
 * gsx_tsit5_adjoint with each save's cotangent added at its bracketing step. ysbar [nsave][D][n]
 * is dL/dys, ybar (NULL: zero) dL/dy_final; save_step / save_t / states as gsx_tsit5_saveat
 * wrote them; tape, ckpt, tape_cap and ckpt_interval as gsx_tsit5_fwd wrote them. Outputs
 * dL/dy0 and dL/dp. flags: 0 or GSX_STRICT. Replaces jax.grad through replay_solve_saveat_jax
 * (gradsolve/solvers/dense.py:318-335). */
int gsx_tsit5_adjoint_saveat(int field, int64_t n, int D, int P, const double* aux,
                             const double* y0, const double* p, double t0,
                             const double* tape, int64_t tape_cap, const int32_t* n_acc,
                             const double* ckpt, int32_t ckpt_interval,
                             const double* ts, int64_t nsave, const int32_t* save_step,
                             const double* save_t, const double* states, const double* ysbar,
                             const double* ybar, double* mu, double* y0bar, double* pbar,
                             int flags, void* stream);

/* ---- Wide-state Tsit5 lane (fields registered at run time, D up to 4096) ------------------
 * The reference routes non-stiff problems above NVAR_CEILING = 64 to its general
 * record-and-replay engine (gradsolve/dispatch.py:256-258: record_tsit5_jax with the
 * vmapped while-loop recorder, gradsolve/solvers/record_jax.py:171-213, and jax.grad of the
 * replay scan, gradsolve/solvers/tsit5_replay.py:127-232). These entry points replace it for
 * fields compiled with gsx_field_register: one trajectory per lane, the state, the stage
 * derivatives and the stage inputs in a caller-owned workspace ws [rows][D][n] f64 (rows from
 * gsx_tsit5_wide_ws_rows: which 0 for the forward and the replay, which 1 with the checkpoint
 * interval for the adjoint). Arguments, layouts, statuses, tape and checkpoints otherwise as
 * gsx_tsit5_fwd / gsx_tsit5_replay / gsx_tsit5_adjoint (default arithmetic only). A field
 * wider than 64 runs on this lane only; built-in fields are rejected (GSX_E_UNSUPPORTED). */
int64_t gsx_tsit5_wide_ws_rows(int which, int32_t ckpt_interval);
int gsx_tsit5_wide_fwd(int field, int64_t n, int D, int P, const double* aux,
                       const double* y0, const double* p, double t0, double t1,
                       const double* t1_per, double rtol, double atol, int64_t max_steps,
                       double* y_final, int32_t* n_acc, int32_t* n_rej, int32_t* status,
                       double* tape, int64_t tape_cap, double* ckpt, int32_t ckpt_interval,
                       double* ws, void* stream);
int gsx_tsit5_wide_replay(int field, int64_t n, int D, int P, const double* aux,
                          const double* y0, const double* p, double t0, const double* tape,
                          const int32_t* n_acc, double* y_final, double* ckpt,
                          int32_t ckpt_interval, double* ws, void* stream);
int gsx_tsit5_wide_adjoint(int field, int64_t n, int D, int P, const double* aux,
                           const double* y0, const double* p, double t0, const double* tape,
                           int64_t tape_cap, const int32_t* n_acc, const double* ckpt,
                           int32_t ckpt_interval, const double* ybar, double* y0bar,
                           double* pbar, double* ws, void* stream);

/* gsx_tsit5_adjoint_saveat without the caller's mu buffer: the re-steps' cotangents go to a
 * library-owned device array of 2^23 doubles (64 MiB) per device, so the call allocates
 * nothing. No counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). nsave * D * n above 2^23 is rejected (GSX_E_INVALID), as are fields compiled at
 * run time (their kernels load on first use). Every call on a device shares that device's
 * array, so calls that overlap on one device (other streams, other threads) race; the array
 * needs no clearing (each call writes every entry it reads). Arguments and results as
 * gsx_tsit5_adjoint_saveat. */
int gsx_tsit5_adjoint_saveat_noalloc(int field, int64_t n, int D, int P, const double* aux,
                                     const double* y0, const double* p, double t0,
                                     const double* tape, int64_t tape_cap,
                                     const int32_t* n_acc, const double* ckpt,
                                     int32_t ckpt_interval, const double* ts, int64_t nsave,
                                     const int32_t* save_step, const double* save_t,
                                     const double* states, const double* ysbar,
                                     const double* ybar, double* y0bar, double* pbar,
                                     int flags, void* stream);

/* Fused adaptive Rosenbrock23 (ode23s, autonomous specialisation) forward + accepted-h
 * record for stiff fields with an analytic Jacobian: W = I - d*h*J(t, y, p) factorised once
 * per step (LU with partial pivoting in registers) and reused by the three stage solves;
 * I-controller with exponent -1/3. As in the reference, every stage evaluates the field at the
 * step's start time t and no df/dt term is added (a field that reads t gets the reference's
 * autonomous formula; Rodas5P carries the df/dt term). Same argument contract as
 * gsx_tsit5_fwd (optional counts and status without a tape, so Rosenbrock23Fwd binds as
 * (y0, p) -> (y_final, nsteps)). Replaces the XLA FFI handler
 * Rosenbrock23Fwd (gradsolve/cuda/kernels/rosenbrock23_template.cu.in:151-223) and the Warp
 * kernel _rosenbrock23_adaptive (gradsolve/warp/_warp_rosenbrock.py:383-452, driven by
 * warp_rosenbrock._launch, gradsolve/warp/warp_rosenbrock.py:166-218). */
int gsx_rb23_fwd(int field, int64_t n, int D, int P, const double* aux,
                 const double* y0, const double* p, double t0, double t1,
                 const double* t1_per, double rtol, double atol, int64_t max_steps,
                 double* y_final, int32_t* n_acc, int32_t* n_rej, int32_t* status,
                 double* tape, int64_t tape_cap, double* ckpt, int32_t ckpt_interval,
                 int flags, void* stream);

/* Frozen-mesh Rosenbrock23 replay (rosenbrock23_step without the error stage): the forward
 * of replay_solve_rosenbrock_jax (gradsolve/warp/warp_replay.py:267-293;
 * gradsolve/solvers/rosenbrock23_step.py:47-66). */
int gsx_rb23_replay(int field, int64_t n, int D, int P, const double* aux,
                    const double* y0, const double* p, double t0,
                    const double* tape, const int32_t* n_acc,
                    double* y_final, double* ckpt, int32_t ckpt_interval, int flags,
                    void* stream);

/* Discrete adjoint of the Rosenbrock23 replay, differentiating through the stage solves and
 * W = I - d*h*J(y, p). Replaces jax.grad through replay_solve_rosenbrock_jax
 * (gradsolve/warp/warp_replay.py:267-293). */
int gsx_rb23_adjoint(int field, int64_t n, int D, int P, const double* aux,
                     const double* y0, const double* p, double t0,
                     const double* tape, int64_t tape_cap, const int32_t* n_acc,
                     const double* ckpt, int32_t ckpt_interval,
                     const double* ybar, double* y0bar, double* pbar, int flags,
                     void* stream);

/* Dense output on a recorded Rosenbrock23 mesh: the bracket-and-re-step of gsx_tsit5_saveat
 * with rosenbrock23_step as the re-step (W built at the re-step's size). Same arguments and
 * layouts as gsx_tsit5_saveat; tape and n_acc as gsx_rb23_fwd wrote them. Replaces
 * replay_solve_rosenbrock_saveat_jax (gradsolve/warp/warp_replay.py:78-84 ->
 * gradsolve/solvers/dense.py:106-163), the warp_rosenbrock lane of solve(saveat=...)
 * (gradsolve/api.py:484-489). */
int gsx_rb23_saveat(int field, int64_t n, int D, int P, const double* aux,
                    const double* y0, const double* p, double t0,
                    const double* tape, const int32_t* n_acc, const double* y_final,
                    const double* ts, int64_t nsave, double* ys, int32_t* save_step,
                    double* save_t, double* states, int flags, void* stream);

/* Gradient of a Rosenbrock23 saveat series: the re-steps' VJPs (through their solves and
 * Jacobians) into mu, then the gsx_rb23_adjoint sweep with each save's cotangent added at its
 * bracketing step. Same arguments as gsx_tsit5_adjoint_saveat; tape, ckpt, tape_cap and
 * ckpt_interval as gsx_rb23_fwd wrote them. Replaces jax.grad through
 * replay_solve_rosenbrock_saveat_jax (gradsolve/warp/warp_replay.py:78-84, 364-382). */
int gsx_rb23_adjoint_saveat(int field, int64_t n, int D, int P, const double* aux,
                            const double* y0, const double* p, double t0,
                            const double* tape, int64_t tape_cap, const int32_t* n_acc,
                            const double* ckpt, int32_t ckpt_interval,
                            const double* ts, int64_t nsave, const int32_t* save_step,
                            const double* save_t, const double* states, const double* ysbar,
                            const double* ybar, double* mu, double* y0bar, double* pbar,
                            int flags, void* stream);

/* ---- Rodas5P: the stiff high-order lane ------------------------------------------------
 * Steinebach's Rodas5P (order-5(4) Rosenbrock-Wanner pair, 8 stages, order-4 continuous
 * extension) for fields with an analytic Jacobian (GSX_FIELD_ROBERTSON,
 * GSX_FIELD_LINEAR_DECAY, GSX_FIELD_FORCED_STIFF, and fields registered with
 * GSX_FIELD_STIFF): one LU of W = I - h*gamma*J per step in registers, reused by the 8 stage
 * solves. A field that reads t adds h^2 g_i df/dt to stage i (g_i = sum_j Gamma_ij,
 * rodas5p_step.py:119-146) and evaluates stage i at t + c_i h; J and df/dt are taken at the
 * step's start. Same argument
 * contracts as the Rosenbrock23 entry points. The forward records with the semantics of
 * gradsolve's device recorder for Rodas5P (gradsolve/solvers/record_jax.py:83-132 with
 * RODAS5P_METHOD): status 2 when a step falls below the underflow floor
 * 4*eps*max(|t0|, |t1 - t0|), a non-finite trial shrinks h by 0.2 without a controller update,
 * max_steps counts attempts. Replaces record_rodas5p / replay_rodas5p_jax and its jax.grad
 * (gradsolve/solvers/rodas5p_replay.py:94-176, rodas5p_step.py:119-168), the rodas5p_replay
 * lane of solve/grad_closure (gradsolve/api.py:509-530, 608-616, 683-691). */
int gsx_rodas5p_fwd(int field, int64_t n, int D, int P, const double* aux,
                    const double* y0, const double* p, double t0, double t1,
                    const double* t1_per, double rtol, double atol, int64_t max_steps,
                    double* y_final, int32_t* n_acc, int32_t* n_rej, int32_t* status,
                    double* tape, int64_t tape_cap, double* ckpt, int32_t ckpt_interval,
                    int flags, void* stream);
int gsx_rodas5p_replay(int field, int64_t n, int D, int P, const double* aux,
                       const double* y0, const double* p, double t0,
                       const double* tape, const int32_t* n_acc,
                       double* y_final, double* ckpt, int32_t ckpt_interval, int flags,
                       void* stream);
int gsx_rodas5p_adjoint(int field, int64_t n, int D, int P, const double* aux,
                        const double* y0, const double* p, double t0,
                        const double* tape, int64_t tape_cap, const int32_t* n_acc,
                        const double* ckpt, int32_t ckpt_interval,
                        const double* ybar, double* y0bar, double* pbar, int flags,
                        void* stream);

/* Rodas5P saveat by bracket-and-re-step (the default of the reference's rodas5p_replay lane,
 * dense.vmap_saveat with rodas5p_advance) and its gradient; arguments as gsx_tsit5_saveat /
 * gsx_tsit5_adjoint_saveat. */
int gsx_rodas5p_saveat(int field, int64_t n, int D, int P, const double* aux,
                       const double* y0, const double* p, double t0,
                       const double* tape, const int32_t* n_acc, const double* y_final,
                       const double* ts, int64_t nsave, double* ys, int32_t* save_step,
                       double* save_t, double* states, int flags, void* stream);
int gsx_rodas5p_adjoint_saveat(int field, int64_t n, int D, int P, const double* aux,
                               const double* y0, const double* p, double t0,
                               const double* tape, int64_t tape_cap, const int32_t* n_acc,
                               const double* ckpt, int32_t ckpt_interval,
                               const double* ts, int64_t nsave, const int32_t* save_step,
                               const double* save_t, const double* states,
                               const double* ysbar, const double* ybar, double* mu,
                               double* y0bar, double* pbar, int flags, void* stream);

/* Rodas5P dense output by its order-4 continuous extension (saveat_dense=True:
 * gradsolve/solvers/dense.py:166-227, rodas5p_step.py:102-226): a save time in
 * (t_s, t_s + h_s] is the interpolant (1-th) y_s + th (y_{s+1} + (1-th)(D1 + th (D2 + th D3)))
 * at th = (ts - t_s)/h_s, D_r = sum_i hd[r][i] k_i of step s; at or past the final time it is
 * y_final, at or before t0 it is y0 (save_step -1 / n_acc). The replay forms each step's
 * stages once and no save takes an extra step. ys [nsave][D][n]; save_step / save_t as
 * gsx_tsit5_saveat. */
int gsx_rodas5p_saveat_dense(int field, int64_t n, int D, int P, const double* aux,
                             const double* y0, const double* p, double t0,
                             const double* tape, const int32_t* n_acc, const double* y_final,
                             const double* ts, int64_t nsave, double* ys, int32_t* save_step,
                             double* save_t, int flags, void* stream);

/* Gradient of a dense Rodas5P series: one reverse sweep over the segments, each bracketed
 * save sending (1-th) of its cotangent to y_s, th to y_{s+1} and th (1-th) th^r to D_{r+1},
 * which the step VJP carries to the stages. ysbar [nsave][D][n]; ybar (NULL: zero) dL/dy_final;
 * save_step / save_t as gsx_rodas5p_saveat_dense wrote them; tape, ckpt as gsx_rodas5p_fwd
 * wrote them. Replaces jax.grad through vmap_saveat_dense (gradsolve/solvers/dense.py:230-260). */
int gsx_rodas5p_adjoint_saveat_dense(int field, int64_t n, int D, int P, const double* aux,
                                     const double* y0, const double* p, double t0,
                                     const double* tape, int64_t tape_cap,
                                     const int32_t* n_acc, const double* ckpt,
                                     int32_t ckpt_interval, const double* ts, int64_t nsave,
                                     const int32_t* save_step, const double* save_t,
                                     const double* ysbar, const double* ybar, double* y0bar,
                                     double* pbar, int flags, void* stream);

/* ---- Persistent work-queue form (ensembles with very uneven step counts) ----------------
 * No reference counterpart as a kernel: the reference integrates one trajectory per thread
 * with one t1 per ensemble (gradsolve/warp/warp_ode.py:292) and leaves a wave idle until its
 * slowest lane finishes. Here a fixed grid of `lanes` lanes (a multiple of 256; see
 * gsx_persist_lanes) takes one trial step per lane per iteration and refills idle lanes from
 * an atomic queue (each wavefront claims 64 indices per atomic). Each lane appends to its own
 * stream of segment records rec[rec_cap][lanes][R], R = gsx_persist_record_len(field, D,
 * ckpt_interval) doubles: record r of lane l, at ((r*lanes) + l)*R, holds the state at the
 * start of one checkpoint segment (ckpt_rows doubles; FP32 stores each (value, compensation)
 * pair in the bytes of one double) and then the step sizes of the segment's ckpt_interval
 * accepted steps. R is ckpt_rows + ckpt_interval rounded up to 8, so a record is whole
 * 64-byte sectors: written while its line is in L2 and read whole, where separately strided
 * checkpoints and step sizes would each cost a memory transaction per lane. The
 * trajectories a lane ran go to order[traj_cap][lanes]; lane_count[lanes] counts them; per
 * trajectory j the lane and its first record are lane_of[j] and rec_start[j], so its accepted
 * step s is slot ckpt_rows + s % ckpt_interval of record rec_start[j] + s / ckpt_interval.
 * `queue` is 8 bytes of device memory the call resets (with a kernel, so the sequence can be
 * captured in a graph). A trajectory that does not fit its lane's streams gets status 3; the
 * caller grows the capacities and calls again. flags: 0 or GSX_FP32. t1_per may be NULL.
 * queue_order (NULL: index order) is the order in which the queue hands trajectories out, a
 * permutation of 0..n-1; handing out the longest first (e.g. by t1_per descending) keeps the
 * lanes' last trajectories short, so fewer lanes idle while the longest ones finish. An entry
 * outside [0, n) retires the lane that draws it. */
int64_t gsx_persist_lanes(int waves_per_simd);
int32_t gsx_persist_record_len(int field, int D, int32_t ckpt_interval);
int gsx_tsit5_persist_fwd(int field, int64_t n, int D, int P, const double* aux,
                          const double* y0, const double* p, double t0, double t1,
                          const double* t1_per, double rtol, double atol, int64_t max_steps,
                          double* y_final, int32_t* n_acc, int32_t* n_rej, int32_t* status,
                          int64_t lanes, double* rec, int64_t rec_cap, int32_t ckpt_interval,
                          int32_t* order, int64_t traj_cap, int32_t* lane_count,
                          int32_t* lane_of, int32_t* rec_start, void* queue,
                          const int32_t* queue_order, int flags, void* stream);

/* Adjoint of a persistent record: lane l walks its trajectories (order) backwards, one
 * segment record per iteration (at ckpt_interval 2 the pair adjoint, which re-integrates
 * nothing). Same lanes, records and ckpt_interval as the forward; call only when every
 * status is 0. */
int gsx_tsit5_persist_adjoint(int field, int64_t n, int D, int P, const double* aux,
                              const double* p, double t0, const int32_t* n_acc, int64_t lanes,
                              const double* rec, int32_t ckpt_interval, const int32_t* order,
                              const int32_t* lane_count, const int32_t* rec_start,
                              const double* ybar, double* y0bar, double* pbar, int flags,
                              void* stream);

/* Kuramoto ensemble, D = P = 32: f = omega + (1/32)((K sin y) cos y - (K cos y) sin y) with one
 * coupling matrix K[32*32] (row-major) shared by the ensemble; the coupling sums run on FP64
 * MFMA, 16 trajectories per wavefront. In the reference this problem has no fused kernel
 * (register_jax_field rejects the closed-over K, gradsolve/warp/jax_field.py:417-421), so
 * grad_closure(..., saveat=ts) takes the general lane: record_tsit5_jax
 * (gradsolve/solvers/tsit5_replay.py:127-189) then replay_solve_saveat_jax
 * (gradsolve/warp/warp_replay.py:58-75 -> gradsolve/solvers/dense.py:106-163, 318-335) and
 * jax.grad. These four calls replace that lane.
 *
 * Layouts: y0, omega, y_final, yfbar, y0bar, omega_bar [32][n]; tape [tape_cap][n];
 * states [tape_cap][n][32] (the state at the start of accepted step s); save_step, save_t
 * [nsave][n]; ys, ysbar, mu [n][nsave][32]. ts[nsave] is sorted non-decreasing (validated by
 * the caller as validate_saveat does, dense.py:81-103).
 *
 * flags: 0 (FP64, coupling on FP64 MFMA) or GSX_FP32 (FP32 stages and coupling on FP32 MFMA,
 * Kahan-compensated state and adjoint accumulators, FP64 t, h, norm, controller and buffers);
 * the four calls on one record must use the same flags.
 *
 * gsx_kuramoto_fwd: adaptive Tsit5 forward (as gsx_tsit5_fwd; t0 and t1 shared). With tape
 * non-NULL it records the accepted steps and, when states is non-NULL, every accepted step's
 * start state; with nsave > 0 it records for each save time q the step s bracketing it
 * (t_s < ts[q] <= t_s + h_s) in save_step[q][j] and t_s in save_t[q][j], -1 when
 * ts[q] <= t0 (the save is y0) and -2 when ts[q] >= the final time (the save is y_final).
 * status as gsx_tsit5_fwd (3: the tape is full; grow tape_cap and call again). */
int gsx_kuramoto_fwd(int64_t n, int D, const double* K, const double* y0, const double* omega,
                     double t0, double t1, double rtol, double atol, int64_t max_steps,
                     const double* ts, int64_t nsave, double* y_final, int32_t* n_acc,
                     int32_t* n_rej, int32_t* status, double* tape, int64_t tape_cap,
                     double* states, int32_t* save_step, double* save_t, int flags,
                     void* stream);

/* Frozen-mesh replay from (y0, omega) over a recorded tape: y_final and, when non-NULL, the
 * states (the save brackets depend on the mesh only and stay valid). */
int gsx_kuramoto_replay(int64_t n, int D, const double* K, const double* y0,
                        const double* omega, const double* tape, const int32_t* n_acc,
                        double* y_final, double* states, int flags, void* stream);

/* Saved states ys[j][q]: one Tsit5 step of size ts[q] - save_t from the bracketing step's
 * start state, y0 for save_step -1, y_final for -2 (dense.py:106-163). */
int gsx_kuramoto_saveat(int64_t n, int D, const double* K, const double* y0,
                        const double* omega, const double* y_final, const double* states,
                        const double* ts, int64_t nsave, const int32_t* save_step,
                        const double* save_t, double* ys, int flags, void* stream);

/* Adjoint of (y_final, ys) with respect to (y0, omega) for cotangents yfbar (may be NULL) and
 * ysbar (may be NULL): y0bar, omega_bar [32][n]. mu [n][nsave][32] is scratch the caller
 * provides when ysbar is given (it must not alias ysbar). */
int gsx_kuramoto_adjoint(int64_t n, int D, const double* K, const double* omega,
                         const double* tape, const int32_t* n_acc, const double* states,
                         const double* ts, int64_t nsave, const int32_t* save_step,
                         const double* save_t, const double* ysbar, const double* yfbar,
                         double* mu, double* y0bar, double* omega_bar, int flags,
                         void* stream);

/* The naive reverse sweep, measured by the benchmark as the plain arm (the fastest "plain"
 * drop-in: an FP64 HIP forward plus a tape-based reverse sweep with no checkpoints). Not part of
 * the drop-in boundary; no reference counterpart. The forward is gsx_tsit5_fwd's algorithm
 * that additionally stores every accepted step's six stage inputs in
 * stages[cap][6][D][n]; the adjoint reads them back instead of recomputing. Lorenz only. */
int gsx_plain_tsit5_fwd(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                        const double* p, double t0, double t1, double rtol, double atol,
                        int64_t max_steps, double* y_final, int32_t* n_acc, int32_t* n_rej,
                        int32_t* status, double* tape, double* stages, int64_t cap,
                        void* stream);
int gsx_plain_tsit5_adjoint(int field, int64_t n, int D, int P, const double* aux,
                            const double* p, const double* tape, const double* stages,
                            const int32_t* n_acc, const double* ybar, double* y0bar,
                            double* pbar, void* stream);

/* Export the tape to the reference's zero-padded trajectory-major form
 * out[n][S] (gradsolve/warp/warp_replay.py:180-182): out[j*S+s] = tape[s*n+j] for
 * s < n_acc[j], else 0. */
int gsx_tape_to_padded(int64_t n, const double* tape, int64_t tape_cap,
                       const int32_t* n_acc, int64_t S, double* out, void* stream);

/* Import a reference-format padded tape in[n][S] (zero past each trajectory's accepted
 * count) into the step-major tape layout; n_acc[j] is set to the index after the last
 * non-zero step of row j. */
int gsx_tape_from_padded(int64_t n, const double* in, int64_t S, double* tape,
                         int64_t tape_cap, int32_t* n_acc, void* stream);

/* The reference's trimmed step-major hand-off (gradsolve/warp/warp_ode.py:164-212 hands
 * dt_rec[:S] to JAX over DLPack, relying on rows past n_acc being 0.0): zeroes rows
 * n_acc[j] <= s < S of each column in place, so that the leading S rows of the tape (a
 * contiguous [S][n] block, exportable without a copy) are that record. */
int gsx_tape_zero_tail(int64_t n, double* tape, int64_t tape_cap, const int32_t* n_acc,
                       int64_t S, void* stream);

/* The inverse for an imported trimmed [S][n] tape: n_acc[j] = 1 + the last row with a nonzero
 * step (accepted steps are positive; zero rows are padding). */
int gsx_tape_count(int64_t n, const double* tape, int64_t S, int32_t* n_acc, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GRADSOLVE_AMD_H */
