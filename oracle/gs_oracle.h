/*
 * gs_oracle.h -- CPU oracle for the ensemble ODE record-and-replay path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: it restates, in plain
 * IEEE-754 binary64 arithmetic with floating-point contraction disabled, the
 * algorithms of gradsolve's adaptive Tsit5 / Rosenbrock23 recorders, their frozen-step
 * replays and the discrete adjoint of those replays. Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py may load it. No product code path calls it.
 *
 * Layout conventions (the reference's host layout, gradsolve/base.py:98-152):
 *   y0, y_final, y0bar  : [n, D] row-major (trajectory-major)
 *   p, pbar             : [n, P] row-major
 *   tape                : [n, cap] row-major, accepted step sizes of trajectory j in
 *                         tape[j*cap + 0 .. n_acc[j]-1], zero past n_acc[j]
 *                         (gradsolve/warp/warp_replay.py:180-182).
 *   t1                  : [n] per-trajectory end time (the reference has one t1 per
 *                         ensemble; a constant array reproduces it).
 *
 * Field ids (GSO_F_*) select the right-hand side; `aux` carries shared field data
 * (the DxD coupling matrix K for Kuramoto, NULL otherwise).
 *
 * Return value of every entry point: 0 on success, a negative value for an invalid
 * argument (unknown field, bad dimension).
 */
#ifndef GS_ORACLE_H
#define GS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GSO_F_LORENZ = 1,        /* D=3, P=1 (rho); sigma=10, beta=8/3 (_warp_kernel.py:111-130) */
  GSO_F_LINEAR_DECAY = 2,  /* any D, P=1: f = -p0*y */
  GSO_F_ROBERTSON = 3,     /* D=3, P=3 (k1,k2,k3) (_warp_rosenbrock.py:96-139) */
  GSO_F_KURAMOTO = 4,      /* any D, P=D (omega); aux = K[D*D] row-major */
  GSO_F_VDP = 5,           /* D=2, P=1 (mu) (_warp_kernel.py:158-174) */
  GSO_F_LORENZ3 = 6,       /* D=3, P=3 (sigma, beta, rho) (warp_ode.py:552-589 columns) */
  GSO_F_LINEAR_LADDER = 7, /* any D, P ignored: f = 1.01*y (_warp_kernel.py:133-155) */
  GSO_F_FORCED_STIFF = 8   /* D=2, P=3 (a, b, c), time-dependent, stiff for large a:
                              f0 = -a (y0 - y1^2) + 0.4 sin 3t,
                              f1 = y0 - y1 - b y1^3 + 0.2 cos t + c sin(t) y0;
                              at (a, b, c) = (30, 0.7, 0) the field of
                              tests/test_rodas5p_dense.py:94-103 */
};

/* Tsitouras (2011) tableau as the oracle holds it: c2..c6, a21..a65 (15), b1..b6, e1..e7. */
int gso_tsit5_tableau(double* c /*5*/, double* a /*15*/, double* b /*6*/, double* e /*7*/);

/* Adaptive Tsit5 forward + accepted-h record (tsit5_replay.py:62-124).
 * dt0: [n] first trial step (NULL -> (t1-t0)/100 per trajectory).
 * tape may be NULL (no record). status: 0 reached t1, 1 exhausted max_steps,
 * 3 tape capacity exceeded (the recorder stops that trajectory). */
int gso_tsit5_record(int field, int64_t n, int D, int P, const double* aux,
                     const double* y0, const double* p, double t0, const double* t1,
                     const double* dt0, double rtol, double atol, int64_t max_steps,
                     double* y_final, int64_t* n_acc, int64_t* n_rej, int32_t* status,
                     double* tape, int64_t cap, int nthreads);

/* Teacher-forced check of a recorded tape: for every accepted step s of trajectory j,
 * starting from the state reached by stepping the *given* tape (tape[j*S + 0..s-1]), runs the
 * oracle's controller (attempts until one is accepted) and writes the step size it accepts
 * to h_oracle[j*S + s]. Each entry of a device tape is thereby compared with the oracle given
 * the same history, independently of how the controller amplifies earlier last-ulp
 * differences along a free-running trajectory. rej_at [n, S] (may be NULL) receives the
 * number of attempts the oracle rejected before accepting entry s. */
int gso_tsit5_teacher_forced(int field, int64_t n, int D, int P, const double* aux,
                             const double* y0, const double* p, double t0, const double* t1,
                             double rtol, double atol, const double* tape, int64_t S,
                             const int64_t* n_acc, double* h_oracle, int64_t* n_rej_oracle,
                             int64_t* rej_at, int nthreads);

/* Frozen-step replay with tsit5_step (tsit5_step.py:80-102; warp_replay.py:30-55). */
int gso_tsit5_replay(int field, int64_t n, int D, int P, const double* aux,
                     const double* y0, const double* p, double t0,
                     const double* tape, int64_t S, double* y_final, int nthreads);

/* Discrete adjoint (vector-Jacobian product) of gso_tsit5_replay:
 * given ybar = dL/dy_final [n,D], returns dL/dy0 [n,D] and dL/dp [n,P]. */
int gso_tsit5_adjoint(int field, int64_t n, int D, int P, const double* aux,
                      const double* y0, const double* p, double t0,
                      const double* tape, int64_t S, const double* ybar,
                      double* y0bar, double* pbar, int nthreads);

/* Bracket-and-re-step dense output on the frozen mesh (dense.py:106-163):
 * ts [k] sorted within [t0, t1]; ys [n, k, D]. */
int gso_tsit5_replay_saveat(int field, int64_t n, int D, int P, const double* aux,
                            const double* y0, const double* p, double t0,
                            const double* tape, int64_t S, const double* ts, int64_t k,
                            double* y_final, double* ys, int nthreads);

/* Adjoint of gso_tsit5_replay_saveat for cotangents ysbar [n,k,D] and yfbar [n,D]
 * (either may be NULL = zero). */
int gso_tsit5_adjoint_saveat(int field, int64_t n, int D, int P, const double* aux,
                             const double* y0, const double* p, double t0,
                             const double* tape, int64_t S, const double* ts, int64_t k,
                             const double* ysbar, const double* yfbar,
                             double* y0bar, double* pbar, int nthreads);

/* Rosenbrock23 dense output: the same bracket-and-re-step with rosenbrock23_step as the
 * re-step (gradsolve/warp/warp_replay.py:78-84), and its adjoint. */
int gso_rb23_replay_saveat(int field, int64_t n, int D, int P, const double* aux,
                           const double* y0, const double* p, double t0,
                           const double* tape, int64_t S, const double* ts, int64_t k,
                           double* y_final, double* ys, int nthreads);
int gso_rb23_adjoint_saveat(int field, int64_t n, int D, int P, const double* aux,
                            const double* y0, const double* p, double t0,
                            const double* tape, int64_t S, const double* ts, int64_t k,
                            const double* ysbar, const double* yfbar,
                            double* y0bar, double* pbar, int nthreads);

/* Rodas5P (Steinebach 2023; gradsolve/solvers/rodas5p_step.py, rodas5p_replay.py) for
 * autonomous stiff fields: record with the device recorder's Rodas5P semantics
 * (gradsolve/solvers/record_jax.py:83-132: status 2 at the underflow floor, a non-finite
 * trial shrinks h by 0.2), replay, adjoint, bracket-and-re-step saveat and its adjoint, and the
 * dense saveat of the order-4 continuous extension (dense.py:166-227) and its adjoint. Same
 * argument conventions as the Tsit5 / Rosenbrock23 functions. */
int gso_rodas5p_record(int field, int64_t n, int D, int P, const double* aux,
                       const double* y0, const double* p, double t0, const double* t1,
                       const double* dt0, double rtol, double atol, int64_t max_steps,
                       double* y_final, int64_t* n_acc, int64_t* n_rej, int32_t* status,
                       double* tape, int64_t cap, int nthreads);
int gso_rodas5p_replay(int field, int64_t n, int D, int P, const double* aux,
                       const double* y0, const double* p, double t0, const double* tape,
                       int64_t S, double* y_final, int nthreads);
int gso_rodas5p_adjoint(int field, int64_t n, int D, int P, const double* aux,
                        const double* y0, const double* p, double t0, const double* tape,
                        int64_t S, const double* ybar, double* y0bar, double* pbar,
                        int nthreads);
int gso_rodas5p_replay_saveat(int field, int64_t n, int D, int P, const double* aux,
                              const double* y0, const double* p, double t0, const double* tape,
                              int64_t S, const double* ts, int64_t k, double* y_final,
                              double* ys, int nthreads);
int gso_rodas5p_adjoint_saveat(int field, int64_t n, int D, int P, const double* aux,
                               const double* y0, const double* p, double t0,
                               const double* tape, int64_t S, const double* ts, int64_t k,
                               const double* ysbar, const double* yfbar, double* y0bar,
                               double* pbar, int nthreads);
int gso_rodas5p_replay_saveat_dense(int field, int64_t n, int D, int P, const double* aux,
                                    const double* y0, const double* p, double t0,
                                    const double* tape, int64_t S, const double* ts, int64_t k,
                                    double* y_final, double* ys, int nthreads);
int gso_rodas5p_adjoint_saveat_dense(int field, int64_t n, int D, int P, const double* aux,
                                     const double* y0, const double* p, double t0,
                                     const double* tape, int64_t S, const double* ts, int64_t k,
                                     const double* ysbar, const double* yfbar, double* y0bar,
                                     double* pbar, int nthreads);
int gso_rodas5p_tableau(double* alpha, double* gamma_off, double* c, double* d, double* m,
                        double* e, double* hd, double* gamma);
/* The eight stage vectors k [8][D] of one Rodas5P step from (t, y) with step h. */
int gso_rodas5p_stages(int field, int D, int P, const double* aux, double t, const double* y,
                       double h, const double* p, double* k);

/* Adaptive Rosenbrock23 forward + record (gradsolve_g2_spike_oracle.py:138-184;
 * _warp_rosenbrock.py:383-452), W = I - d*h*J factored once per step. */
int gso_rb23_record(int field, int64_t n, int D, int P, const double* aux,
                    const double* y0, const double* p, double t0, const double* t1,
                    const double* dt0, double rtol, double atol, int64_t max_steps,
                    double* y_final, int64_t* n_acc, int64_t* n_rej, int32_t* status,
                    double* tape, int64_t cap, int nthreads);

/* Frozen-step Rosenbrock23 replay (rosenbrock23_step.py:47-66). */
int gso_rb23_replay(int field, int64_t n, int D, int P, const double* aux,
                    const double* y0, const double* p, double t0,
                    const double* tape, int64_t S, double* y_final, int nthreads);

/* Discrete adjoint of gso_rb23_replay (differentiates through W, J(y,p)). */
int gso_rb23_adjoint(int field, int64_t n, int D, int P, const double* aux,
                     const double* y0, const double* p, double t0,
                     const double* tape, int64_t S, const double* ybar,
                     double* y0bar, double* pbar, int nthreads);

/* Field evaluation and its vector-Jacobian product at one point (for tests). */
int gso_field(int field, int D, int P, const double* aux, double t, const double* y,
              const double* p, double* dy);
int gso_field_vjp(int field, int D, int P, const double* aux, double t, const double* y,
                  const double* p, const double* v, double* ybar, double* pbar);

/* Algorithmic operation counts with an op-counting scalar: add/sub/mul/div/sqrt = 1,
 * a fused multiply-add would be 2 (none is formed: contraction is off); pow counted
 * apart. which: 0 = one Tsit5 trial step (FSAL: 6 field calls), 1 = one replay step
 * (tsit5_step, 7 field calls), 2 = one adjoint step (stage recompute + VJP, no k7),
 * 3 = one Rosenbrock23 trial step, 4 = one Rosenbrock23 adjoint step, 5 = one Rodas5P trial
 * step (J, one LU, 8 stages, error norm and controller), 6 = one Rodas5P adjoint step. */
int gso_count_ops(int which, int field, int D, int P, const double* aux,
                  double* flops, double* pows);

#ifdef __cplusplus
}
#endif
#endif
