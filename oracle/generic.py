"""Pure-Python Tsit5 record, replay and adjoint for a right-hand side given as a Python
function (the fields compiled at run time have no counterpart in the C oracle).

TEST INFRASTRUCTURE ONLY: the checker for gsx_field_register's kernels; only ``tests/`` may
import it. Small cases only (Python loops). It restates the same algorithm as the C oracle
(``gs_oracle.cpp``: ``record_one``, ``tsit5_trial``, ``tsit5_step``, ``tsit5_step_vjp``), which
follows gradsolve/solvers/tsit5_replay.py:100-124 (record), gradsolve/solvers/method.py:85-108
(controller) and gradsolve/solvers/tsit5_step.py:80-102 (step): the same textual operation
order on IEEE doubles (Python floats never contract into fused multiply-adds), the WRMS norm
summed by ``numpy.sum`` (the reference's own reduction), ``**`` for the controller power (the C
library's pow, as NumPy's scalar power). The field is evaluated by calling ``f(t, y, p)`` on
float64 NumPy arrays; its vector-Jacobian products are taken by complex-step differentiation
(exact to rounding for analytic right-hand sides), so the adjoint does not depend on the
generated derivative code it checks. Pinned against the C oracle on Lorenz: records equal bit
for bit, adjoints to 1e-13 (tests/test_oracle_generic.py).
"""
from __future__ import annotations

import math

import numpy as np

from . import oracle as _O

_C, _A, _B, _E = _O.tableau()
C2, C3, C4, C5, C6 = (float(x) for x in _C)
(A21, A31, A32, A41, A42, A43, A51, A52, A53, A54,
 A61, A62, A63, A64, A65) = (float(x) for x in _A)
B1, B2, B3, B4, B5, B6 = (float(x) for x in _B)
E1, E2, E3, E4, E5, E6, E7 = (float(x) for x in _E)
SAFETY, ERR_FLOOR, FAC_MIN, FAC_MAX, TSIT5_EXP = 0.9, 1e-10, 0.2, 5.0, -0.2


def _f(f, t, y, p):
    return [float(v) for v in np.asarray(f(t, np.asarray(y, np.float64),
                                           np.asarray(p, np.float64)), np.float64)]


def _stages(f, t, y, h, p, k0):
    """k1..k6 and y5 of one step from k1 = k0 (gs_oracle.cpp tsit5_stages)."""
    D = len(y)
    k = [k0]
    z = [y[i] + h * (A21 * k[0][i]) for i in range(D)]
    k.append(_f(f, t + C2 * h, z, p))
    z = [y[i] + h * (A31 * k[0][i] + A32 * k[1][i]) for i in range(D)]
    k.append(_f(f, t + C3 * h, z, p))
    z = [y[i] + h * (A41 * k[0][i] + A42 * k[1][i] + A43 * k[2][i]) for i in range(D)]
    k.append(_f(f, t + C4 * h, z, p))
    z = [y[i] + h * (A51 * k[0][i] + A52 * k[1][i] + A53 * k[2][i] + A54 * k[3][i])
         for i in range(D)]
    k.append(_f(f, t + C5 * h, z, p))
    z = [y[i] + h * (A61 * k[0][i] + A62 * k[1][i] + A63 * k[2][i] + A64 * k[3][i] +
                     A65 * k[4][i]) for i in range(D)]
    k.append(_f(f, t + C6 * h, z, p))
    y5 = [y[i] + h * (B1 * k[0][i] + B2 * k[1][i] + B3 * k[2][i] + B4 * k[3][i] +
                      B5 * k[4][i] + B6 * k[5][i]) for i in range(D)]
    return k, y5


def _wrms(e, y, y5, rtol, atol):
    D = len(e)
    q = np.empty(D)
    for i in range(D):
        sc = atol + rtol * max(abs(y[i]), abs(y5[i]))
        r = e[i] / sc
        q[i] = r * r
    return math.sqrt(float(np.sum(q)) / float(D))


def tsit5_record_one(f, y0, p, t0, t1, rtol, atol, max_steps=50000):
    """One trajectory: (y_final, tape list, n_acc, n_rej, status) (gs_oracle.cpp record_one)."""
    y = [float(v) for v in y0]
    D = len(y)
    t, dt = float(t0), (float(t1) - float(t0)) / 100.0
    k0 = _f(f, t, y, p)
    tape, acc, rej = [], 0, 0
    for _ in range(max_steps):
        if t >= t1:
            break
        h = min(dt, t1 - t)
        k, y5 = _stages(f, t, y, h, p, k0)
        k.append(_f(f, t + h, y5, p))
        e = [h * (E1 * k[0][i] + E2 * k[1][i] + E3 * k[2][i] + E4 * k[3][i] + E5 * k[4][i] +
                  E6 * k[5][i] + E7 * k[6][i]) for i in range(D)]
        err = _wrms(e, y, y5, rtol, atol)
        if err <= 1.0:
            t = t + h
            y = y5
            k0 = k[6]
            tape.append(h)
            acc += 1
        else:
            rej += 1
        dt = h * min(max(SAFETY * max(err, ERR_FLOOR) ** TSIT5_EXP, FAC_MIN), FAC_MAX)
    status = 0 if t >= t1 else 1
    return np.array(y), np.array(tape), acc, rej, status


def tsit5_record(f, y0, params, t0, t1, *, rtol, atol, max_steps=50000):
    """Ensemble record in the reference's host layout: (y_final [n,D], tape [n,S] zero-padded,
    n_acc, n_rej, status)."""
    n = y0.shape[0]
    outs = [tsit5_record_one(f, y0[j], params[j], t0, t1, rtol, atol, max_steps)
            for j in range(n)]
    S = max((len(o[1]) for o in outs), default=0)
    tape = np.zeros((n, S))
    for j, o in enumerate(outs):
        tape[j, :len(o[1])] = o[1]
    return (np.stack([o[0] for o in outs]), tape, np.array([o[2] for o in outs]),
            np.array([o[3] for o in outs]), np.array([o[4] for o in outs], np.int32))


def _step(f, t, y, h, p):
    """tsit5_step (gs_oracle.cpp): all seven stages, B7 * k7 kept."""
    D = len(y)
    k, y5 = _stages(f, t, y, h, p, _f(f, t, y, p))
    k.append(_f(f, t + C6 * h, y5, p))
    return [y[i] + h * (B1 * k[0][i] + B2 * k[1][i] + B3 * k[2][i] + B4 * k[3][i] +
                        B5 * k[4][i] + B6 * k[5][i] + 0.0 * k[6][i]) for i in range(D)]


def _cs_vjp(f, t, z, p, v):
    """v^T df/dz and v^T df/dp at (t, z, p) by complex steps."""
    hcs = 1e-30
    z = np.asarray(z, np.float64)
    p = np.asarray(p, np.float64)
    gz = np.zeros(z.shape[0])
    gp = np.zeros(p.shape[0])
    for j in range(z.shape[0]):
        zc = z.astype(complex)
        zc[j] += 1j * hcs
        gz[j] = float(np.dot(v, np.imag(np.asarray(f(t, zc, p.astype(complex)))) / hcs))
    for q in range(p.shape[0]):
        pc = p.astype(complex)
        pc[q] += 1j * hcs
        gp[q] = float(np.dot(v, np.imag(np.asarray(f(t, z.astype(complex), pc))) / hcs))
    return gz, gp


def _step_vjp(f, t, y, h, p, lam):
    """(d y_next/d y)^T lam and (d y_next/d p)^T lam for one replay step (gs_oracle.cpp
    tsit5_step_vjp restated on the stage inputs)."""
    D = len(y)
    A = [[0.0] * 6, [A21] + [0.0] * 5, [A31, A32] + [0.0] * 4, [A41, A42, A43, 0.0, 0.0, 0.0],
         [A51, A52, A53, A54, 0.0, 0.0], [A61, A62, A63, A64, A65, 0.0]]
    B = [B1, B2, B3, B4, B5, B6]
    Cn = [0.0, C2, C3, C4, C5, C6]
    z = [list(y)]
    k = [_f(f, t, y, p)]
    for s in range(1, 6):
        zs = [y[i] + h * sum(A[s][r] * k[r][i] for r in range(s)) for i in range(D)]
        z.append(zs)
        k.append(_f(f, t + Cn[s] * h, zs, p))
    kb = [[B[s] * h * lam[i] for i in range(D)] for s in range(6)]
    lam_new = np.array(lam, np.float64)
    pb = np.zeros(len(p))
    for s in range(5, -1, -1):
        zb, pbs = _cs_vjp(f, t + Cn[s] * h, z[s], p, np.array(kb[s]))
        lam_new += zb
        pb += pbs
        for r in range(s):
            for i in range(D):
                kb[r][i] += A[s][r] * h * zb[i]
    return lam_new, pb


def tsit5_replay(f, y0, params, t0, tape):
    """Frozen-mesh replay: y_final [n, D] (padded tape entries 0 are identity steps)."""
    out = []
    for j in range(y0.shape[0]):
        y, t = [float(v) for v in y0[j]], float(t0)
        for h in tape[j]:
            if h == 0.0:
                continue
            y = _step(f, t, y, float(h), params[j])
            t = t + float(h)
        out.append(y)
    return np.array(out)


def tsit5_adjoint(f, y0, params, t0, tape, ybar):
    """dL/dy0 [n, D], dL/dp [n, P] of the frozen-mesh replay for ybar = dL/dy_final."""
    y0b, pb_all = [], []
    for j in range(y0.shape[0]):
        ys, ts, hs = [], [], []
        y, t = [float(v) for v in y0[j]], float(t0)
        for h in tape[j]:
            if h == 0.0:
                continue
            ys.append(y)
            ts.append(t)
            hs.append(float(h))
            y = _step(f, t, y, float(h), params[j])
            t = t + float(h)
        lam = np.array(ybar[j], np.float64)
        pb = np.zeros(params.shape[1])
        for s in range(len(hs) - 1, -1, -1):
            lam, pbs = _step_vjp(f, ts[s], ys[s], hs[s], params[j], lam)
            pb += pbs
        y0b.append(lam)
        pb_all.append(pb)
    return np.array(y0b), np.array(pb_all)
