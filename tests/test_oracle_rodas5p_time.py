"""Pins the oracle's Rodas5P on a right-hand side that reads t (the df/dt term of
gradsolve/solvers/rodas5p_step.py:119-146; CPU only). The field is the oracle's
FORCED_STIFF: f0 = -a (y0 - y1^2) + 0.4 sin 3t, f1 = y0 - y1 - b y1^3 + 0.2 cos t + c sin(t) y0,
which at (a, b, c) = (30, 0.7, 0) is the nonlinear, time-dependent, stiff field of the
reference's tests/test_rodas5p_dense.py:94-103; c != 0 makes J depend on t and df/dt on y
and on a parameter.

* stage vectors: Gamma @ k of one oracle step equals the stage vectors of an independent
  driver in the published transformed (a, C) form (W = J - I/(h gamma), linsolve_tmp =
  f(t + c_i h, u_i) + h d_i f_t + sum_j C_ij/h ks_j, ks_i = W \\ -linsolve_tmp), the check of
  tests/test_rodas5p_dense.py:106-161 with its bounds (abs 1e-14, rel 1e-13);
* observed order 5 on a non-stiff instance (a = 1) with fixed steps against scipy DOP853 at
  rtol 1e-13 (pairwise rates in [4.6, 5.6] above the round-off floor, the bounds of
  tests/test_rodas5p_step.py:145-171); a step without the df/dt term is first order in it;
* the record on the stiff instance against scipy Radau;
* the adjoints (final state, re-step saveat, dense saveat) against torch.autograd through a
  torch restatement whose J and df/dt come from torch.autograd itself (not from the oracle's
  hand-written derivatives), and central differences on the frozen tape in c, which enters
  only through the time-dependent terms.
"""
import importlib.util
import os

import numpy as np
import scipy.integrate as si
import torch

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GEN = os.path.join(HERE, "..", "gradsolve_amd", "csrc", "gen_rodas5p_tableau.py")


def _gen():
    spec = importlib.util.spec_from_file_location("gen_rodas5p_tableau", GEN)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


G = _gen()
ALPHA, GAMMA, MM, _E, HD = G.convert()
GSUM = GAMMA.sum(axis=1)
C_ABS = G.ABSCISSAE
GAM = G.GAMMA_DIAG


def _f_np(t, y, p):
    a, b, c = p
    return np.array([-a * (y[0] - y[1] ** 2) + 0.4 * np.sin(3.0 * t),
                     y[0] - y[1] - b * y[1] ** 3 + 0.2 * np.cos(t) + c * np.sin(t) * y[0]])


def _jac_np(t, y, p):
    a, b, c = p
    return np.array([[-a, 2.0 * a * y[1]], [1.0 + c * np.sin(t), -1.0 - 3.0 * b * y[1] ** 2]])


def _dfdt_np(t, y, p):
    c = p[2]
    return np.array([1.2 * np.cos(3.0 * t), -0.2 * np.sin(t) + c * np.cos(t) * y[0]])


def _ac_stage_vectors(t, y, dt, p):
    """The transformed (a, C) driver from the published stored tables."""
    Cs = np.zeros((8, 8))
    Cs[:, :7] = G.C7
    A = G.A
    W = _jac_np(t, y, p) - np.eye(2) / (dt * GAM)
    dT = _dfdt_np(t, y, p)
    ks = []
    for i in range(8):
        u = y.copy()
        for j in range(i):
            u = u + A[i, j] * ks[j]
        lt = _f_np(t + C_ABS[i] * dt, u, p) + (dt * G.GAMMA_SUMS[i]) * dT
        for j in range(i):
            lt = lt + (Cs[i, j] / dt) * ks[j]
        ks.append(np.linalg.solve(W, -lt))
    return np.array(ks)


def test_stage_vectors_match_transformed_form():
    y = np.array([0.4, 0.6])
    for p in ((30.0, 0.7, 0.0), (30.0, 0.7, 0.5)):
        for t, dt in ((0.3, 0.02), (1.7, 0.1)):
            ks = _ac_stage_vectors(t, y, dt, p)
            k = O.rodas5p_stages(O.FORCED_STIFF, t, y, dt, np.array(p))
            gk = GAMMA @ k
            abs_err = float(np.abs(ks - gk).max())
            assert abs_err <= 1e-14, (p, t, abs_err)
            assert abs_err / float(np.abs(ks).max()) <= 1e-13, (p, t)


def test_observed_order_five_time_dependent():
    p = np.array([[1.0, 0.7, 0.5]])
    y0 = np.array([[0.5, 0.3]])
    T = 2.0
    ref = si.solve_ivp(lambda t, y: _f_np(t, y, p[0]), (0.0, T), y0[0], method="DOP853",
                       rtol=1e-13, atol=1e-15).y[:, -1]
    errs = []
    for n in (4, 8, 16, 32, 64, 128):
        tape = np.full((1, n), T / n)
        yf = O.rodas5p_replay(O.FORCED_STIFF, y0, p, 0.0, tape)
        errs.append(float(np.linalg.norm(yf[0] - ref)))
    rates = [np.log(errs[i] / errs[i + 1]) / np.log(2.0) for i in range(len(errs) - 1)
             if errs[i] > 1e-11 and errs[i + 1] > 1e-11]
    assert len(rates) >= 3, errs
    for r in rates:
        assert 4.6 <= r <= 5.6, rates


def test_stiff_record_against_radau():
    rng = np.random.default_rng(3)
    n = 6
    p = np.column_stack([np.full(n, 1000.0), rng.uniform(0.5, 0.9, n), rng.uniform(-0.5, 0.5, n)])
    y0 = np.column_stack([rng.uniform(0.2, 0.6, n), rng.uniform(0.4, 0.8, n)])
    yf, tape, na, nr, st = O.rodas5p_record(O.FORCED_STIFF, y0, p, 0.0, 5.0, rtol=1e-8,
                                            atol=1e-10)
    assert np.all(st == 0)
    for j in range(n):
        r = si.solve_ivp(lambda t, y: _f_np(t, y, p[j]), (0.0, 5.0), y0[j], method="Radau",
                         rtol=1e-12, atol=1e-14, jac=lambda t, y: _jac_np(t, y, p[j]))
        assert np.max(np.abs(yf[j] - r.y[:, -1])) < 2e-6, (j, yf[j], r.y[:, -1])
        assert abs(tape[j, :na[j]].sum() - 5.0) < 1e-12


def _tf(t, y, p):
    return torch.stack([-p[0] * (y[0] - y[1] * y[1]) + 0.4 * torch.sin(3.0 * t),
                        y[0] - y[1] - p[1] * y[1] ** 3 + 0.2 * torch.cos(t) +
                        p[2] * torch.sin(t) * y[0]])


def _t_step(t, y, h, p):
    tt = torch.tensor(float(t), dtype=torch.float64)
    ft, J = torch.autograd.functional.jacobian(lambda a, b: _tf(a, b, p), (tt, y),
                                               create_graph=True)
    Wm = torch.eye(2, dtype=y.dtype) - (h * GAM) * J
    ks = []
    for i in range(8):
        u, acc = y, torch.zeros_like(y)
        for j in range(i):
            u = u + float(ALPHA[i, j]) * ks[j]
            acc = acc + float(GAMMA[i, j]) * ks[j]
        rhs = h * _tf(tt + float(C_ABS[i]) * h, u, p) + h * (J @ acc) + \
            (h * h * float(GSUM[i])) * ft
        ks.append(torch.linalg.solve(Wm, rhs))
    out = y
    for i in range(8):
        out = out + float(MM[i]) * ks[i]
    return out, ks


def _t_saveat(y0, p, tape, ts, dense):
    y, t = y0, 0.0
    hits = {}
    for s in range(len(tape)):
        h = float(tape[s])
        yn, ks = _t_step(t, y, h, p)
        for q, tq in enumerate(ts):
            if t < tq <= t + h:
                hits[q] = (y, yn, ks, t, h)
        y, t = yn, t + h
    out = []
    for q, tq in enumerate(ts):
        if tq >= t:
            out.append(y)
        elif q in hits:
            yl, yr, ks, tl, h = hits[q]
            if dense:
                dv = [sum(float(HD[r, i]) * ks[i] for i in range(8)) for r in range(3)]
                th = (tq - tl) / h
                om = 1.0 - th
                out.append(om * yl + th * (yr + om * (dv[0] + th * (dv[1] + th * dv[2]))))
            else:
                out.append(_t_step(tl, yl, tq - tl, p)[0])
        else:
            out.append(y0)
    return y, torch.stack(out)


def test_adjoints_against_autograd_and_fd():
    p = np.array([[30.0, 0.7, 0.5], [200.0, 0.6, -0.4]])
    y0 = np.array([[0.4, 0.6], [0.3, 0.5]])
    t1 = 2.0
    _, tape, na, _, st = O.rodas5p_record(O.FORCED_STIFF, y0, p, 0.0, t1, rtol=1e-7, atol=1e-9)
    assert np.all(st == 0)
    rng = np.random.default_rng(5)
    ybar = rng.standard_normal((2, 2))
    g0, gp = O.rodas5p_adjoint(O.FORCED_STIFF, y0, p, 0.0, tape, ybar)
    ts = np.array([0.0, 1e-3, 0.45, 1.1, 1.95, 2.0])
    ysbar = rng.standard_normal((2, ts.shape[0], 2))
    sg = {d: O.rodas5p_adjoint_saveat(O.FORCED_STIFF, y0, p, 0.0, tape, ts, ysbar, ybar, dense=d)
          for d in (False, True)}
    for j in range(2):
        z = torch.tensor(y0[j], requires_grad=True)
        pt = torch.tensor(p[j], requires_grad=True)
        y, t = z, 0.0
        for s in range(na[j]):
            y = _t_step(t, y, float(tape[j, s]), pt)[0]
            t += float(tape[j, s])
        yo = O.rodas5p_replay(O.FORCED_STIFF, y0[j:j + 1], p[j:j + 1], 0.0, tape[j:j + 1])
        assert np.max(np.abs(y.detach().numpy() - yo[0])) < 1e-13
        gz, gpt = torch.autograd.grad((y * torch.tensor(ybar[j])).sum(), (z, pt))
        assert np.max(np.abs(gpt.numpy() - gp[j])) < 1e-9 * np.max(np.abs(gp[j]))
        assert np.max(np.abs(gz.numpy() - g0[j])) < 1e-9 * np.max(np.abs(g0[j]))
        for dense in (False, True):
            z = torch.tensor(y0[j], requires_grad=True)
            pt = torch.tensor(p[j], requires_grad=True)
            yf, ys = _t_saveat(z, pt, tape[j, :na[j]], ts, dense)
            _, oys = O.rodas5p_replay_saveat(O.FORCED_STIFF, y0[j:j + 1], p[j:j + 1], 0.0,
                                             tape[j:j + 1], ts, dense=dense)
            assert np.max(np.abs(ys.detach().numpy() - oys[0])) < 1e-13, dense
            loss = (ys * torch.tensor(ysbar[j])).sum() + (yf * torch.tensor(ybar[j])).sum()
            gz, gpt = torch.autograd.grad(loss, (z, pt))
            o0, op = sg[dense]
            assert np.max(np.abs(gpt.numpy() - op[j])) < 1e-9 * np.max(np.abs(op[j])), dense
            assert np.max(np.abs(gz.numpy() - o0[j])) < 1e-9 * np.max(np.abs(o0[j])), dense

    def loss_of(pp, dense):
        yf_, ys_ = O.rodas5p_replay_saveat(O.FORCED_STIFF, y0[:1], pp, 0.0, tape[:1], ts,
                                           dense=dense)
        return np.sum(ys_[0] * ysbar[0]) + np.sum(yf_[0] * ybar[0])

    for dense in (False, True):
        errs = []
        for eps in (1e-4, 1e-5, 1e-6):
            pp, pm = p[:1].copy(), p[:1].copy()
            pp[0, 2] += eps
            pm[0, 2] -= eps
            fd = (loss_of(pp, dense) - loss_of(pm, dense)) / (2 * eps)
            errs.append(abs(fd - sg[dense][1][0, 2]) / abs(sg[dense][1][0, 2]))
        assert min(errs) < 1e-6, (dense, errs)
