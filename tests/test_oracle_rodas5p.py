"""Pins the oracle's Rodas5P (gradsolve/solvers/rodas5p_step.py, rodas5p_replay.py; CPU only):

* the tableau: the committed header is the generator's output; alpha's row sums are the
  published abscissae and Gamma's (with gamma on the diagonal) the published d; sum m = 1;
* order: one step on y' = -k y against exp(-k h) has local error ~ h^6 (order 5), and one
  nonlinear Robertson step against scipy Radau at rtol 1e-13 likewise, which a transposed or
  misplaced alpha/Gamma entry breaks;
* the record: Robertson on [0, 100] (rtol 1e-6, atol 1e-10) against Radau; the recorder's
  statuses (1 on max_steps, 2 at the underflow floor after non-finite trials, record_jax.py:83-132);
* the dense output: a save on a mesh node equals the replayed state (<= 4 eps relative,
  rodas5p_step.py:108-112), and the interpolant at theta = 1/2 of one step on y' = -k y has
  error ~ h^5 (order 4); dense and re-step saves agree to the tolerance on Robertson;
* the adjoints (final state, re-step saveat, dense saveat) against torch.autograd through a
  torch restatement of the same steps (float64, torch.linalg.solve, the analytic Jacobian
  written in torch so second derivatives flow) and central differences on the frozen tape.
"""
import importlib.util
import os

import numpy as np
import scipy.integrate as si
import torch

from gradsolve_amd import workloads as W
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GEN = os.path.join(HERE, "..", "gradsolve_amd", "csrc", "gen_rodas5p_tableau.py")


def _gen():
    spec = importlib.util.spec_from_file_location("gen_rodas5p_tableau", GEN)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_tableau_header_and_published_sums():
    G = _gen()
    with open(os.path.join(os.path.dirname(GEN), "gsx_rodas5p_tab.h")) as fh:
        assert fh.read() == G.header()
    alpha, goff, c, d, m, e, hd, gam = O.rodas5p_tableau()
    ga, gg, gm, ge, ghd = G.convert()
    assert np.array_equal(np.tril(alpha, -1), np.tril(ga, -1))
    assert np.array_equal(goff, np.tril(gg, -1)) and np.array_equal(m, gm)
    assert np.array_equal(e, ge) and np.array_equal(hd, ghd)
    assert gam == G.GAMMA_DIAG
    assert np.max(np.abs(np.tril(alpha, -1).sum(1) - c)) < 1e-13
    assert np.max(np.abs(goff.sum(1) + gam - d)) < 1e-13
    assert abs(m.sum() - 1.0) < 1e-13 and abs(e.sum()) < 1e-13
    assert np.max(np.abs(hd.sum(1))) < 1e-12  # the extension adds nothing to a constant


def _slope(hs, errs):
    return np.polyfit(np.log(hs), np.log(errs), 1)[0]


def test_order_five_linear_and_nonlinear():
    k = np.array([[2.0]])
    y0 = np.array([[1.0, -0.5]])
    hs = np.array([0.2, 0.1, 0.05, 0.025])
    errs = [np.max(np.abs(O.rodas5p_replay(O.LINEAR_DECAY, y0, k, 0.0, np.array([[h]]))[0] -
                          y0[0] * np.exp(-2.0 * h))) for h in hs]
    assert _slope(hs, errs) > 5.6, errs  # order >= 5 (the linear stability function matches e^z further)
    # a nonlinear step: Robertson kinetics with non-stiff rates (stiff rates would put these
    # steps many time constants deep, where Rosenbrock methods lose order by design)
    p = np.array([[0.6, 3.0, 1.5]])
    yA = np.array([[0.7, 0.2, 0.1]])

    def f(t, y):
        return [-p[0, 0] * y[0] + p[0, 2] * y[1] * y[2],
                p[0, 0] * y[0] - p[0, 2] * y[1] * y[2] - p[0, 1] * y[1] ** 2,
                p[0, 1] * y[1] ** 2]

    hs = np.array([0.2, 0.1, 0.05, 0.025])
    errs = []
    for h in hs:
        ref = si.solve_ivp(f, (0.0, h), yA[0], method="Radau", rtol=1e-14, atol=1e-16).y[:, -1]
        got = O.rodas5p_replay(O.ROBERTSON, yA, p, 0.0, np.array([[h]]))[0]
        errs.append(np.max(np.abs(got - ref)))
    assert _slope(hs, errs) > 5.5, errs


def test_robertson_record_against_radau_and_statuses():
    e = W.robertson(3, seed=5)
    yf, tape, na, nr, st = O.rodas5p_record(O.ROBERTSON, e.y0, e.params, 0.0, 100.0, rtol=1e-6,
                                            atol=1e-10)
    assert np.all(st == 0) and np.all(na > 10)
    assert np.allclose(tape.sum(axis=1), 100.0, rtol=1e-12)
    for j in range(3):
        k1, k2, k3 = e.params[j]
        ref = si.solve_ivp(lambda t, y: [-k1 * y[0] + k3 * y[1] * y[2],
                                         k1 * y[0] - k3 * y[1] * y[2] - k2 * y[1] ** 2,
                                         k2 * y[1] ** 2], (0.0, 100.0), e.y0[j],
                           method="Radau", rtol=1e-12, atol=1e-14).y[:, -1]
        assert abs(yf[j, 0] - ref[0]) / ref[0] < 1e-5 and abs(yf[j, 2] - ref[2]) / ref[2] < 1e-5
    assert np.array_equal(yf, O.rodas5p_replay(O.ROBERTSON, e.y0, e.params, 0.0, tape))
    _, _, _, _, st1 = O.rodas5p_record(O.ROBERTSON, e.y0, e.params, 0.0, 100.0, rtol=1e-6,
                                       atol=1e-10, max_steps=5)
    assert np.all(st1 == 1)
    # y' = 800 y overflows: the non-finite trials shrink h by 0.2 until the underflow floor
    _, _, na2, nr2, st2 = O.rodas5p_record(O.LINEAR_DECAY, np.array([[1.0]]), np.array([[-800.0]]),
                                           0.0, 10.0, rtol=1e-6, atol=1e-6)
    assert st2[0] == 2 and nr2[0] > 0


def test_dense_output_nodes_and_order():
    e = W.robertson(2, seed=6)
    _, tape, na, _, _ = O.rodas5p_record(O.ROBERTSON, e.y0, e.params, 0.0, 10.0, rtol=1e-7,
                                         atol=1e-11)
    j = 0
    mesh = np.cumsum(tape[j, :na[j]])
    ts = np.sort(np.array([0.0, 1e-4, mesh[4], 0.37, mesh[30], 4.2, 10.0]))
    yf, ys = O.rodas5p_replay_saveat(O.ROBERTSON, e.y0, e.params, 0.0, tape, ts, dense=True)
    _, ysr = O.rodas5p_replay_saveat(O.ROBERTSON, e.y0, e.params, 0.0, tape, ts)
    assert np.array_equal(ys[:, 0], e.y0) and np.array_equal(ys[:, -1], yf)
    for s in (4, 30):
        q = int(np.nonzero(ts == mesh[s])[0][0])
        node = O.rodas5p_replay(O.ROBERTSON, e.y0[j:j + 1], e.params[j:j + 1], 0.0,
                                tape[j:j + 1, :s + 1])[0]
        assert np.max(np.abs(ys[j, q] - node) / np.abs(node)) <= 4 * np.finfo(float).eps
    assert np.max(np.abs(ys - ysr)) < 1e-7
    # the interpolant's order: one step on y' = -k y, theta = 1/2
    k = np.array([[2.0]])
    y0 = np.array([[1.0, -0.5]])
    hs = np.array([0.2, 0.1, 0.05, 0.025])
    errs = []
    for h in hs:
        _, yd = O.rodas5p_replay_saveat(O.LINEAR_DECAY, y0, k, 0.0, np.array([[h]]),
                                        np.array([h / 2]), dense=True)
        errs.append(np.max(np.abs(yd[0, 0] - y0[0] * np.exp(-2.0 * h / 2))))
    assert 4.6 < _slope(hs, errs) < 5.6, errs


# ---- torch restatement for the adjoint tests -----------------------------------------------
ALPHA, GOFF, CC, DD, MM, EE, HD, GAM = O.rodas5p_tableau()


def _rob_f(y, p):
    return torch.stack([-p[0] * y[0] + p[2] * y[1] * y[2],
                        p[0] * y[0] - p[2] * y[1] * y[2] - p[1] * y[1] * y[1],
                        p[1] * y[1] * y[1]])


def _rob_jac(y, p):
    z = torch.zeros((), dtype=y.dtype)
    return torch.stack([
        torch.stack([-p[0], p[2] * y[2], p[2] * y[1]]),
        torch.stack([p[0], -p[2] * y[2] - 2.0 * p[1] * y[1], -p[2] * y[1]]),
        torch.stack([z, 2.0 * p[1] * y[1], z])])


def _t_stages(y, h, p):
    J = _rob_jac(y, p)
    Wm = torch.eye(3, dtype=y.dtype) - (h * GAM) * J
    ks = []
    for i in range(8):
        u, acc = y, torch.zeros_like(y)
        for j in range(i):
            u = u + float(ALPHA[i, j]) * ks[j]
            acc = acc + float(GOFF[i, j]) * ks[j]
        ks.append(torch.linalg.solve(Wm, h * _rob_f(u, p) + h * (J @ acc)))
    return ks


def _t_step(y, h, p):
    ks = _t_stages(y, h, p)
    out = y
    for i in range(8):
        out = out + float(MM[i]) * ks[i]
    return out, ks


def _t_saveat(y0, p, tape, ts, dense):
    y, t = y0, 0.0
    hits = {}
    for s in range(len(tape)):
        h = float(tape[s])
        yn, ks = _t_step(y, h, p)
        for q, tq in enumerate(ts):
            if h > 0 and t < tq <= t + h:
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
                out.append(_t_step(yl, tq - tl, p)[0])
        else:
            out.append(y0)
    return y, torch.stack(out)


def test_adjoints_against_autograd_and_fd():
    e = W.robertson(2, seed=7)
    _, tape, na, _, _ = O.rodas5p_record(O.ROBERTSON, e.y0, e.params, 0.0, 3.0, rtol=1e-6,
                                         atol=1e-10)
    rng = np.random.default_rng(8)
    ybar = rng.standard_normal((2, 3))
    g0, gp = O.rodas5p_adjoint(O.ROBERTSON, e.y0, e.params, 0.0, tape, ybar)
    ts = np.array([0.0, 1e-3, 0.3, 1.7, 2.9, 3.0])
    ysbar = rng.standard_normal((2, ts.shape[0], 3))
    sg = {d: O.rodas5p_adjoint_saveat(O.ROBERTSON, e.y0, e.params, 0.0, tape, ts, ysbar, ybar,
                                      dense=d) for d in (False, True)}
    for j in range(2):
        z = torch.tensor(e.y0[j], requires_grad=True)
        p = torch.tensor(e.params[j], requires_grad=True)
        y = z
        for s in range(na[j]):
            y = _t_step(y, float(tape[j, s]), p)[0]
        gz, gpt = torch.autograd.grad((y * torch.tensor(ybar[j])).sum(), (z, p))
        assert np.max(np.abs(gpt.numpy() - gp[j]) / np.abs(gp[j])) < 1e-9
        assert np.max(np.abs(gz.numpy() - g0[j])) < 1e-9 * np.max(np.abs(g0[j]))
        for dense in (False, True):
            z = torch.tensor(e.y0[j], requires_grad=True)
            p = torch.tensor(e.params[j], requires_grad=True)
            yf, ys = _t_saveat(z, p, tape[j, :na[j]], ts, dense)
            _, oys = O.rodas5p_replay_saveat(O.ROBERTSON, e.y0[j:j + 1], e.params[j:j + 1], 0.0,
                                             tape[j:j + 1], ts, dense=dense)
            assert np.max(np.abs(ys.detach().numpy() - oys[0])) < 1e-13
            loss = (ys * torch.tensor(ysbar[j])).sum() + (yf * torch.tensor(ybar[j])).sum()
            gz, gpt = torch.autograd.grad(loss, (z, p))
            o0, op = sg[dense]
            assert np.max(np.abs(gpt.numpy() - op[j]) / np.abs(op[j])) < 1e-9, dense
            assert np.max(np.abs(gz.numpy() - o0[j])) < 1e-9 * np.max(np.abs(o0[j])), dense

    def loss_of(pp, dense):
        yf_, ys_ = O.rodas5p_replay_saveat(O.ROBERTSON, e.y0[:1], pp, 0.0, tape[:1], ts,
                                           dense=dense)
        return np.sum(ys_[0] * ysbar[0]) + np.sum(yf_[0] * ybar[0])

    for dense in (False, True):
        for q in (0, 2):
            errs = []
            for rel in (1e-4, 1e-5, 1e-6):
                eps = rel * abs(e.params[0, q])
                pp, pm = e.params[:1].copy(), e.params[:1].copy()
                pp[0, q] += eps
                pm[0, q] -= eps
                fd = (loss_of(pp, dense) - loss_of(pm, dense)) / (2 * eps)
                errs.append(abs(fd - sg[dense][1][0, q]) / abs(sg[dense][1][0, q]))
            assert min(errs) < 1e-5, (dense, q, errs)
