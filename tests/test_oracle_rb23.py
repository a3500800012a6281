"""Pins the oracle's Rosenbrock23 record, replay and adjoint (CPU only): order of the
propagated solution against expm, dt = 0 identity, a stiff reference solution (scipy Radau),
and the adjoint against an independent reverse mode (torch.autograd through the replay with
torch.linalg.solve, float64) and central differences."""
import numpy as np
import scipy.integrate as si
import scipy.linalg as sl
import torch

from gradsolve_amd import workloads as W
from oracle import oracle as O

D_COEF = 1.0 / (2.0 + np.sqrt(2.0))


def _torch_rb23_step(f, jac, y, h, p):
    """rosenbrock23_step (gradsolve/solvers/rosenbrock23_step.py:47-66) in torch."""
    Dm = y.shape[0]
    W_ = torch.eye(Dm, dtype=y.dtype) - h * D_COEF * jac(y, p)
    F0 = f(y, p)
    k1 = torch.linalg.solve(W_, F0)
    F1 = f(y + 0.5 * h * k1, p)
    k2 = torch.linalg.solve(W_, F1 - k1) + k1
    return y + h * k2


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


def test_local_order_two_linear():
    """Propagated solution is order 2: local error ~ h^3 on y' = -k y
    (gradsolve tests/test_steps.py:114-125 uses expm)."""
    k = np.array([[3.0]])
    y0 = np.array([[1.0, 2.0]])
    hs = np.array([0.04, 0.02, 0.01, 0.005])
    errs = []
    for h in hs:
        y1 = O.rb23_replay(O.LINEAR_DECAY, y0, k, 0.0, np.array([[h]]))[0]
        ref = sl.expm(-3.0 * h * np.eye(2)) @ y0[0]
        errs.append(np.linalg.norm(y1 - ref))
    slope = np.polyfit(np.log(hs), np.log(errs), 1)[0]
    assert 2.7 < slope < 3.3, slope


def test_dt_zero_identity():
    e = W.robertson(4, seed=1)
    yf = O.rb23_replay(O.ROBERTSON, e.y0, e.params, 0.0, np.zeros((4, 3)))
    assert np.array_equal(yf, e.y0)


def test_robertson_against_radau():
    e = W.robertson(8, seed=2)
    yf, tape, na, nr, st = O.rb23_record(O.ROBERTSON, e.y0, e.params, 0.0, 100.0, rtol=1e-6,
                                         atol=1e-10)
    assert np.all(st == 0)
    assert np.max(np.abs(tape.sum(axis=1) - 100.0)) < 1e-10
    for j in range(8):
        k1, k2, k3 = e.params[j]

        def f(t, y):
            return [-k1 * y[0] + k3 * y[1] * y[2], k1 * y[0] - k3 * y[1] * y[2] - k2 * y[1] ** 2,
                    k2 * y[1] ** 2]
        ref = si.solve_ivp(f, (0, 100), e.y0[j], method="Radau", rtol=1e-12,
                           atol=1e-14).y[:, -1]
        assert abs(yf[j, 0] - ref[0]) / ref[0] < 1e-4
        assert abs(yf[j, 2] - ref[2]) / ref[2] < 1e-4
        assert abs(yf[j, 1] - ref[1]) / ref[1] < 1e-3


def test_adjoint_matches_torch_autograd():
    e = W.robertson(4, seed=3)
    yf, tape, na, nr, st = O.rb23_record(O.ROBERTSON, e.y0, e.params, 0.0, 100.0, rtol=1e-6,
                                         atol=1e-10)
    w = np.array([1.0, 10.0, 1.0])
    g0, gp = O.rb23_adjoint(O.ROBERTSON, e.y0, e.params, 0.0, tape, np.tile(w, (4, 1)))
    for j in range(4):
        y = torch.tensor(e.y0[j], requires_grad=True)
        p = torch.tensor(e.params[j], requires_grad=True)
        z = y
        for s in range(na[j]):
            z = _torch_rb23_step(_rob_f, _rob_jac, z, float(tape[j, s]), p)
        (z * torch.tensor(w)).sum().backward()
        assert np.max(np.abs(p.grad.numpy() - gp[j]) / np.abs(gp[j])) < 1e-9
        assert np.max(np.abs(y.grad.numpy() - g0[j]) / np.abs(g0[j]).max()) < 1e-9


def test_adjoint_finite_differences_converge():
    """Central differences on the frozen tape converge to the adjoint as eps shrinks; at the
    reference's eps = 1e-6 relative the agreement is within 1e-4 (the Rosenbrock replay's
    d/dk2 of a quantity of size 1e-12 limits the FD accuracy, not the adjoint)."""
    e = W.robertson(2, seed=4)
    yf, tape, na, nr, st = O.rb23_record(O.ROBERTSON, e.y0, e.params, 0.0, 100.0, rtol=1e-6,
                                         atol=1e-10)
    w = np.array([1.0, 10.0, 1.0])
    g0, gp = O.rb23_adjoint(O.ROBERTSON, e.y0, e.params, 0.0, tape, np.tile(w, (2, 1)))
    j = 0

    def loss(pp):
        return float(O.rb23_replay(O.ROBERTSON, e.y0[j:j + 1], pp, 0.0, tape[j:j + 1])[0] @ w)

    for q in (0, 2):
        errs = []
        for rel in (1e-4, 1e-5, 1e-6):
            eps = rel * abs(e.params[j, q])
            pp, pm = e.params[j:j + 1].copy(), e.params[j:j + 1].copy()
            pp[0, q] += eps
            pm[0, q] -= eps
            fd = (loss(pp) - loss(pm)) / (2 * eps)
            errs.append(abs(fd - gp[j, q]) / abs(gp[j, q]))
        assert min(errs) < 1e-4, errs


def _torch_rb23_saveat(y0, p, tape, ts):
    """Bracket-and-re-step (gradsolve/solvers/dense.py:106-163) with rosenbrock23_step as the
    re-step (gradsolve/warp/warp_replay.py:78-84), in torch on the Robertson field."""
    y, t = y0, 0.0
    left = [None] * len(ts)
    for s in range(len(tape)):
        h = float(tape[s])
        for q, tq in enumerate(ts):
            if h > 0 and t < tq <= t + h:
                left[q] = (y, t)
        y = _torch_rb23_step(_rob_f, _rob_jac, y, h, p)
        t = t + h
    out = []
    for q, tq in enumerate(ts):
        if tq >= t:
            out.append(y)
        elif left[q] is not None:
            yl, tl = left[q]
            out.append(_torch_rb23_step(_rob_f, _rob_jac, yl, tq - tl, p))
        else:
            out.append(y0)
    return y, torch.stack(out)


def _rob_saveat_setup(n, seed, t1=100.0):
    e = W.robertson(n, seed=seed)
    yf, tape, na, nr, st = O.rb23_record(O.ROBERTSON, e.y0, e.params, 0.0, t1, rtol=1e-6,
                                         atol=1e-10)
    assert np.all(st == 0)
    return e, yf, tape, na


def test_rb23_saveat_semantics():
    """Rosenbrock23 dense output: saves at t0 are y0, the final save is y_final bit for bit, a
    save on a mesh node equals the replayed state there (the re-step is the step itself), the
    linear invariant y1 + y2 + y3 = 1 of Robertson holds at every save (a W-method conserves
    what J's column sums conserve), and on y' = -k y the saves match exp(-k t) to the
    method's accuracy."""
    e, yf, tape, na = _rob_saveat_setup(3, 21)
    j = 1
    mesh = np.cumsum(tape[j, :na[j]])
    ts = np.sort(np.array([0.0, 0.0, 1e-4, mesh[5], 0.37, mesh[40], 12.5, 63.0, 100.0]))
    ryf, ys = O.rb23_replay_saveat(O.ROBERTSON, e.y0, e.params, 0.0, tape, ts)
    assert np.array_equal(ryf, O.rb23_replay(O.ROBERTSON, e.y0, e.params, 0.0, tape))
    assert np.array_equal(ys[:, 0], e.y0) and np.array_equal(ys[:, 1], e.y0)
    assert np.array_equal(ys[:, -1], ryf)
    for s in (5, 40):
        q = int(np.nonzero(ts == mesh[s])[0][0])
        stepped = O.rb23_replay(O.ROBERTSON, e.y0[j:j + 1], e.params[j:j + 1], 0.0,
                                tape[j:j + 1, :s + 1])
        assert np.max(np.abs(ys[j, q] - stepped[0])) < 1e-15
    assert np.max(np.abs(ys.sum(axis=2) - 1.0)) < 1e-13
    # the saves follow the stiff reference solution (scipy Radau at tight tolerance)
    for jj in range(3):
        ref = si.solve_ivp(lambda t, y, k=e.params[jj]: [
            -k[0] * y[0] + k[2] * y[1] * y[2],
            k[0] * y[0] - k[2] * y[1] * y[2] - k[1] * y[1] ** 2,
            k[1] * y[1] ** 2], (0.0, 100.0), e.y0[jj], method="Radau", t_eval=ts[2:],
            rtol=1e-12, atol=1e-14).y.T
        assert np.max(np.abs(ys[jj, 2:, 0] - ref[:, 0]) / ref[:, 0]) < 1e-3
        assert np.max(np.abs(ys[jj, 2:, 2] - ref[:, 2]) / np.maximum(ref[:, 2], 1e-3)) < 1e-3
    y0 = np.array([[1.0, 2.0], [0.5, -1.0]])
    k = np.array([[0.7], [1.3]])
    _, tape2, _, _, st2 = O.rb23_record(O.LINEAR_DECAY, y0, k, 0.0, 2.0, rtol=1e-9, atol=1e-12)
    assert np.all(st2 == 0)
    ts2 = np.linspace(0.0, 2.0, 9)
    _, ys2 = O.rb23_replay_saveat(O.LINEAR_DECAY, y0, k, 0.0, tape2, ts2)
    exact = y0[:, None, :] * np.exp(-k[:, None, :] * ts2[None, :, None])
    assert np.max(np.abs(ys2 - exact) / np.abs(exact)) < 1e-6


def test_rb23_saveat_adjoint_against_autograd_and_fd():
    """The oracle's Rosenbrock23 saveat gradient against torch autograd through the torch
    bracket-and-re-step (float64, torch.linalg.solve), and against central differences."""
    e, yf, tape, na = _rob_saveat_setup(2, 22, t1=5.0)
    ts = np.array([0.0, 1e-3, 0.2, 1.1, 2.5, 4.0, 5.0])
    ryf, ys = O.rb23_replay_saveat(O.ROBERTSON, e.y0, e.params, 0.0, tape, ts)
    rng = np.random.default_rng(23)
    ysbar = rng.standard_normal(ys.shape)
    yfbar = rng.standard_normal(ryf.shape)
    g0, gp = O.rb23_adjoint_saveat(O.ROBERTSON, e.y0, e.params, 0.0, tape, ts, ysbar, yfbar)
    for j in range(2):
        z = torch.tensor(e.y0[j], requires_grad=True)
        p = torch.tensor(e.params[j], requires_grad=True)
        tyf, tys = _torch_rb23_saveat(z, p, tape[j, :na[j]], ts)
        assert np.max(np.abs(tys.detach().numpy() - ys[j])) < 1e-13
        loss = (tys * torch.tensor(ysbar[j])).sum() + (tyf * torch.tensor(yfbar[j])).sum()
        gz, gpt = torch.autograd.grad(loss, (z, p))
        assert np.max(np.abs(gpt.numpy() - gp[j]) / np.abs(gp[j])) < 1e-9
        assert np.max(np.abs(gz.numpy() - g0[j])) < 1e-9 * np.max(np.abs(g0[j]))

    def loss_of(pp):
        yf_, ys_ = O.rb23_replay_saveat(O.ROBERTSON, e.y0[:1], pp, 0.0, tape[:1], ts)
        return np.sum(ys_[0] * ysbar[0]) + np.sum(yf_[0] * yfbar[0])

    for q in (0, 2):
        errs = []
        for rel in (1e-4, 1e-5, 1e-6):
            eps = rel * abs(e.params[0, q])
            pp, pm = e.params[:1].copy(), e.params[:1].copy()
            pp[0, q] += eps
            pm[0, q] -= eps
            errs.append(abs((loss_of(pp) - loss_of(pm)) / (2 * eps) - gp[0, q]) /
                        abs(gp[0, q]))
        assert min(errs) < 1e-5, errs
