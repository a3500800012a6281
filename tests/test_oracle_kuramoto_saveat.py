"""Pins the oracle's Kuramoto field and its saveat (bracket-and-re-step) record/replay/adjoint.

* The field against the problem's own NumPy expression and its VJP against torch autograd.
* The D = 32 record against an independent NumPy restatement of tsit5_adaptive
  (gradsolve/solvers/tsit5_replay.py:62-124: np.mean for the norm, Python pow for the
  controller) with the coupling sums taken left to right (np.cumsum) and math.sin/math.cos:
  the tapes agree bit for bit, which fixes the 32-element NumPy summation order of the norm.
* saveat semantics (gradsolve/solvers/dense.py:106-163, tests/test_saveat.py:146-305 of the
  reference): times <= t0 give y0 and t1 gives y_final bit for bit; a save on a mesh point
  equals the stepped state; interior saves of y' = -k y match exp(-k t).
* The saveat adjoint against torch autograd through a torch restatement of the re-steps, and
  against central finite differences on the frozen tape.
"""
import math

import numpy as np
import torch

from gradsolve_amd import problems as P
from gradsolve_amd import workloads as W
from oracle import oracle as O

torch.set_default_dtype(torch.float64)


def _kura(n, seed, D=32):
    e = W.kuramoto32(n, seed=seed, D=D)
    return e.y0, e.params, e.aux


def test_kuramoto_field_and_vjp():
    y0, om, K = _kura(4, 0)
    prob = P.Kuramoto(K)
    for j in range(4):
        dy = O.field(O.KURAMOTO, 0.0, y0[j], om[j], aux=K)
        ref = prob.f_jax(0.0, y0[j], om[j])
        assert np.max(np.abs(dy - ref)) < 1e-13 * max(1.0, np.max(np.abs(ref)))
        v = np.random.default_rng(j).standard_normal(32)
        yb, pb = O.field_vjp(O.KURAMOTO, 0.0, y0[j], om[j], v, aux=K)
        yt = torch.tensor(y0[j], requires_grad=True)
        pt = torch.tensor(om[j], requires_grad=True)
        out = prob.f_jax(0.0, yt, pt)
        gy, gp = torch.autograd.grad(out, (yt, pt), torch.tensor(v))
        assert np.max(np.abs(yb - gy.numpy())) < 1e-13 * np.max(np.abs(gy.numpy()))
        assert np.array_equal(pb, gp.numpy())


def _numpy_tsit5_adaptive(K, y0, om, t0, t1, rtol, atol, max_steps=100000):
    """tsit5_adaptive restated in NumPy for one Kuramoto trajectory."""
    c, a, b, e = O.tableau()
    A = np.zeros((7, 7))
    A[np.tril_indices(6, -1)] = a  # rows 1..5 of the lower triangle, row-major
    B = b
    D = y0.shape[0]

    def f(y):
        s = np.array([math.sin(v) for v in y])
        co = np.array([math.cos(v) for v in y])
        sa = np.cumsum(K * s[None, :], axis=1)[:, -1]
        sb = np.cumsum(K * co[None, :], axis=1)[:, -1]
        return om + (1.0 / D) * (sa * co - sb * s)

    y = y0.copy()
    t = t0
    dt = (t1 - t0) / 100.0
    k1 = f(y)
    tape = []
    nrej = 0
    for _ in range(max_steps):
        if t >= t1:
            break
        h = min(dt, t1 - t)
        k = [k1]
        for i in range(1, 6):
            acc = A[i, 0] * k[0]
            for r in range(1, i):
                acc = acc + A[i, r] * k[r]
            k.append(f(y + h * acc))
        y5 = y + h * (B[0] * k[0] + B[1] * k[1] + B[2] * k[2] + B[3] * k[3] + B[4] * k[4] +
                      B[5] * k[5])
        k7 = f(y5)
        k.append(k7)
        err_vec = h * (e[0] * k[0] + e[1] * k[1] + e[2] * k[2] + e[3] * k[3] + e[4] * k[4] +
                       e[5] * k[5] + e[6] * k[6])
        sc = atol + rtol * np.maximum(np.abs(y), np.abs(y5))
        err = float(np.sqrt(np.mean((err_vec / sc) ** 2)))
        if err <= 1.0:
            t = t + h
            y = y5
            k1 = k7
            tape.append(h)
        else:
            nrej += 1
        fac = 0.9 * max(err, 1e-10) ** -0.2
        dt = h * min(max(fac, 0.2), 5.0)
    return y, np.array(tape), nrej


def test_record_matches_numpy_restatement_bitwise():
    y0, om, K = _kura(1, 3)
    t1 = 10.0
    yf, tape, na, nr, st = O.tsit5_record(O.KURAMOTO, y0, om, 0.0, t1, rtol=1e-8, atol=1e-8,
                                          aux=K)
    ny, ntape, nrej = _numpy_tsit5_adaptive(K, y0[0], om[0], 0.0, t1, 1e-8, 1e-8)
    assert st[0] == 0 and na[0] == len(ntape) > 20
    assert np.array_equal(tape[0, :na[0]], ntape) and nr[0] == nrej
    assert np.array_equal(yf[0], ny)


def _decay_setup():
    y0 = np.array([[1.0, 2.0], [0.5, -1.0], [3.0, 0.25]])
    k = np.array([[0.7], [1.3], [2.0]])
    yf, tape, na, nr, st = O.tsit5_record(O.LINEAR_DECAY, y0, k, 0.0, 2.0, rtol=1e-10,
                                          atol=1e-12)
    return y0, k, yf, tape, na


def test_saveat_semantics_decay():
    y0, k, yf, tape, na = _decay_setup()
    j = 1
    mesh = np.cumsum(tape[j, :na[j]])  # the accepted times of trajectory j (t0 = 0)
    ts = np.sort(np.array([0.0, 0.0, 0.3, mesh[3], 0.77, mesh[10], 1.5, 2.0]))
    ryf, ys = O.tsit5_replay_saveat(O.LINEAR_DECAY, y0, k, 0.0, tape, ts)
    assert np.array_equal(ryf, O.tsit5_replay(O.LINEAR_DECAY, y0, k, 0.0, tape))
    assert np.array_equal(ys[:, 0], y0) and np.array_equal(ys[:, 1], y0)
    assert np.array_equal(ys[:, -1], ryf)
    exact = y0[:, None, :] * np.exp(-k[:, None, :] * ts[None, :, None])
    assert np.max(np.abs(ys - exact) / np.abs(exact)) < 1e-8
    # a save on a mesh point equals the stepped state there
    for s in (3, 10):
        q = int(np.nonzero(ts == mesh[s])[0][0])
        stepped = O.tsit5_replay(O.LINEAR_DECAY, y0[j:j + 1], k[j:j + 1], 0.0,
                                 tape[j:j + 1, :s + 1])
        assert np.max(np.abs(ys[j, q] - stepped[0]) / np.abs(stepped[0])) < 1e-12


def _torch_saveat(K, y0, om, tape, ts):
    """Bracket-and-re-step in torch (dense.py:106-163) with the Kuramoto field."""
    c, a, b, e = O.tableau()
    A = np.zeros((7, 7))
    A[np.tril_indices(6, -1)] = a
    Kt = torch.tensor(K)
    D = y0.shape[0]

    def f(y):
        s, co = torch.sin(y), torch.cos(y)
        return om + (1.0 / D) * ((Kt @ s) * co - (Kt @ co) * s)

    def step(y, h):
        k = [f(y)]
        for i in range(1, 7):
            acc = A[i, 0] * k[0] if i < 6 else b[0] * k[0]
            for r in range(1, i):
                acc = acc + (A[i, r] if i < 6 else b[r]) * k[r]
            if i < 6:
                k.append(f(y + h * acc))
            else:
                return y + h * acc
        return None

    y = y0
    t = 0.0
    left = [None] * len(ts)
    for s in range(len(tape)):
        h = float(tape[s])
        for q, tq in enumerate(ts):
            if h > 0 and t < tq <= t + h:
                left[q] = (y, t)
        y = step(y, h)
        t = t + h
    out = []
    for q, tq in enumerate(ts):
        if tq >= t:
            out.append(y)
        elif left[q] is not None:
            yl, tl = left[q]
            out.append(step(yl, tq - tl))
        else:
            out.append(y0)
    return y, torch.stack(out)


def test_saveat_adjoint_against_autograd_and_fd():
    y0, om, K = _kura(2, 5)
    t1 = 1.5
    _, tape, na, _, _ = O.tsit5_record(O.KURAMOTO, y0, om, 0.0, t1, rtol=1e-8, atol=1e-8, aux=K)
    ts = np.linspace(0.0, t1, 7)
    ryf, ys = O.tsit5_replay_saveat(O.KURAMOTO, y0, om, 0.0, tape, ts, aux=K)
    rng = np.random.default_rng(7)
    target = rng.standard_normal(ys.shape)
    ysbar = 2.0 * (ys - target) / ys.size
    yfbar = rng.standard_normal(ryf.shape)
    g0, gp = O.tsit5_adjoint_saveat(O.KURAMOTO, y0, om, 0.0, tape, ts, ysbar, yfbar, aux=K)
    for j in range(2):
        z = torch.tensor(y0[j], requires_grad=True)
        p = torch.tensor(om[j], requires_grad=True)
        tyf, tys = _torch_saveat(K, z, p, tape[j, :na[j]], ts)
        assert np.max(np.abs(tys.detach().numpy() - ys[j])) < 1e-12
        loss = (tys * torch.tensor(ysbar[j])).sum() + (tyf * torch.tensor(yfbar[j])).sum()
        gz, gpt = torch.autograd.grad(loss, (z, p))
        assert np.max(np.abs(gz.numpy() - g0[j])) < 1e-11 * np.max(np.abs(g0[j]))
        assert np.max(np.abs(gpt.numpy() - gp[j])) < 1e-11 * np.max(np.abs(gp[j]))

    def loss_of(omj, j):
        o = om.copy()
        o[j] = omj
        yf_, ys_ = O.tsit5_replay_saveat(O.KURAMOTO, y0[j:j + 1], o[j:j + 1], 0.0,
                                         tape[j:j + 1], ts, aux=K)
        return np.sum(ys_ * ysbar[j]) + np.sum(yf_ * yfbar[j])

    j = 0
    fd = np.zeros(32)
    for i in range(32):
        eps = 1e-6 * max(1.0, abs(om[j, i]))
        up, dn = om[j].copy(), om[j].copy()
        up[i] += eps
        dn[i] -= eps
        fd[i] = (loss_of(up, j) - loss_of(dn, j)) / (2 * eps)
    assert np.linalg.norm(fd - gp[j]) / np.linalg.norm(gp[j]) < 1e-6
