"""GPU parity of dense output (saveat) on the Tsit5 lanes: gsx_tsit5_saveat and
gsx_tsit5_adjoint_saveat against the oracle's replay_saveat / adjoint_saveat
(gradsolve/solvers/dense.py:106-163, 318-335).

  * Lorenz, N = 2048, on the device's own recorded mesh: the saved states bit for bit in the
    strict variant and within 1e-12 relative in the default one (the saves come from the same
    replay arithmetic as the final state); saves before t0, at t1 and repeated times included;
  * the gradient of a random save-series cotangent plus a final-state cotangent within 1e-10
    of the oracle's on the same mesh, through the pair (C = 2) and segment (C = 8) adjoints;
  * through the API: solve(saveat=...) and grad_closure(saveat=...) on Lorenz, and on a field
    compiled at run time (a time-dependent forced pendulum): the closure's autograd gradient
    against central differences of the same frozen-mesh closure (rel 1e-6);
  * y' = -k y: interior saves against exp(-k t) (rel 2e-7 at rtol = atol = 1e-8).
"""
import dataclasses

import numpy as np
import pytest
import torch

import gradsolve_amd as gs
from gradsolve_amd import _native as N
from gradsolve_amd import hip_tsit5 as H
from gradsolve_amd import problems as P
from gradsolve_amd import workloads as W
from gradsolve_amd.fields import register_field
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TS = np.array([-0.25, 0.0, 1e-3, 0.1, 0.3, 0.3, 0.5, 0.61, 0.9, 0.999, 1.0, 1.2])


def _soa(a, dev):
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a, np.float64).T), device=dev)


def _record(e, dev, **kw):
    rec = H.record(N.FIELD_LORENZ, _soa(e.y0, dev), _soa(e.params, dev), t0=e.t0, t1=e.t1,
                   rtol=e.rtol, atol=e.atol, max_steps=100000, **kw)
    torch.cuda.synchronize()
    assert np.all(rec.status.cpu().numpy() == 0)
    return rec


@pytest.mark.parametrize("strict", [True, False])
def test_saveat_matches_oracle(gpu, strict):
    e = W.lorenz(2048, seed=41, noise=0.05)
    print(f"case saveat lorenz n=2048 strict={strict}", flush=True)
    rec = _record(e, gpu, strict=strict, ckpt_interval=8)
    ts = torch.as_tensor(TS, device=gpu)
    sv = H.saveat(rec, ts, strict=strict)
    torch.cuda.synchronize()
    ys = sv.ys.permute(2, 0, 1).cpu().numpy()
    tape = rec.padded_tape().cpu().numpy()
    oyf, oys = O.tsit5_replay_saveat(O.LORENZ, e.y0, e.params, 0.0, tape, TS)
    if strict:
        assert np.array_equal(ys, oys)
    else:
        np.testing.assert_allclose(ys, oys, rtol=1e-12, atol=1e-12 * np.abs(oys).max())
    yf = rec.y_final.t().cpu().numpy()
    assert np.array_equal(ys[:, -1], yf) and np.array_equal(ys[:, -2], yf)
    assert np.array_equal(ys[:, 0], e.y0) and np.array_equal(ys[:, 1], e.y0)
    assert np.array_equal(ys[:, 4], ys[:, 5])
    step = sv.save_step.cpu().numpy()
    assert np.all(step[0] == -1) and np.all(step[-1] == rec.n_acc.cpu().numpy())


@pytest.mark.parametrize("C", [2, 8])
def test_saveat_gradient_matches_oracle(gpu, C):
    e = W.lorenz(1024, seed=42, noise=0.05)
    print(f"case saveat adjoint lorenz n=1024 C={C}", flush=True)
    rec = _record(e, gpu, ckpt_interval=C)
    ts = torch.as_tensor(TS, device=gpu)
    sv = H.saveat(rec, ts)
    rng = np.random.default_rng(43)
    ysbar = rng.standard_normal((1024, TS.shape[0], 3))
    ybar = rng.standard_normal((1024, 3))
    y0b, pb = H.adjoint_saveat(rec, ts, sv,
                               torch.as_tensor(ysbar.transpose(1, 2, 0).copy(), device=gpu),
                               torch.as_tensor(ybar.T.copy(), device=gpu))
    torch.cuda.synchronize()
    tape = rec.padded_tape().cpu().numpy()
    og0, ogp = O.tsit5_adjoint_saveat(O.LORENZ, e.y0, e.params, 0.0, tape, TS, ysbar, ybar)
    g0, gp = y0b.cpu().numpy().T, pb.cpu().numpy().T
    assert np.max(np.abs(gp - ogp) / np.abs(ogp).max(axis=1, keepdims=True)) < 1e-10
    assert np.max(np.abs(g0 - og0) / np.abs(og0).max(axis=1, keepdims=True)) < 1e-10


def test_saveat_through_the_api(gpu):
    prob = P.DiffEqGPULorenz()
    e = W.lorenz(256, seed=44, noise=0.05)
    ts = np.linspace(0.0, 1.0, 11)
    print("case saveat api lorenz n=256", flush=True)
    res = gs.solve(prob, e.y0, e.params, saveat=ts, rtol=1e-8, atol=1e-8)
    assert res.y_saved.shape == (256, 11, 3) and res.route.actual == "hip_replay"
    np.testing.assert_array_equal(res.y_saved[:, -1], res.y_final)
    clo = gs.grad_closure(prob, e.y0, e.params, saveat=ts, rtol=1e-8, atol=1e-8)
    pt = torch.tensor(e.params, dtype=torch.float64, device=gpu, requires_grad=True)
    ys = clo(pt)
    assert ys.shape == (256, 11, 3)
    (ys ** 2).sum().backward()
    tape = clo.dts_padded()
    og0, ogp = O.tsit5_adjoint_saveat(O.LORENZ, e.y0, e.params, 0.0, tape, ts,
                                      2.0 * ys.detach().cpu().numpy())
    g = pt.grad.cpu().numpy()
    assert np.max(np.abs(g - ogp) / np.abs(ogp).max(axis=1, keepdims=True)) < 1e-10


@dataclasses.dataclass
class Prob:
    name: str
    dim: int
    is_stiff: bool
    f_jax: object
    t0: float = 0.0
    t1: float = 2.0


def pendulum(t, y, p):
    return np.array([y[1], -p[0] * np.sin(y[0]) - p[1] * y[1] + 0.5 * np.cos(t)])


def test_saveat_on_a_runtime_field_matches_finite_differences(gpu):
    register_field("sv_pendulum", pendulum, 2, 2)
    prob = Prob("sv_pendulum", 2, False, pendulum)
    rng = np.random.default_rng(45)
    n = 32
    y0 = np.column_stack([rng.uniform(-1.0, 1.0, n), rng.uniform(-0.5, 0.5, n)])
    params = np.column_stack([rng.uniform(0.5, 2.0, n), rng.uniform(0.05, 0.3, n)])
    ts = np.array([0.0, 0.37, 0.8, 1.5, 2.0])
    print("case saveat runtime pendulum n=32", flush=True)
    clo = gs.grad_closure(prob, y0, params, saveat=ts, rtol=1e-8, atol=1e-8)
    w = rng.standard_normal((n, ts.shape[0], 2))
    pt = torch.tensor(params, dtype=torch.float64, device=gpu, requires_grad=True)
    (clo(pt) * torch.as_tensor(w, device=gpu)).sum().backward()
    g = pt.grad.cpu().numpy()
    g_fd = np.zeros_like(params)
    for c in range(2):
        eps = 1e-6
        pp, pm = params.copy(), params.copy()
        pp[:, c] += eps
        pm[:, c] -= eps
        g_fd[:, c] = ((clo(pp) - clo(pm)) * w).sum(axis=(1, 2)) / (2 * eps)
    assert np.max(np.abs(g - g_fd) / np.abs(g_fd).max()) < 1e-6


def test_saveat_linear_decay_closed_form(gpu):
    prob = P.LinearDecay(2, t1=2.0)
    rng = np.random.default_rng(46)
    n = 64
    y0 = rng.uniform(0.5, 2.0, (n, 2))
    k = rng.uniform(0.3, 1.5, (n, 1))
    ts = np.linspace(0.0, 2.0, 17)
    res = gs.solve(prob, y0, k, saveat=ts, rtol=1e-8, atol=1e-8)
    exact = y0[:, None, :] * np.exp(-k[:, None, :] * ts[None, :, None])
    np.testing.assert_allclose(res.y_saved, exact, rtol=2e-7)
