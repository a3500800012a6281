"""GPU parity of dense output on the stiff lane: gsx_rb23_saveat and gsx_rb23_adjoint_saveat
against the oracle's rb23_replay_saveat / rb23_adjoint_saveat (bracket-and-re-step with
rosenbrock23_step, gradsolve/warp/warp_replay.py:78-84 -> gradsolve/solvers/dense.py:106-163;
the gradient is jax.grad of the same, gradsolve/warp/warp_replay.py:364-382).

  * Robertson, N = 2048 (k = [0.04, 3e7, 1e4] * exp(U(log .8, log 1.25)), t in [0, 100],
    rtol 1e-6, atol 1e-10), on the device's own recorded mesh: saves bit for bit in the strict
    variant and within 1e-12 relative in the default one; saves before t0, at a repeated time,
    at t1 and past it; y1 + y2 + y3 = 1 at every save (to 1e-12);
  * the gradient of a random save-series cotangent plus a final-state cotangent within 1e-9 of
    the oracle's (relative to each trajectory's largest component), through the segment
    adjoint at C = 1 and C = 8;
  * through the API: solve(robertson, saveat=...) routes to hip_rosenbrock23 with
    y_saved[-1] == y_final, and grad_closure(saveat=...)'s autograd gradient matches the
    oracle's on the closure's own mesh;
  * a stiff field compiled at run time: the closure's gradient against central differences of
    the same frozen-mesh closure (rel 1e-6 beyond the differences' own rounding).
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

TS = np.array([-1.0, 0.0, 1e-5, 0.01, 0.4, 3.0, 3.0, 17.0, 55.5, 99.9, 100.0, 120.0])


def _soa(a, dev):
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a, np.float64).T), device=dev)


def _record(e, dev, **kw):
    rec = H.record(N.FIELD_ROBERTSON, _soa(e.y0, dev), _soa(e.params, dev), t0=0.0, t1=100.0,
                   rtol=1e-6, atol=1e-10, max_steps=50000, entry="gsx_rb23_fwd", **kw)
    torch.cuda.synchronize()
    assert int((rec.status != 0).sum().item()) == 0
    return rec


@pytest.mark.parametrize("strict", [True, False])
def test_rb23_saveat_matches_oracle(gpu, strict):
    e = W.robertson(2048, seed=61)
    print(f"case rb23 saveat robertson n=2048 strict={strict}", flush=True)
    rec = _record(e, gpu, strict=strict, ckpt_interval=8)
    ts = torch.as_tensor(TS, device=gpu)
    sv = H.saveat(rec, ts, strict=strict, method="rb23")
    torch.cuda.synchronize()
    ys = sv.ys.permute(2, 0, 1).cpu().numpy()
    tape = rec.padded_tape().cpu().numpy()
    _, oys = O.rb23_replay_saveat(O.ROBERTSON, e.y0, e.params, 0.0, tape, TS)
    if strict:
        assert np.array_equal(ys, oys)
    else:
        np.testing.assert_allclose(ys, oys, rtol=1e-12, atol=1e-12 * np.abs(oys).max())
    yf = rec.y_final.t().cpu().numpy()
    assert np.array_equal(ys[:, -1], yf) and np.array_equal(ys[:, -2], yf)
    assert np.array_equal(ys[:, 0], e.y0) and np.array_equal(ys[:, 1], e.y0)
    assert np.array_equal(ys[:, 5], ys[:, 6])
    assert np.max(np.abs(ys.sum(axis=2) - 1.0)) < 1e-12
    step = sv.save_step.cpu().numpy()
    assert np.all(step[0] == -1) and np.all(step[-1] == rec.n_acc.cpu().numpy())


@pytest.mark.parametrize("C", [1, 2, 8])
def test_rb23_saveat_gradient_matches_oracle(gpu, C):
    e = W.robertson(1024, seed=62)
    print(f"case rb23 saveat adjoint robertson n=1024 C={C}", flush=True)
    rec = _record(e, gpu, ckpt_interval=C)
    ts = torch.as_tensor(TS, device=gpu)
    sv = H.saveat(rec, ts, method="rb23")
    rng = np.random.default_rng(63)
    ysbar = rng.standard_normal((1024, TS.shape[0], 3))
    ybar = rng.standard_normal((1024, 3))
    y0b, pb = H.adjoint_saveat(rec, ts, sv,
                               torch.as_tensor(ysbar.transpose(1, 2, 0).copy(), device=gpu),
                               torch.as_tensor(ybar.T.copy(), device=gpu), method="rb23")
    torch.cuda.synchronize()
    tape = rec.padded_tape().cpu().numpy()
    og0, ogp = O.rb23_adjoint_saveat(O.ROBERTSON, e.y0, e.params, 0.0, tape, TS, ysbar, ybar)
    g0, gp = y0b.cpu().numpy().T, pb.cpu().numpy().T
    assert np.max(np.abs(gp - ogp) / np.abs(ogp).max(axis=1, keepdims=True)) < 1e-9
    assert np.max(np.abs(g0 - og0) / np.abs(og0).max(axis=1, keepdims=True)) < 1e-9


def test_rb23_saveat_through_the_api(gpu):
    prob = P.Robertson()
    e = W.robertson(256, seed=64)
    ts = np.linspace(0.0, 100.0, 21)
    print("case rb23 saveat api robertson n=256", flush=True)
    res = gs.solve(prob, e.y0, e.params, saveat=ts, rtol=1e-6, atol=1e-10)
    assert res.y_saved.shape == (256, 21, 3) and res.route.actual == "hip_rosenbrock23"
    np.testing.assert_array_equal(res.y_saved[:, -1], res.y_final)
    np.testing.assert_array_equal(res.y_saved[:, 0], e.y0)
    clo = gs.grad_closure(prob, e.y0, e.params, saveat=ts, rtol=1e-6, atol=1e-10)
    pt = torch.tensor(e.params, dtype=torch.float64, device=gpu, requires_grad=True)
    ys = clo(pt)
    assert ys.shape == (256, 21, 3)
    wts = torch.as_tensor(np.tile([1.0, 1e4, 1.0], (256, 21, 1)), device=gpu)
    (ys * wts).sum().backward()
    tape = clo.dts_padded()
    og0, ogp = O.rb23_adjoint_saveat(O.ROBERTSON, e.y0, e.params, 0.0, tape, ts,
                                     wts.cpu().numpy())
    g = pt.grad.cpu().numpy()
    assert np.max(np.abs(g - ogp) / np.abs(ogp).max(axis=1, keepdims=True)) < 1e-9


@dataclasses.dataclass
class Prob:
    name: str
    dim: int
    is_stiff: bool
    f_jax: object
    t0: float = 0.0
    t1: float = 3.0


def stiff_pair(t, y, p):
    # a fast linear mode coupled to a slow nonlinear one (stiffness ratio ~ p[0])
    return np.array([-p[0] * (y[0] - np.cos(y[1])), -p[1] * y[1] + 0.1 * y[0] * y[0]])


def test_rb23_saveat_on_a_runtime_field_matches_finite_differences(gpu):
    register_field("sv_stiff_pair", stiff_pair, 2, 2, stiff=True)
    prob = Prob("sv_stiff_pair", 2, True, stiff_pair)
    rng = np.random.default_rng(65)
    n = 32
    y0 = np.column_stack([rng.uniform(-1.0, 1.0, n), rng.uniform(-0.5, 0.5, n)])
    params = np.column_stack([rng.uniform(500.0, 2000.0, n), rng.uniform(0.5, 1.5, n)])
    ts = np.array([0.0, 0.05, 0.8, 1.9, 3.0])
    print("case rb23 saveat runtime stiff_pair n=32", flush=True)
    clo = gs.grad_closure(prob, y0, params, saveat=ts, rtol=1e-7, atol=1e-9)
    assert clo.route.actual == "hip_rosenbrock23"
    w = rng.standard_normal((n, ts.shape[0], 2))
    pt = torch.tensor(params, dtype=torch.float64, device=gpu, requires_grad=True)
    (clo(pt) * torch.as_tensor(w, device=gpu)).sum().backward()
    g = pt.grad.cpu().numpy()
    g_fd = np.zeros_like(params)
    for c in range(2):
        eps = 1e-6 * np.abs(params[:, c])
        pp, pm = params.copy(), params.copy()
        pp[:, c] += eps
        pm[:, c] -= eps
        g_fd[:, c] = ((clo(pp) - clo(pm)) * w).sum(axis=(1, 2)) / (2 * eps)
    # Central differences carry ~ eps_mach * |loss| / eps of rounding (|loss| ~ 10, eps = 1e-6
    # relative of a rate ~ 1e3): ~ 1e-12 absolute, which matters for the stiff rate, whose
    # gradient is only ~ 1e-7 because its fast mode equilibrates at once.
    loss = np.abs((clo(params) * w).sum(axis=(1, 2)))[:, None]
    fd_noise = 2.2e-16 * loss / (2e-6 * params)
    err = np.abs(g - g_fd) - 4.0 * fd_noise
    for c in range(2):
        assert np.max(err[:, c]) / np.abs(g_fd[:, c]).max() < 1e-6, c
