"""The wide-state Tsit5 lane (gsx_tsit5_wide_*): fields compiled at run time with the state,
stages and stage inputs in an HBM workspace, the native engine of the reference's high-NVAR
cells (gradsolve/dispatch.py:256-258; record_jax.py:171-213, tsit5_replay.py:127-232).

* A generated Lorenz (D = 3) through the wide lane against the oracle (the register lanes'
  arithmetic): step-count mismatches < 0.1 %, final states 1e-10, the wide replay of the
  wide tape within 1e-13 of the forward, gradients against the oracle's adjoint on the same tape
  1e-10 (checkpoint intervals 1, 5 and 8).
* A 160-wide Allen-Cahn lattice (a * discrete Laplacian + y - y^3 + b; 160 > 128 takes the
  split branch of NumPy's pairwise WRMS sum) against the Python oracle (oracle/generic.py,
  NumPy's own np.sum and complex-step VJPs): counts, final states 1e-10, gradients on the
  device tape 1e-9, and central finite differences on the frozen tape.
* Through the API: solve / grad_closure route the 160-wide problem to hip_tsit5_wide with the
  reason, and the closure's gradient equals the oracle's.
"""
import numpy as np
import pytest
import torch

import gradsolve_amd as gs
from gradsolve_amd import hip_tsit5 as H
from gradsolve_amd import hip_wide as HW
from gradsolve_amd import workloads as W
from gradsolve_amd.fields import register_field
from oracle import generic as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def lorenz(t, y, p):
    return np.stack([10.0 * (y[1] - y[0]), p[0] * y[0] - y[1] - y[0] * y[2],
                     y[0] * y[1] - (8.0 / 3.0) * y[2]])


def allen_cahn(t, y, p):
    lap = np.roll(y, -1) - 2.0 * y + np.roll(y, 1)
    return p[0] * lap + y - y * y * y + p[1]


class AllenCahn160:
    name = "allen_cahn_160"
    dim = 160
    t0 = 0.0
    t1 = 3.0
    is_stiff = False

    def f_jax(self, t, y, p):
        return allen_cahn(t, y, p)


def _ac_inputs(n, seed):
    rng = np.random.default_rng(seed)
    p = np.stack([rng.uniform(0.3, 0.6, n), rng.uniform(-0.1, 0.1, n)], axis=1)
    y0 = rng.uniform(-1.0, 1.0, (n, 160))
    return y0, p


def _soa(a, dev):
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a, np.float64).T), device=dev)


def _rel_rows(a, b):
    return np.abs(a - b) / np.abs(b).max(axis=1, keepdims=True)


def _adjoint(lane, rec, ybar, C, dev):
    """Replay (fresh checkpoints at interval C) then the wide adjoint."""
    rec.ckpt_interval = C
    rec.ckpt = H._alloc_ckpt(rec.field, rec.D, rec.n, rec.cap, C, dev)
    yf, ck = lane.replay(rec, rec.y0, rec.p)
    g0, gp = lane.adjoint(rec, ck, rec.p, ybar)
    torch.cuda.synchronize()
    return yf, g0, gp


def test_wide_lane_generated_lorenz_against_oracle(gpu):
    b = register_field("wide_lorenz_3", lorenz, 3, 1)
    e = W.lorenz(2048, seed=31)
    print("case wide lane lorenz n=2048", flush=True)
    rec = HW.record(b.field_id, _soa(e.y0, gpu), _soa(e.params, gpu), t0=0.0, t1=1.0,
                    rtol=1e-8, atol=1e-8, max_steps=100000, ckpt_interval=8)
    torch.cuda.synchronize()
    assert int((rec.status != 0).sum().item()) == 0
    oyf, _, ona, onr, _ = O.tsit5_record(O.LORENZ, e.y0, e.params, 0.0, 1.0, rtol=1e-8,
                                         atol=1e-8)
    na = rec.n_acc.cpu().numpy()
    same = na == ona
    print(f"count mismatches {(~same).sum()}", flush=True)
    assert (~same).mean() < 1e-3
    yf = rec.y_final.cpu().numpy().T
    assert np.max(_rel_rows(yf[same], oyf[same])) < 1e-10
    tape = rec.padded_tape().cpu().numpy()
    ybar = (2.0 * rec.y_final).contiguous()
    og0, ogp = O.tsit5_adjoint(O.LORENZ, e.y0, e.params, 0.0, tape, 2.0 * yf)
    lane = HW._WideLane(b, None)
    for C in (1, 5, 8):
        yr, g0, gp = _adjoint(lane, rec, ybar, C, gpu)
        # the replay re-forms each step (contracted products may round differently)
        assert torch.max(torch.abs(yr - rec.y_final) / torch.abs(rec.y_final)).item() < 1e-13
        er = np.max(np.abs(gp.cpu().numpy()[0] - ogp[:, 0]) / np.abs(ogp[:, 0]))
        ey = np.max(_rel_rows(g0.cpu().numpy().T, og0))
        print(f"C={C}: gradient vs oracle on the same tape: rho {er:.2e}, y0 {ey:.2e}",
              flush=True)
        assert er < 1e-10 and ey < 1e-10


def test_wide_lane_lattice_160_against_python_oracle(gpu):
    b = register_field("wide_allen_cahn_160", allen_cahn, 160, 2)
    n = 64
    y0, p = _ac_inputs(n, 32)
    print(f"case wide lane allen-cahn D=160 n={n}", flush=True)
    rec = HW.record(b.field_id, _soa(y0, gpu), _soa(p, gpu), t0=0.0, t1=3.0, rtol=1e-8,
                    atol=1e-8, max_steps=100000, ckpt_interval=6)
    ybar = (2.0 * rec.y_final).contiguous()
    lane = HW._WideLane(b, None)
    g0, gp = lane.adjoint(rec, rec.ckpt, rec.p, ybar)
    torch.cuda.synchronize()
    assert int((rec.status != 0).sum().item()) == 0
    oyf, _, ona, onr, _ = G.tsit5_record(allen_cahn, y0, p, 0.0, 3.0, rtol=1e-8, atol=1e-8)
    na = rec.n_acc.cpu().numpy()
    same = (na == ona) & (rec.n_rej.cpu().numpy() == onr)
    yf = rec.y_final.cpu().numpy().T
    es = np.max(_rel_rows(yf[same], oyf[same]))
    print(f"count mismatches {(~same).sum()} of {n}; mean accepted {na.mean():.1f}; "
          f"final states {es:.2e}", flush=True)
    assert same.mean() > 0.95 and es < 1e-10
    tape = rec.padded_tape().cpu().numpy()
    k = 6
    og0, ogp = G.tsit5_adjoint(allen_cahn, y0[:k], p[:k], 0.0, tape[:k], 2.0 * yf[:k])
    eg = np.max(_rel_rows(gp.cpu().numpy().T[:k], ogp))
    ey = np.max(_rel_rows(g0.cpu().numpy().T[:k], og0))
    print(f"gradients vs the Python oracle on the device tape: p {eg:.2e}, y0 {ey:.2e}",
          flush=True)
    assert eg < 1e-9 and ey < 1e-9
    # central differences in p on the frozen tape (one trajectory)
    j = 0
    tp = tape[j:j + 1]

    def loss(pp):
        return float(np.sum(G.tsit5_replay(allen_cahn, y0[j:j + 1], pp[None, :], 0.0, tp) ** 2))

    g = gp.cpu().numpy()[:, j]
    for q in range(2):
        eps = 1e-6 * max(1.0, abs(p[j, q]))
        pp, pm = p[j].copy(), p[j].copy()
        pp[q] += eps
        pm[q] -= eps
        fd = (loss(pp) - loss(pm)) / (2 * eps)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(g[q])), (q, fd, g[q])


def test_wide_lane_through_the_api(gpu):
    prob = AllenCahn160()
    y0, p = _ac_inputs(16, 33)
    print("case api allen-cahn D=160 n=16", flush=True)
    res = gs.solve(prob, y0, p, rtol=1e-8, atol=1e-8, device="cuda")
    assert res.route.actual == "hip_tsit5_wide"
    assert "diffrax-not-installed" in res.route.reason
    cl = gs.grad_closure(prob, y0, p, wrt=("y0", "params"), rtol=1e-8, atol=1e-8,
                         device="cuda")
    assert cl.route.actual == "hip_tsit5_wide" and "alias:tsit5_replay" in cl.route.reason
    z = torch.tensor(y0, device="cuda", requires_grad=True)
    pt = torch.tensor(p, device="cuda", requires_grad=True)
    yf = cl(z, pt)
    (yf * yf).sum().backward()
    assert np.array_equal(yf.detach().cpu().numpy(), res.y_final) or \
        np.max(np.abs(yf.detach().cpu().numpy() - res.y_final)) < 1e-12
    tape = cl.dts_padded()
    oyf = G.tsit5_replay(allen_cahn, y0[:3], p[:3], 0.0, tape[:3])
    og0, ogp = G.tsit5_adjoint(allen_cahn, y0[:3], p[:3], 0.0, tape[:3], 2.0 * oyf)
    assert np.max(_rel_rows(pt.grad.cpu().numpy()[:3], ogp)) < 1e-9
    assert np.max(_rel_rows(z.grad.cpu().numpy()[:3], og0)) < 1e-9
