"""The measured configurations at full size, through the exact launches the benchmarks time.

Lorenz, N = 2^20 (rho ~ U(20, 30), y0 = [1, 0, 0] + 0.1 N(0, 1), t in [0, 1],
rtol = atol = 1e-8): bench.py's path -- a pilot record sizes the tape (longest trajectory + 16),
then a preallocated ``Workspace`` at checkpoint interval 2 (the pair adjoint
``tsit5_adj_pair_kernel``), default (FMA) arithmetic, cotangent 2 y_final. Checked:
* every trajectory reaches t1 and its tape sums to t1 - t0;
* a 2048-trajectory sample: gradients w.r.t. rho and y0 against the oracle's adjoint on the
  device tape (1e-10 relative); the oracle's own record (count mismatches < 0.1 %, final states
  1e-10 relative on count-matched trajectories) and the gradients on the oracle's own tape;
* central finite differences (eps = 1e-6 max(1, |x|)) on the frozen device tape for 64
  trajectories drawn by ``rng.choice(2**20, 64)`` from the same ensemble (1e-6 relative).

Uneven lengths, N = 2^22 (rho ~ U(0, 60), t1 ~ U(0.5, 10), y0 noise 0.5, rtol = atol = 1e-6):
the persistent work-queue engine at the setting the autotuner picks at this size (four waves
per SIMD, checkpoint interval 4, longest-first queue): every status 0, every tape sums to its
own t1, and a 2048-trajectory sample's gradients against the oracle's adjoint on the device
tape.
"""
import numpy as np
import pytest
import torch

from gradsolve_amd import _native as N
from gradsolve_amd import hip_tsit5 as H
from gradsolve_amd import persist as PS
from gradsolve_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rel_rows(a, b):
    """Per-trajectory relative error, normalised by the reference row's largest entry."""
    return np.abs(a - b) / np.abs(b).max(axis=1, keepdims=True)


def test_headline_path_full_size(gpu):
    ens = W.lorenz_1m(seed=0)
    n = ens.y0.shape[0]
    print(f"case headline lorenz n={n} C=2 fma", flush=True)
    y0 = torch.as_tensor(np.ascontiguousarray(ens.y0.T), device=gpu)
    p = torch.as_tensor(np.ascontiguousarray(ens.params.T), device=gpu)
    kw = dict(t0=0.0, t1=1.0, rtol=1e-8, atol=1e-8, max_steps=100000)
    rec = H.record(N.FIELD_LORENZ, y0, p, tape_cap=256, ckpt_interval=8, **kw)
    torch.cuda.synchronize()
    cap = int(rec.max_steps_taken() + 16)
    del rec
    ws = H.Workspace(N.FIELD_LORENZ, 3, 1, n, cap, 2, gpu)
    ws.forward(y0, p, **kw)
    torch.mul(ws.y_final, 2.0, out=ws.ybar)
    g0, gp = ws.adjoint(p, ws.ybar, t0=0.0)
    torch.cuda.synchronize()
    assert int((ws.status != 0).sum().item()) == 0
    na_all = ws.n_acc.to(torch.int64)
    rows = torch.arange(cap, device=gpu)[:, None]
    hsum = torch.where(rows < na_all[None, :], ws.tape, 0.0).sum(dim=0)
    assert torch.max(torch.abs(hsum - 1.0)).item() < 1e-12

    idx = np.random.default_rng(0).choice(n, 2048, replace=False)
    it = torch.as_tensor(idx, device=gpu)
    na = na_all[it].cpu().numpy()
    S = int(na.max())
    tp = ws.tape[:S, it].t().cpu().numpy()
    tp[np.arange(S)[None, :] >= na[:, None]] = 0.0
    yf = ws.y_final[:, it].t().cpu().numpy()
    dr = gp[0, it].cpu().numpy()
    dy = g0[:, it].t().cpu().numpy()
    oy0b, opb = O.tsit5_adjoint(O.LORENZ, ens.y0[idx], ens.params[idx], 0.0, tp, 2.0 * yf)
    e_r = np.abs(dr - opb[:, 0]) / np.abs(opb[:, 0])
    e_y = _rel_rows(dy, oy0b)
    print(f"same tape: d rho max {e_r.max():.2e}, d y0 max {e_y.max():.2e}", flush=True)
    assert e_r.max() < 1e-10 and e_y.max() < 1e-10

    oyf, otape, ona, onr, ost = O.tsit5_record(O.LORENZ, ens.y0[idx], ens.params[idx], 0.0,
                                               1.0, rtol=1e-8, atol=1e-8)
    same = na == ona
    print(f"count mismatches vs the oracle's record: {(~same).sum()} of {same.size}",
          flush=True)
    assert (~same).mean() < 1e-3
    e_s = _rel_rows(yf[same], oyf[same])
    print(f"final states vs the oracle's record: max {e_s.max():.2e}", flush=True)
    assert e_s.max() < 1e-10
    # gradients on the oracle's own (free-running) tape: the two meshes differ in the
    # controller's rounding, so this bounds the whole path, mesh included
    og0, ogp = O.tsit5_adjoint(O.LORENZ, ens.y0[idx], ens.params[idx], 0.0, otape,
                               2.0 * oyf)
    f_r = (np.abs(dr - ogp[:, 0]) / np.abs(ogp[:, 0]))[same]
    f_y = _rel_rows(dy, og0)[same]
    print(f"own tapes: d rho max {f_r.max():.2e} median {np.median(f_r):.2e}; d y0 max "
          f"{f_y.max():.2e} median {np.median(f_y):.2e}", flush=True)
    assert f_r.max() < 1e-8 and f_y.max() < 1e-8

    fd_idx = np.random.default_rng(1).choice(n, 64, replace=False)
    ft = torch.as_tensor(fd_idx, device=gpu)
    fna = na_all[ft].cpu().numpy()
    FS = int(fna.max())
    ftp = ws.tape[:FS, ft].t().cpu().numpy()
    ftp[np.arange(FS)[None, :] >= fna[:, None]] = 0.0
    ys, ps = ens.y0[fd_idx], ens.params[fd_idx]
    g_r = gp[0, ft].cpu().numpy()
    g_y = g0[:, ft].t().cpu().numpy()

    def loss(y, q):
        return np.sum(O.tsit5_replay(O.LORENZ, y, q, 0.0, ftp) ** 2, axis=1)

    eps = 1e-6 * np.maximum(1.0, np.abs(ps[:, 0]))
    fd = (loss(ys, ps + eps[:, None]) - loss(ys, ps - eps[:, None])) / (2 * eps)
    fd_err = np.abs(fd - g_r) / np.abs(g_r)
    for d in range(3):
        e = 1e-6 * np.maximum(1.0, np.abs(ys[:, d]))
        yp, ym = ys.copy(), ys.copy()
        yp[:, d] += e
        ym[:, d] -= e
        fdd = (loss(yp, ps) - loss(ym, ps)) / (2 * e)
        fd_err = np.maximum(fd_err, np.abs(fdd - g_y[:, d]) /
                            np.maximum(np.abs(g_y).max(axis=1), 1.0))
    print(f"frozen-tape central FD on 64 trajectories: max {fd_err.max():.2e}", flush=True)
    assert fd_err.max() < 1e-6


def test_uneven_persistent_full_size(gpu):
    e = W.lorenz_uneven(1 << 22, seed=0)
    n = e.y0.shape[0]
    print(f"case uneven persistent n={n} waves=4 C=4", flush=True)
    y0 = torch.as_tensor(np.ascontiguousarray(e.y0.T), device=gpu)
    p = torch.as_tensor(np.ascontiguousarray(e.params.T), device=gpu)
    t1 = torch.as_tensor(e.t1_per, device=gpu)
    kw = dict(t0=0.0, t1=float(e.t1_per.max()), rtol=1e-6, atol=1e-6, max_steps=100000,
              t1_per=t1)
    est = PS.estimate_steps(N.FIELD_LORENZ, y0, p, t0=0.0, t1=kw["t1"], t1_per=t1, rtol=1e-6,
                            atol=1e-6)
    eng = PS.PersistentTsit5(N.FIELD_LORENZ, 3, 1, n, est=est, waves_per_simd=4,
                             ckpt_interval=4, device=gpu)
    eng.plan_longest_first(t1)
    eng.run_forward(y0, p, **kw)
    torch.mul(eng.y_final, 2.0, out=eng.ybar)
    g0, gp = eng.adjoint(p, eng.ybar, t0=0.0)
    torch.cuda.synchronize()
    assert int((eng.status != 0).sum().item()) == 0
    na = eng.n_acc
    spread = na.max().item() / max(1, na.min().item())
    print(f"accepted steps {na.min().item()}..{na.max().item()} ({spread:.0f}x)", flush=True)
    assert spread > 20
    worst = 0.0
    for lo in range(0, n, 1 << 16):
        rows = torch.arange(lo, min(n, lo + (1 << 16)), device=gpu)
        hs = eng.padded_tape(rows).sum(dim=1)
        worst = max(worst, torch.max(torch.abs(hs - t1[rows]) / t1[rows]).item())
    print(f"tape sums vs t1: max rel {worst:.2e}", flush=True)
    assert worst < 1e-12

    idx = np.random.default_rng(0).choice(n, 2048, replace=False)
    it = torch.as_tensor(idx, device=gpu)
    tp = eng.padded_tape(it).cpu().numpy()
    yf = eng.y_final[:, it].t().cpu().numpy()
    oy0b, opb = O.tsit5_adjoint(O.LORENZ, e.y0[idx], e.params[idx], 0.0, tp, 2.0 * yf)
    oyf = O.tsit5_replay(O.LORENZ, e.y0[idx], e.params[idx], 0.0, tp)
    s_err = _rel_rows(yf, oyf)
    e_r = np.abs(gp[0, it].cpu().numpy() - opb[:, 0]) / np.abs(opb[:, 0])
    e_y = _rel_rows(g0[:, it].t().cpu().numpy(), oy0b)
    print(f"replay of the device tape: state max {s_err.max():.2e}; gradients d rho median "
          f"{np.median(e_r):.2e} max {e_r.max():.2e}, d y0 median {np.median(e_y):.2e} max "
          f"{e_y.max():.2e}", flush=True)
    assert np.median(e_r) < 1e-12 and np.median(e_y) < 1e-12
    assert np.percentile(e_r, 99) < 1e-8 and np.percentile(e_y, 99) < 1e-8
