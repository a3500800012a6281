"""GPU parity of the fused HIP Rosenbrock23 forward+record and adjoint against the oracle
(Robertson kinetics, the stiff workload: k = [0.04, 3e7, 1e4] * exp(U(log .8, log 1.25)),
y0 = [1, 0, 0], t in [0, 100], rtol 1e-6, atol 1e-10).

Tolerances: step-count mismatch < 0.1 %; final states on count-matched trajectories rel 1e-8
(the reference's own value-parity floor for its Rosenbrock kernel, which factorises W with
a different elimination, gradsolve/warp/_warp_rosenbrock.py:35-42); gradients of
sum(y_final * [1, 10, 1]) vs the oracle's adjoint on the device's own tape rel 1e-9.
"""
import numpy as np
import pytest
import torch

from gradsolve_amd import _native as N
from gradsolve_amd import hip_tsit5 as H
from gradsolve_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

W_LOSS = np.array([1.0, 10.0, 1.0])


def _soa(a, dev):
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a, np.float64).T), device=dev)


def _record(e, dev, **kw):
    rec = H.record(N.FIELD_ROBERTSON, _soa(e.y0, dev), _soa(e.params, dev), t0=0.0, t1=100.0,
                   rtol=1e-6, atol=1e-10, max_steps=50000, entry="gsx_rb23_fwd", **kw)
    torch.cuda.synchronize()
    return rec


def _adjoint(rec, dev):
    ybar = torch.as_tensor(np.tile(W_LOSS[:, None], (1, rec.n)), device=dev).contiguous()
    y0b = torch.empty((3, rec.n), dtype=torch.float64, device=dev)
    pb = torch.empty((3, rec.n), dtype=torch.float64, device=dev)
    N.call("gsx_rb23_adjoint", N.FIELD_ROBERTSON, rec.n, 3, 3, None, None, rec.p.data_ptr(),
           0.0, rec.tape.data_ptr(), rec.cap, rec.n_acc.data_ptr(), rec.ckpt.data_ptr(),
           rec.ckpt_interval, ybar.data_ptr(), y0b.data_ptr(), pb.data_ptr(), 0,
           torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return y0b.cpu().numpy().T, pb.cpu().numpy().T


@pytest.mark.parametrize("strict", [True, False])
def test_rb23_forward_matches_oracle(gpu, strict):
    e = W.robertson(2048, seed=0)
    print(f"case robertson n=2048 strict={strict}", flush=True)
    rec = _record(e, gpu, ckpt_interval=8, tape_cap=512, strict=strict)
    assert int((rec.status != 0).sum().item()) == 0
    oyf, otape, ona, onr, ost = O.rb23_record(O.ROBERTSON, e.y0, e.params, 0.0, 100.0,
                                              rtol=1e-6, atol=1e-10)
    na = rec.n_acc.cpu().numpy()
    nr = rec.n_rej.cpu().numpy()
    same = (na == ona) & (nr == onr)
    print(f"count mismatches: {(~same).sum()}", flush=True)
    assert (~same).mean() < 1e-3
    yf = rec.y_final.cpu().numpy().T
    rel = np.abs(yf[same] - oyf[same]) / np.abs(oyf[same])
    print(f"final state max rel {rel.max():.3e}", flush=True)
    assert rel.max() < 1e-8
    tape = rec.padded_tape().cpu().numpy()
    assert np.max(np.abs(tape.sum(axis=1) - 100.0)) < 1e-9


@pytest.mark.parametrize("C", [1, 2, 8, 32])
def test_rb23_adjoint_matches_oracle_on_same_tape(gpu, C):
    """C = 2 runs the pair adjoint (rb23_adj_pair_kernel), the others the segment adjoint."""
    e = W.robertson(512, seed=1)
    print(f"case robertson adjoint C={C}", flush=True)
    rec = _record(e, gpu, ckpt_interval=C, tape_cap=512)
    g0, gp = _adjoint(rec, gpu)
    tape = rec.padded_tape().cpu().numpy()
    og0, ogp = O.rb23_adjoint(O.ROBERTSON, e.y0, e.params, 0.0, tape,
                              np.tile(W_LOSS, (512, 1)))
    relp = np.abs(gp - ogp) / np.abs(ogp)
    print(f"param gradient max rel {relp.max():.3e}", flush=True)
    assert relp.max() < 1e-9
    assert np.max(np.abs(g0 - og0) / np.abs(og0).max(axis=1, keepdims=True)) < 1e-9


def test_rb23_full_size_properties(gpu):
    """N = 262,144: every trajectory reaches t1 and a 1024-trajectory sample matches the
    oracle (counts, states, gradients on the device's tape)."""
    e = W.robertson(1 << 18, seed=0)
    print("case robertson n=2^18", flush=True)
    rec = _record(e, gpu, ckpt_interval=16, tape_cap=1024)
    assert int((rec.status != 0).sum().item()) == 0
    g0, gp = _adjoint(rec, gpu)
    idx = np.random.default_rng(0).choice(e.y0.shape[0], 1024, replace=False)
    oyf, otape, ona, onr, _ = O.rb23_record(O.ROBERTSON, e.y0[idx], e.params[idx], 0.0, 100.0,
                                            rtol=1e-6, atol=1e-10)
    na = rec.n_acc.cpu().numpy()[idx]
    same = na == ona
    assert (~same).mean() < 1e-3
    yf = rec.y_final.cpu().numpy().T[idx]
    assert np.max(np.abs(yf[same] - oyf[same]) / np.abs(oyf[same])) < 1e-8
    tape = rec.padded_tape().cpu().numpy()[idx]
    og0, ogp = O.rb23_adjoint(O.ROBERTSON, e.y0[idx], e.params[idx], 0.0, tape,
                              np.tile(W_LOSS, (1024, 1)))
    assert np.max(np.abs(gp[idx] - ogp) / np.abs(ogp)) < 1e-9
