"""The FP32 Rosenbrock23 lane (FP32 field and stage vectors on a compensated state; FP64 W, LU,
solves, error estimate, norm, controller, time and tape), with the reference's f32 acceptance
(gradsolve tests/test_f32_replay.py:13-24, 88-94, 154-156):

* the reference's stiff cell -- Robertson k = [0.04, 3e7, 1e4], y0 = [1, 0, 0], n = 256,
  t in [0, 1e4], rtol 1e-5, atol 1e-8, loss sum(y_final^2): the gradient against the FP64
  oracle's adjoint on the same FP32-recorded tape, rel-L2 <= 1e-3 and cosine >= 1 - 1e-6; the
  primal against a tight reference (the oracle's Rodas5P at rtol 1e-12) <= 1e-3 of max|y|;
* the stiff workload (k = [0.04, 3e7, 1e4] * exp(U(log .8, log 1.25)), t in [0, 100], rtol 1e-6,
  atol 1e-10, n = 2048, loss sum(y_final * [1, 10, 1])): the same gradient bounds, and the
  final-state error against the FP64 oracle's own record stated beside the FP64 lane's
  discretization error.
"""
import numpy as np
import pytest
import torch

import gradsolve_amd as gs
from gradsolve_amd import _native as N
from gradsolve_amd import hip_tsit5 as H
from gradsolve_amd import problems as P
from gradsolve_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _soa(a, dev):
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a, np.float64).T), device=dev)


def _rel_cos(g, ref):
    g, ref = np.ravel(g), np.ravel(ref)
    rel = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    cos = g @ ref / (np.linalg.norm(g) * np.linalg.norm(ref))
    return rel, cos


def test_reference_f32_cell(gpu):
    n = 256
    y0 = np.tile(np.array([1.0, 0.0, 0.0]), (n, 1))
    k = np.tile(np.array([0.04, 3.0e7, 1.0e4]), (n, 1))
    prob = P.Robertson(t0=0.0, t1=1.0e4)
    print("case robertson f32 cell n=256 t1=1e4", flush=True)
    cl = gs.grad_closure(prob, y0, k, wrt="params", engine="hip_rosenbrock23",
                         precision="float32", rtol=1e-5, atol=1e-8, device="cuda")
    assert cl.route.actual == "hip_rosenbrock23"
    p = torch.tensor(k, device="cuda", requires_grad=True)
    yf = cl(p)
    (yf * yf).sum().backward()
    g32 = p.grad.cpu().numpy()
    tape = cl.dts_padded()
    assert np.allclose(tape.sum(axis=1), 1.0e4, rtol=0, atol=1e-8)
    oyf = O.rb23_replay(O.ROBERTSON, y0, k, 0.0, tape)
    _, g64 = O.rb23_adjoint(O.ROBERTSON, y0, k, 0.0, tape, 2.0 * oyf)
    rel, cos = _rel_cos(g32, g64)
    tyf, _, _, _, tst = O.rodas5p_record(O.ROBERTSON, y0[:1], k[:1], 0.0, 1.0e4, rtol=1e-12,
                                         atol=1e-16)
    assert tst[0] == 0
    y32 = yf.detach().cpu().numpy()
    prim = np.max(np.abs(y32 - tyf[0]) / np.abs(tyf[0]).max())
    res64 = gs.solve(prob, y0[:1], k[:1], engine="hip_rosenbrock23", rtol=1e-5, atol=1e-8,
                     device="cuda")
    prim64 = np.max(np.abs(res64.y_final[0] - tyf[0]) / np.abs(tyf[0]).max())
    print(f"FP32 gradient vs FP64 adjoint on the FP32 tape: rel-L2 {rel:.2e}, cosine "
          f"1-{1 - cos:.1e}; primal vs tight reference {prim:.2e} (FP64 lane {prim64:.2e})",
          flush=True)
    assert rel <= 1e-3 and cos >= 1 - 1e-6 and prim <= 1e-3
    res = gs.solve(prob, y0, k, engine="hip_rosenbrock23", precision="float32", rtol=1e-5,
                   atol=1e-8, device="cuda")
    assert np.array_equal(res.y_final, y32)


def test_stiff_workload_f32_against_fp64_oracle(gpu):
    e = W.robertson(2048, seed=3)
    n = e.y0.shape[0]
    wl = np.array([1.0, 10.0, 1.0])
    print(f"case robertson workload f32 n={n}", flush=True)
    rec = H.record(N.FIELD_ROBERTSON, _soa(e.y0, gpu), _soa(e.params, gpu), t0=0.0, t1=100.0,
                   rtol=1e-6, atol=1e-10, max_steps=50000, tape_cap=1024, ckpt_interval=8,
                   fp32=True, entry="gsx_rb23_fwd")
    ybar = torch.as_tensor(np.tile(wl[:, None], (1, n)), device=gpu).contiguous()
    y0b = torch.empty((3, n), dtype=torch.float64, device=gpu)
    pb = torch.empty((3, n), dtype=torch.float64, device=gpu)
    N.call("gsx_rb23_adjoint", N.FIELD_ROBERTSON, n, 3, 3, None, None, rec.p.data_ptr(), 0.0,
           rec.tape.data_ptr(), rec.cap, rec.n_acc.data_ptr(), rec.ckpt.data_ptr(),
           rec.ckpt_interval, ybar.data_ptr(), y0b.data_ptr(), pb.data_ptr(), N.FP32,
           torch.cuda.current_stream().cuda_stream)
    # the FP32 replay reproduces the FP32 forward on its own tape
    yr = torch.empty_like(rec.y_final)
    N.call("gsx_rb23_replay", N.FIELD_ROBERTSON, n, 3, 3, None, rec.y0.data_ptr(),
           rec.p.data_ptr(), 0.0, rec.tape.data_ptr(), rec.n_acc.data_ptr(), yr.data_ptr(),
           None, 1, N.FP32, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert int((rec.status != 0).sum().item()) == 0
    assert torch.equal(yr, rec.y_final)
    tape = rec.padded_tape().cpu().numpy()
    g32 = pb.cpu().numpy().T
    oyf = O.rb23_replay(O.ROBERTSON, e.y0, e.params, 0.0, tape)
    _, g64 = O.rb23_adjoint(O.ROBERTSON, e.y0, e.params, 0.0, tape, np.tile(wl, (n, 1)))
    rel, cos = _rel_cos(g32, g64)
    per = np.abs(g32 - g64) / np.abs(g64)
    yf32 = rec.y_final.cpu().numpy().T
    ryf, _, rna, _, _ = O.rb23_record(O.ROBERTSON, e.y0, e.params, 0.0, 100.0, rtol=1e-6,
                                      atol=1e-10)
    tyf, _, _, _, _ = O.rodas5p_record(O.ROBERTSON, e.y0, e.params, 0.0, 100.0, rtol=1e-12,
                                       atol=1e-16)
    sc = np.abs(tyf).max(axis=1, keepdims=True)
    e32 = np.max(np.abs(yf32 - ryf) / sc, axis=1)
    d64 = np.max(np.abs(ryf - tyf) / sc, axis=1)
    d32 = np.max(np.abs(yf32 - tyf) / sc, axis=1)
    na = rec.n_acc.cpu().numpy()
    print(f"FP32 gradient vs FP64 adjoint on the FP32 tape: rel-L2 {rel:.2e}, cosine "
          f"1-{1 - cos:.1e}, per-entry median {np.median(per):.2e} max {per.max():.2e}",
          flush=True)
    print(f"final state vs the FP64 oracle's record: median {np.median(e32):.2e} max "
          f"{e32.max():.2e}; discretization error vs rtol-1e-12 Rodas5P: FP64 median "
          f"{np.median(d64):.2e} max {d64.max():.2e}, FP32 median {np.median(d32):.2e} max "
          f"{d32.max():.2e}; accepted steps FP32 {na.mean():.1f} vs FP64 {rna.mean():.1f}",
          flush=True)
    assert rel <= 1e-3 and cos >= 1 - 1e-6
    # the FP32 solution is as accurate as the FP64 one at this tolerance
    assert np.median(d32) < 2.0 * np.median(d64) and e32.max() < 1e-3
