"""gsx_tsit5_adjoint_saveat_noalloc (the saveat gradient with its re-step cotangents in the
library's own device array, no allocation) against gsx_tsit5_adjoint_saveat with a caller
buffer: bitwise equal on two calls in a row, the second with a smaller input (the array keeps
the first call's contents, which the second must not read)."""
import numpy as np
import pytest
import torch

from gradsolve_amd import _native as N
from gradsolve_amd import hip_tsit5 as H
from gradsolve_amd import workloads as W

pytestmark = pytest.mark.gpu


def _case(dev, n, nsave, seed):
    e = W.lorenz(n, seed=seed)
    y0 = torch.as_tensor(np.ascontiguousarray(e.y0.T), device=dev)
    p = torch.as_tensor(np.ascontiguousarray(e.params.T), device=dev)
    rec = H.record(N.FIELD_LORENZ, y0, p, t0=0.0, t1=1.0, rtol=1e-8, atol=1e-8,
                   max_steps=100000, tape_cap=256, ckpt_interval=8)
    ts = torch.linspace(0.0, 1.0, nsave, dtype=torch.float64, device=dev)
    sv = H.saveat(rec, ts)
    g = torch.Generator(device="cpu").manual_seed(seed)
    ysbar = torch.randn((nsave, 3, n), generator=g, dtype=torch.float64).to(dev)
    ybar = torch.randn((3, n), generator=g, dtype=torch.float64).to(dev)
    return rec, ts, sv, ysbar, ybar


def _noalloc(rec, ts, sv, ysbar, ybar):
    dev = rec.y0.device
    y0bar = torch.full((rec.D, rec.n), np.nan, dtype=torch.float64, device=dev)
    pbar = torch.full((rec.P, rec.n), np.nan, dtype=torch.float64, device=dev)
    N.call("gsx_tsit5_adjoint_saveat_noalloc", rec.field, rec.n, rec.D, rec.P, None,
           rec.y0.data_ptr(), rec.p.data_ptr(), rec.t0, rec.tape.data_ptr(), rec.cap,
           rec.n_acc.data_ptr(), rec.ckpt.data_ptr(), rec.ckpt_interval, ts.data_ptr(),
           int(ts.shape[0]), sv.save_step.data_ptr(), sv.save_t.data_ptr(),
           sv.states.data_ptr(), ysbar.data_ptr(), ybar.data_ptr(), y0bar.data_ptr(),
           pbar.data_ptr(), 0, torch.cuda.current_stream().cuda_stream)
    return y0bar, pbar


def test_noalloc_saveat_adjoint_matches_caller_buffer(gpu):
    for n, nsave, seed in ((6000, 33, 1), (700, 5, 2)):
        print(f"case noalloc saveat adjoint n={n} nsave={nsave}", flush=True)
        rec, ts, sv, ysbar, ybar = _case(gpu, n, nsave, seed)
        ref0, refp = H.adjoint_saveat(rec, ts, sv, ysbar, ybar)
        got0, gotp = _noalloc(rec, ts, sv, ysbar, ybar)
        torch.cuda.synchronize()
        assert int((rec.status != 0).sum().item()) == 0
        assert torch.equal(got0, ref0) and torch.equal(gotp, refp)


def test_noalloc_rejects_inputs_beyond_its_scratch(gpu):
    fake = 4096  # never dereferenced: the size check comes first
    with pytest.raises(ValueError, match="exceeds"):
        N.call("gsx_tsit5_adjoint_saveat_noalloc", N.FIELD_LORENZ, 1 << 20, 3, 1, None, fake,
               fake, 0.0, fake, 16, fake, fake, 8, fake, 3, fake, fake, fake, fake, fake,
               fake, fake, 0, None)
