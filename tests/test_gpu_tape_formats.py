"""Tape interchange formats on the GPU (gradsolve/warp/warp_ode.py:164-212 trimmed [S, n]
DLPack hand-off; gradsolve/warp/warp_replay.py:180-182 zero-padded [n, S]):

  * Tsit5Record.trimmed_tape() is a zero-copy [S, n] view of the record's own tape, zero past
    each trajectory's n_acc, equal to the transpose of the padded [n, S] export bit for bit,
    and crosses DLPack without a copy;
  * a trimmed tape imported with record_from_trimmed (from a device tensor, a DLPack capsule
    holder or a host NumPy array, as the reference's two transports deliver it) recovers n_acc
    exactly, replays to the recorded final state (rel 1e-13), and its adjoint gives the
    recorded run's gradients (rel 1e-12);
  * the padded import (gsx_tape_from_padded) gives back the same tape rows and counts.
"""
import numpy as np
import pytest
import torch
import torch.utils.dlpack

from gradsolve_amd import _native as N
from gradsolve_amd import hip_tsit5 as H
from gradsolve_amd import workloads as W

pytestmark = pytest.mark.gpu


def _soa(a, dev):
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a, np.float64).T), device=dev)


def test_trimmed_export_import_round_trip(gpu):
    e = W.lorenz(4096, seed=51, noise=0.1)
    print("case tape formats lorenz n=4096", flush=True)
    y0, p = _soa(e.y0, gpu), _soa(e.params, gpu)
    rec = H.record(N.FIELD_LORENZ, y0, p, t0=0.0, t1=1.0, rtol=1e-8, atol=1e-8,
                   max_steps=100000, ckpt_interval=8)
    torch.cuda.synchronize()
    trimmed = rec.trimmed_tape()
    S = rec.max_steps_taken()
    assert tuple(trimmed.shape) == (S, 4096)
    assert trimmed.data_ptr() == rec.tape.data_ptr()
    padded = rec.padded_tape(S)
    torch.cuda.synchronize()
    assert torch.equal(trimmed.t(), padded)
    via = torch.utils.dlpack.from_dlpack(trimmed.__dlpack__())
    assert via.data_ptr() == trimmed.data_ptr() and torch.equal(via, trimmed)
    gy, gp = H.adjoint(rec, 2.0 * rec.y_final)
    for src in (trimmed.clone(), trimmed.cpu().numpy()):
        rec2 = H.record_from_trimmed(N.FIELD_LORENZ, y0, p, src, t0=0.0, t1=1.0,
                                     ckpt_interval=8)
        torch.cuda.synchronize()
        assert torch.equal(rec2.n_acc, rec.n_acc)
        assert torch.max(torch.abs(rec2.y_final - rec.y_final) /
                         torch.abs(rec.y_final)).item() < 1e-13
        hy, hp = H.adjoint(rec2, 2.0 * rec.y_final)
        torch.cuda.synchronize()
        assert torch.max(torch.abs(hp - gp) / torch.abs(gp)).item() < 1e-12
        scale = torch.abs(gy).max(dim=0, keepdim=True).values
        assert torch.max(torch.abs(hy - gy) / scale).item() < 1e-12
    tape_p, nacc_p = H.record_from_padded(N.FIELD_LORENZ, y0, p, padded, t0=0.0, t1=1.0)
    torch.cuda.synchronize()
    assert torch.equal(nacc_p, rec.n_acc)
    assert torch.equal(tape_p[:S], trimmed)
