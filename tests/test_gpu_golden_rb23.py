"""The HIP Rosenbrock23 lane against vectors written by the reference's own numpy oracle
(``tests/_oracles/gradsolve_g2_spike_oracle.py:138-184``; fixtures from
``tests/golden/make_rb23_golden.py``): the stiff Robertson workload (1152 trajectories), the
reference's own three-trajectory Robertson batch and a constant-Jacobian linear case.

Tolerances (the reference's own claim for its Rosenbrock kernel, ``_warp_rosenbrock.py:35-42``,
is the accept/reject sequence): accepted/rejected-count mismatches < 0.1 % of trajectories;
final states within 1e-8 of the state's magnitude; tape entries of count-matched trajectories
within 1e-5 relative.

Also checks that ``strict=True`` reaches the Rosenbrock23 kernels through ``solve`` and
``grad_closure`` (the contraction-off variant, whose rounding differs from the default one).
"""
import os

import numpy as np
import pytest
import torch

import gradsolve_amd as gs
from gradsolve_amd import _native as N
from gradsolve_amd import hip_tsit5 as H
from gradsolve_amd import problems as P

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "rb23_reference_oracle.npz")


def _soa(a, dev):
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a, np.float64).T), device=dev)


def _case(g, pre):
    t0, t1 = g[pre + "_t"]
    rtol, atol = g[pre + "_tol"]
    return g[pre + "_y0"], g[pre + "_params"], float(t0), float(t1), float(rtol), float(atol)


@pytest.mark.parametrize("strict", [False, True])
def test_rb23_lane_matches_reference_oracle_vectors(gpu, strict):
    g = np.load(GOLD)
    total = mism = 0
    for pre, fid in (("c3", N.FIELD_ROBERTSON), ("c3n", N.FIELD_ROBERTSON),
                     ("rish", N.FIELD_ROBERTSON), ("lin", N.FIELD_LINEAR_DECAY)):
        y0, p, t0, t1, rtol, atol = _case(g, pre)
        print(f"case {pre} n={y0.shape[0]} strict={strict}", flush=True)
        rec = H.record(fid, _soa(y0, gpu), _soa(p, gpu), t0=t0, t1=t1, rtol=rtol, atol=atol,
                       max_steps=50000, tape_cap=1024, ckpt_interval=8, strict=strict,
                       entry="gsx_rb23_fwd")
        torch.cuda.synchronize()
        assert int((rec.status != 0).sum().item()) == 0
        na = rec.n_acc.cpu().numpy()
        nr = rec.n_rej.cpu().numpy()
        same = (na == g[pre + "_n_acc"]) & (nr == g[pre + "_n_rej"])
        total += same.size
        mism += int((~same).sum())
        yf = rec.y_final.cpu().numpy().T
        ref = g[pre + "_y_final"]
        scale = np.abs(ref).max(axis=1, keepdims=True)
        err = np.abs(yf - ref) / scale
        print(f"  count mismatches {(~same).sum()}, state err {err[same].max():.2e}",
              flush=True)
        assert err[same].max() < 1e-8
        if pre + "_tape" in g:
            rt = g[pre + "_tape"]
            tape = rec.padded_tape().cpu().numpy()
            S = rt.shape[1]
            assert tape.shape[1] >= S
            a, b = tape[same, :S], rt[same]
            nz = b != 0
            assert np.array_equal(a[~nz], np.zeros(int((~nz).sum())))
            rel = np.abs(a[nz] - b[nz]) / b[nz]
            print(f"  tape max rel {rel.max():.2e}", flush=True)
            assert rel.max() < 1e-5
    assert mism / total < 1e-3, f"{mism} of {total} trajectories changed their counts"


def test_strict_reaches_rb23_kernels(gpu):
    g = np.load(GOLD)
    y0, k = g["c3_y0"], g["c3_params"]
    prob = P.Robertson()
    direct = {}
    for strict in (False, True):
        rec = H.record(N.FIELD_ROBERTSON, _soa(y0, gpu), _soa(k, gpu), t0=0.0, t1=100.0,
                       rtol=1e-6, atol=1e-10, max_steps=50000, tape_cap=1024,
                       strict=strict, entry="gsx_rb23_fwd")
        torch.cuda.synchronize()
        direct[strict] = (rec.y_final.cpu().numpy().T, rec.padded_tape().cpu().numpy())
    # the two variants round differently somewhere in 128 stiff trajectories
    assert not np.array_equal(direct[False][1], direct[True][1])
    for strict in (False, True):
        print(f"case api strict={strict}", flush=True)
        res = gs.solve(prob, y0, k, engine="hip_rosenbrock23", rtol=1e-6, atol=1e-10,
                       device="cuda", strict=strict)
        assert res.route.actual == "hip_rosenbrock23"
        assert np.array_equal(res.y_final, direct[strict][0])
        cl = gs.grad_closure(prob, y0, k, engine="hip_rosenbrock23", rtol=1e-6, atol=1e-10,
                             device="cuda", strict=strict)
        tape = cl.dts_padded()
        ref = direct[strict][1]
        assert np.array_equal(tape, ref[:, :tape.shape[1]])
