"""The C-ABI forward entry points bound the way the reference's forward handlers are
(``Tsit5Fwd``: y0 [D, n], rho [n] -> y_final [D, n], nsteps [n];
gradsolve/cuda/kernels/tsit5_template.cu.in:127-178; ``Rosenbrock23Fwd``:
rosenbrock23_template.cu.in:151-223): no tape, the rejected count and the status left null
(and the accepted count too). Results equal the full call's bit for bit."""
import numpy as np
import pytest
import torch

from gradsolve_amd import _native as N
from gradsolve_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = [("gsx_tsit5_fwd", N.FIELD_LORENZ, lambda: W.lorenz(4099, seed=8)),
         ("gsx_rb23_fwd", N.FIELD_ROBERTSON, lambda: W.robertson(1031, seed=8)),
         ("gsx_rodas5p_fwd", N.FIELD_ROBERTSON, lambda: W.robertson(1031, seed=9))]


@pytest.mark.parametrize("entry,field,make", CASES, ids=[c[0] for c in CASES])
def test_forward_binds_like_the_reference_handler(gpu, entry, field, make):
    e = make()
    n, D = e.y0.shape
    P = e.params.shape[1]
    print(f"case {entry} n={n} null counts", flush=True)
    y0 = torch.as_tensor(np.ascontiguousarray(e.y0.T), device=gpu)
    p = torch.as_tensor(np.ascontiguousarray(e.params.T), device=gpu)
    f64 = dict(dtype=torch.float64, device=gpu)
    i32 = dict(dtype=torch.int32, device=gpu)
    s = torch.cuda.current_stream().cuda_stream
    args = (field, n, D, P, None, y0.data_ptr(), p.data_ptr(), e.t0, e.t1, None, e.rtol,
            e.atol, 50000)
    yf_full = torch.empty((D, n), **f64)
    na, nr, st = torch.empty(n, **i32), torch.empty(n, **i32), torch.empty(n, **i32)
    N.call(entry, *args, yf_full.data_ptr(), na.data_ptr(), nr.data_ptr(), st.data_ptr(),
           None, 0, None, 1, 0, s)
    yf = torch.full((D, n), -1.0, **f64)
    ns = torch.full((n,), -7, **i32)
    N.call(entry, *args, yf.data_ptr(), ns.data_ptr(), None, None, None, 0, None, 1, 0, s)
    yf_only = torch.full((D, n), -1.0, **f64)
    N.call(entry, *args, yf_only.data_ptr(), None, None, None, None, 0, None, 1, 0, s)
    torch.cuda.synchronize()
    assert int((st != 0).sum().item()) == 0
    assert torch.equal(yf, yf_full) and torch.equal(yf_only, yf_full)
    assert torch.equal(ns, na)
