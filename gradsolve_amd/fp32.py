"""``precision="float32"``: the FP32 lanes behind ``solve`` and ``grad_closure``
(gradsolve/api.py:253-274, gradsolve/warp/warp_replay.py:87-118).

The reference's float32 route is all-FP32 (time, controller and tape included) and covers
its registered-field lanes, the non-stiff ``warp_ode``/``warp_replay`` and the stiff
``warp_rosenbrock`` (``_F32_CAPABLE``, gradsolve/api.py:259). These lanes keep what decides
the mesh in FP64 -- time, the error norm, the controller and the tape -- and run the stages in
FP32 with the state carried as a compensated (Kahan) pair, so the error stays below the FP64
solution's own discretization error at the same tolerance. The Rosenbrock23 lane also keeps
W = I - d h J, its LU factorisation and the triangular solves in FP64. The adjoints are the
FP32 step VJPs with compensated accumulators. Results are returned as float64.

Lanes without an FP32 variant (Rodas5P) raise ``ValueError`` naming float32, as the reference
does for its f64-only lanes (gradsolve/api.py:262-274).
"""
from __future__ import annotations

import numpy as np
import torch

from . import dispatch as D
from . import hip_tsit5 as H
from .base import SolveResult
from .closure import make_rb23_closure, make_tsit5_f32_closure

F32_CAPABLE = frozenset({D.HIP_TSIT5, D.HIP_REPLAY, D.HIP_KURAMOTO, D.HIP_ROSENBROCK23})


def reject_incapable(name: str) -> None:
    if name not in F32_CAPABLE:
        raise ValueError(
            f"precision='float32' is not supported on lane {name!r}: the FP32 variant is built "
            "for the Tsit5 lanes (hip_tsit5, hip_replay, hip_kuramoto) and the Rosenbrock23 "
            "lane (hip_rosenbrock23). Use precision='float64'.")


def solve(b, problem, y0, params, *, rtol, atol, device, name=D.HIP_TSIT5,
          max_steps=100000) -> SolveResult:
    from .api import _device, _solve_kuramoto
    reject_incapable(name)
    if name == D.HIP_KURAMOTO:
        return _solve_kuramoto(problem, y0, params, rtol=rtol, atol=atol, device=device,
                               max_steps=max_steps, fp32=True)
    dev = _device(device)
    y0_t = torch.as_tensor(np.asarray(y0, np.float64), device=dev) \
        if not isinstance(y0, torch.Tensor) else y0.to(dev, torch.float64)
    p_t = torch.as_tensor(np.asarray(params, np.float64), device=dev) \
        if not isinstance(params, torch.Tensor) else params.to(dev, torch.float64)
    aux = None
    if b.aux_of is not None:
        aux = torch.as_tensor(b.aux_of(problem), device=dev)
    stiff = name == D.HIP_ROSENBROCK23
    yf, na, nr, st = H.forward_only(b.field_id, y0_t.t().contiguous(),
                                    b.native_params(p_t).t().contiguous(), t0=problem.t0,
                                    t1=problem.t1, rtol=rtol, atol=atol,
                                    max_steps=50000 if stiff else max_steps, aux=aux,
                                    fp32=True, entry="gsx_rb23_fwd" if stiff else "gsx_tsit5_fwd")
    y_final = yf.t().cpu().numpy().copy()
    status = st.cpu().numpy()
    y_final[status != 0] = np.nan
    return SolveResult(y_final=y_final, accepted_steps=na.cpu().numpy().astype(np.int64),
                       rejected_steps=nr.cpu().numpy().astype(np.int64), solver=name)


def make_closure(name, b, problem, y0, params, *, wrt, rtol, atol, device, aux=None,
                 ckpt_interval=None):
    reject_incapable(name)
    if name == D.HIP_KURAMOTO:
        from .closure import KuramotoClosure
        return KuramotoClosure(b, problem, y0, params, wrt=wrt, rtol=rtol, atol=atol,
                               device=device, aux=aux, fp32=True)
    if name == D.HIP_ROSENBROCK23:
        return make_rb23_closure(b, problem, y0, params, wrt=wrt, rtol=rtol, atol=atol,
                                 device=device, ckpt_interval=ckpt_interval, aux=aux, fp32=True)
    return make_tsit5_f32_closure(b, problem, y0, params, wrt=wrt, rtol=rtol, atol=atol,
                                  device=device, ckpt_interval=ckpt_interval, aux=aux)
