"""``precision="float32"``: the FP32 Tsit5 lane behind ``solve`` and ``grad_closure``
(gradsolve/api.py:253-274, gradsolve/warp/warp_replay.py:87-118).

The reference's float32 route is all-FP32 (time, controller and tape included). This lane
keeps what decides the mesh in FP64 -- time, the error norm, the controller and the tape --
and runs the stages in FP32 with the state carried as a compensated (Kahan) pair, so its
error stays below the FP64 solution's own discretization error at the same tolerance. The
adjoint is the FP32 step VJP with compensated accumulators. Results are returned as float64.

Only the non-stiff Tsit5 lane has an FP32 variant; asking for float32 on the Rosenbrock23
lane raises ``ValueError`` as the reference does for lanes without an f32 adjoint
(gradsolve/api.py:262-274).
"""
from __future__ import annotations

import numpy as np
import torch

from . import dispatch as D
from . import hip_tsit5 as H
from .base import SolveResult
from .closure import make_tsit5_f32_closure

F32_CAPABLE = frozenset({D.HIP_TSIT5, D.HIP_REPLAY, D.HIP_KURAMOTO})


def reject_incapable(name: str) -> None:
    if name not in F32_CAPABLE:
        raise ValueError(
            f"precision='float32' is not supported on lane {name!r}: the FP32 variant is built "
            "for the non-stiff Tsit5 lanes (hip_tsit5, hip_replay, hip_kuramoto) only. Use "
            "precision='float64'.")


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
    yf, na, nr, st = H.forward_only(b.field_id, y0_t.t().contiguous(),
                                    b.native_params(p_t).t().contiguous(), t0=problem.t0,
                                    t1=problem.t1, rtol=rtol, atol=atol, max_steps=max_steps,
                                    aux=aux, fp32=True)
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
    return make_tsit5_f32_closure(b, problem, y0, params, wrt=wrt, rtol=rtol, atol=atol,
                                  device=device, ckpt_interval=ckpt_interval, aux=aux)
