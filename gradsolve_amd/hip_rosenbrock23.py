"""The stiff lane: fused Rosenbrock23 forward (+ record) and its adjoint on the GPU.

Mirrors the reference's stiff record-and-replay path (``warp_rosenbrock`` forward,
gradsolve/warp/warp_rosenbrock.py:166-351; ``record_rosenbrock`` and
``make_rosenbrock_replay_kernel``, gradsolve/warp/warp_replay.py:296-384) with the native
entry points ``gsx_rb23_fwd`` / ``gsx_rb23_replay`` / ``gsx_rb23_adjoint``. ``max_steps``
defaults to the reference's 50000 (stiff solves take many steps).
"""
from __future__ import annotations

import numpy as np
import torch

from . import fields as F
from . import hip_tsit5 as H
from .base import SolveResult
from .closure import make_rb23_closure

MAX_STEPS = 50000


def _tensors(problem, y0, params, dev):
    b = F.binding_for(problem)
    y0_t = y0.to(dev, torch.float64) if isinstance(y0, torch.Tensor) else \
        torch.as_tensor(np.asarray(y0, np.float64), device=dev)
    p_t = params.to(dev, torch.float64) if isinstance(params, torch.Tensor) else \
        torch.as_tensor(np.asarray(params, np.float64), device=dev)
    return b, y0_t.t().contiguous(), b.native_params(p_t).t().contiguous()


def solve(problem, y0, params, *, rtol, atol, device, strict=False,
          max_steps=MAX_STEPS) -> SolveResult:
    from .api import _device
    dev = _device(device)
    b, y0s, ps = _tensors(problem, y0, params, dev)
    yf, na, nr, st = H.forward_only(b.field_id, y0s, ps, t0=problem.t0, t1=problem.t1,
                                    rtol=rtol, atol=atol, max_steps=max_steps, strict=strict,
                                    entry="gsx_rb23_fwd")
    y_final = yf.t().cpu().numpy().copy()
    y_final[st.cpu().numpy() != 0] = np.nan
    return SolveResult(y_final=y_final, accepted_steps=na.cpu().numpy().astype(np.int64),
                       rejected_steps=nr.cpu().numpy().astype(np.int64),
                       solver="hip_rosenbrock23")


def make_closure(binding, problem, y0, params, *, wrt, rtol, atol, device, ckpt_interval=None,
                 strict=False):
    return make_rb23_closure(binding, problem, y0, params, wrt=wrt, rtol=rtol, atol=atol,
                             device=device, ckpt_interval=ckpt_interval, strict=strict,
                             max_steps=MAX_STEPS)
