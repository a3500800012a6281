"""The stiff high-order lane: Rodas5P forward (+ record), its adjoint and its dense output on
the GPU.

Mirrors the reference's ``rodas5p_replay`` engine (gradsolve/solvers/rodas5p_replay.py:
94-251: ``record_rodas5p`` through the device recorder, ``replay_rodas5p_jax`` and its
``jax.grad``; dense output by the continuous extension, gradsolve/solvers/dense.py:166-227)
with the native entry points ``gsx_rodas5p_*``. Fields need an analytic Jacobian (the
built-in stiff fields, or a field registered with ``stiff=True``); a field that reads t
brings its generated df/dt, which the stages carry as the reference's do. ``max_steps``
defaults to the reference's 50000.
"""
from __future__ import annotations

import numpy as np
import torch

from . import fields as F
from . import hip_tsit5 as H
from .base import SolveResult

MAX_STEPS = 50000
FWD = "gsx_rodas5p_fwd"


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
    aux = None if b.aux_of is None else torch.as_tensor(b.aux_of(problem), device=dev)
    yf, na, nr, st = H.forward_only(b.field_id, y0s, ps, t0=problem.t0, t1=problem.t1,
                                    rtol=rtol, atol=atol, max_steps=max_steps, strict=strict,
                                    entry=FWD, aux=aux)
    y_final = yf.t().cpu().numpy().copy()
    y_final[st.cpu().numpy() != 0] = np.nan
    return SolveResult(y_final=y_final, accepted_steps=na.cpu().numpy().astype(np.int64),
                       rejected_steps=nr.cpu().numpy().astype(np.int64), solver="hip_rodas5p")


def make_closure(binding, problem, y0, params, *, wrt, rtol, atol, device, ckpt_interval=None,
                 strict=False, aux=None):
    from .closure import make_rodas5p_closure
    return make_rodas5p_closure(binding, problem, y0, params, wrt=wrt, rtol=rtol, atol=atol,
                                device=device, ckpt_interval=ckpt_interval, strict=strict,
                                aux=aux, max_steps=MAX_STEPS)
