"""The wide-state Tsit5 lane: the native engine of the high-NVAR cells.

The reference routes non-stiff problems above ``NVAR_CEILING`` (64) to its general
record-and-replay engine -- ``record_tsit5_jax`` with the vmapped while-loop recorder and a
``jax.grad`` of the replay scan (gradsolve/dispatch.py:256-258, gradsolve/solvers/
record_jax.py:171-213, tsit5_replay.py:127-232) -- and forward-only solves to diffrax, or to
that same lane when diffrax is absent. Here such a problem's right-hand side is compiled at
run time (``register_field`` / ``fused='auto'``) and runs on ``gsx_tsit5_wide_*``: one
trajectory per lane like the register lanes, with the state, stage derivatives and stage
inputs in a per-trajectory workspace in HBM laid out [row][D][n], so any D up to 4096 fits.
The recorder keeps the register lanes' contract (status codes 0/1/3, tape capacity doubling
on status 3 as record_jax.py:192-199 does, sqrt(S)-spaced checkpoints) and the adjoint is the
exact discrete adjoint of the frozen-step replay.
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from . import _native as N
from . import hip_tsit5 as H
from .base import SolveResult
from .closure import ReplayClosure, _Lane

MAX_STEPS = 100000


def _ptr(t):
    return None if t is None else t.data_ptr()


def workspace(which: int, D: int, n: int, device, ckpt_interval: int = 1) -> torch.Tensor:
    """Workspace of the forward / replay (which 0) or the adjoint (which 1)."""
    rows = int(N.call("gsx_tsit5_wide_ws_rows", which, int(ckpt_interval)))
    return torch.empty((rows, D, n), dtype=torch.float64, device=device)


def record(field: int, y0_soa: torch.Tensor, p_soa: torch.Tensor, *, t0: float, t1: float,
           rtol: float, atol: float, max_steps: int = MAX_STEPS,
           t1_per: Optional[torch.Tensor] = None, aux: Optional[torch.Tensor] = None,
           tape_cap: Optional[int] = None, ckpt_interval: Optional[int] = None,
           grow: bool = True) -> H.Tsit5Record:
    """Forward with the accepted-h record and checkpoints (gsx_tsit5_wide_fwd); the tape grows
    by doubling while a trajectory reports status 3."""
    D, n = y0_soa.shape
    P = p_soa.shape[0]
    dev = y0_soa.device
    cap = max(1, int(min(max_steps, tape_cap if tape_cap else H.DEFAULT_TAPE_CAP)))
    ws = workspace(0, D, n, dev)
    while True:
        C = int(ckpt_interval) if ckpt_interval else H.default_ckpt_interval(cap // 2)
        C = max(1, min(H.MAX_CKPT_INTERVAL, C))
        y_final = torch.empty((D, n), dtype=torch.float64, device=dev)
        n_acc = torch.empty(n, dtype=torch.int32, device=dev)
        n_rej = torch.empty(n, dtype=torch.int32, device=dev)
        status = torch.empty(n, dtype=torch.int32, device=dev)
        tape = torch.empty((cap, n), dtype=torch.float64, device=dev)
        ckpt = H._alloc_ckpt(field, D, n, cap, C, dev)
        N.call("gsx_tsit5_wide_fwd", field, n, D, P, _ptr(aux), _ptr(y0_soa), _ptr(p_soa),
               float(t0), float(t1), _ptr(t1_per), float(rtol), float(atol), int(max_steps),
               _ptr(y_final), _ptr(n_acc), _ptr(n_rej), _ptr(status), _ptr(tape), cap,
               _ptr(ckpt), C, _ptr(ws), H._stream_handle(None))
        rec = H.Tsit5Record(field, n, D, P, float(t0), float(t1), y0_soa, p_soa, t1_per, aux,
                            y_final, n_acc, n_rej, status, tape, cap, ckpt, C)
        if not grow or cap >= max_steps or n == 0:
            return rec
        if not bool((status == 3).any().item()):
            return rec
        cap = int(min(max_steps, cap * 2))


def forward_only(field: int, y0_soa: torch.Tensor, p_soa: torch.Tensor, *, t0: float,
                 t1: float, rtol: float, atol: float, max_steps: int = MAX_STEPS,
                 aux: Optional[torch.Tensor] = None):
    """(y_final, n_acc, n_rej, status) without a tape."""
    D, n = y0_soa.shape
    P = p_soa.shape[0]
    dev = y0_soa.device
    ws = workspace(0, D, n, dev)
    y_final = torch.empty((D, n), dtype=torch.float64, device=dev)
    n_acc = torch.empty(n, dtype=torch.int32, device=dev)
    n_rej = torch.empty(n, dtype=torch.int32, device=dev)
    status = torch.empty(n, dtype=torch.int32, device=dev)
    N.call("gsx_tsit5_wide_fwd", field, n, D, P, _ptr(aux), _ptr(y0_soa), _ptr(p_soa),
           float(t0), float(t1), None, float(rtol), float(atol), int(max_steps), _ptr(y_final),
           _ptr(n_acc), _ptr(n_rej), _ptr(status), None, 0, None, 1, _ptr(ws),
           H._stream_handle(None))
    return y_final, n_acc, n_rej, status


class _WideLane(_Lane):
    """The record / replay / adjoint calls of ReplayClosure, on the wide entry points."""

    FWD, REPLAY, ADJ = "gsx_tsit5_wide_fwd", "gsx_tsit5_wide_replay", "gsx_tsit5_wide_adjoint"

    def __init__(self, binding, aux):
        super().__init__(binding, aux, False)

    def record(self, y0_soa, p_soa, *, t0, t1, rtol, atol, max_steps, ckpt_interval):
        return record(self.b.field_id, y0_soa, p_soa, t0=t0, t1=t1, rtol=rtol, atol=atol,
                      max_steps=max_steps, aux=self.aux, ckpt_interval=ckpt_interval)

    def replay(self, rec, y0_soa, p_soa):
        y_final = torch.empty_like(rec.y_final)
        ckpt = torch.empty_like(rec.ckpt)
        ws = workspace(0, rec.D, rec.n, y0_soa.device)
        N.call(self.REPLAY, rec.field, rec.n, rec.D, rec.P, _ptr(self.aux), _ptr(y0_soa),
               _ptr(p_soa), rec.t0, _ptr(rec.tape), _ptr(rec.n_acc), _ptr(y_final),
               _ptr(ckpt), rec.ckpt_interval, _ptr(ws), H._stream_handle(None))
        return y_final, ckpt

    def adjoint(self, rec, ckpt, p_soa, ybar_soa):
        dev = ybar_soa.device
        y0bar = torch.empty((rec.D, rec.n), dtype=torch.float64, device=dev)
        pbar = torch.empty((rec.P, rec.n), dtype=torch.float64, device=dev)
        ws = workspace(1, rec.D, rec.n, dev, rec.ckpt_interval)
        N.call(self.ADJ, rec.field, rec.n, rec.D, rec.P, _ptr(self.aux), None, _ptr(p_soa),
               rec.t0, _ptr(rec.tape), rec.cap, _ptr(rec.n_acc), _ptr(ckpt),
               rec.ckpt_interval, _ptr(ybar_soa.contiguous()), _ptr(y0bar), _ptr(pbar),
               _ptr(ws), H._stream_handle(None))
        return y0bar, pbar


def solve(problem, y0, params, *, rtol, atol, device, max_steps=MAX_STEPS) -> SolveResult:
    from . import fields as F
    from .api import _device
    b = F.binding_for(problem)
    dev = _device(device)
    y0_t = y0.to(dev, torch.float64) if isinstance(y0, torch.Tensor) else \
        torch.as_tensor(np.asarray(y0, np.float64), device=dev)
    p_t = params.to(dev, torch.float64) if isinstance(params, torch.Tensor) else \
        torch.as_tensor(np.asarray(params, np.float64), device=dev)
    aux = torch.as_tensor(b.aux_of(problem), device=dev) if b.aux_of is not None else None
    yf, na, nr, st = forward_only(b.field_id, y0_t.t().contiguous(),
                                  b.native_params(p_t).t().contiguous(), t0=problem.t0,
                                  t1=problem.t1, rtol=rtol, atol=atol, max_steps=max_steps,
                                  aux=aux)
    y_final = yf.t().cpu().numpy().copy()
    y_final[st.cpu().numpy() != 0] = np.nan
    return SolveResult(y_final=y_final, accepted_steps=na.cpu().numpy().astype(np.int64),
                       rejected_steps=nr.cpu().numpy().astype(np.int64),
                       solver="hip_tsit5_wide")


def make_closure(binding, problem, y0, params, *, wrt, rtol, atol, device, ckpt_interval=None,
                 aux=None, max_steps=MAX_STEPS) -> ReplayClosure:
    return ReplayClosure(_WideLane(binding, aux), binding, problem, y0, params, wrt=wrt,
                         rtol=rtol, atol=atol, device=device, ckpt_interval=ckpt_interval,
                         max_steps=max_steps)
