"""Host side of the Kuramoto ensemble engine (``gsx_kuramoto_*``): D = P = 32, one coupling
matrix K shared by the ensemble, coupling sums on FP64 MFMA with 16 trajectories per
wavefront.

Record -> saveat -> adjoint replaces the reference's general lane for this problem
(record_tsit5_jax, gradsolve/solvers/tsit5_replay.py:127-189; replay_solve_saveat_jax,
gradsolve/warp/warp_replay.py:58-75; jax.grad): the forward records every accepted step's
size and start state (HBM holds them: 256 bytes per trajectory per step), so the adjoint reads
each step's state instead of re-integrating, and the save times' bracketing steps, so saveat is
one batched re-step per save time.

Layouts: y0, omega, y_final [32, n] (device, float64); ys [n, k, 32].
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _native as N
from .hip_tsit5 import _ptr
from .hip_tsit5 import _stream_handle as _stream

D = 32


def validate_saveat(saveat, t0: float, t1: float) -> np.ndarray:
    """Sorted (non-decreasing), finite, inside the domain (gradsolve/solvers/dense.py:81-103)."""
    ts = np.asarray(saveat, dtype=np.float64)
    if ts.ndim != 1:
        raise ValueError(f"saveat must be 1-D (k,); got shape {ts.shape}")
    if ts.size == 0:
        raise ValueError("saveat must contain at least one time")
    if not np.all(np.isfinite(ts)):
        raise ValueError("saveat contains non-finite times")
    if np.any(np.diff(ts) < 0):
        raise ValueError(f"saveat must be sorted non-decreasing; got {ts}")
    lo, hi = (t0, t1) if t1 >= t0 else (t1, t0)
    if ts[0] < lo or ts[-1] > hi:
        raise ValueError(f"saveat times must lie within the problem domain [{lo}, {hi}]; got "
                         f"[{ts[0]}, {ts[-1]}]")
    return ts


@dataclass
class KuramotoRecord:
    n: int
    K: torch.Tensor          # [32, 32]
    y0: torch.Tensor         # [32, n]
    omega: torch.Tensor      # [32, n]
    t0: float
    t1: float
    ts: Optional[torch.Tensor]
    y_final: torch.Tensor
    n_acc: torch.Tensor
    n_rej: torch.Tensor
    status: torch.Tensor
    tape: torch.Tensor       # [cap, n]
    cap: int
    states: torch.Tensor     # [cap, n, 32]
    save_step: Optional[torch.Tensor]
    save_t: Optional[torch.Tensor]
    flags: int = 0           # 0 (FP64) or N.FP32

    @property
    def nsave(self) -> int:
        return 0 if self.ts is None else int(self.ts.numel())

    def max_steps_taken(self) -> int:
        return int(self.n_acc.max().item()) if self.n else 0

    def padded_tape(self) -> torch.Tensor:
        """The reference's zero-padded ``[n, S]`` tape (S = max accepted steps)."""
        S = self.max_steps_taken()
        s = torch.arange(S, device=self.tape.device)[:, None]
        t = self.tape[:S]
        return torch.where(s < self.n_acc[None, :].to(torch.int64), t,
                           torch.zeros((), dtype=t.dtype, device=t.device)).t().contiguous()


def forward_only(K, y0_soa, om_soa, *, t0, t1, rtol, atol, max_steps=100000, stream=None,
                 fp32=False):
    """Forward without a record: (y_final [32, n], n_acc, n_rej, status)."""
    n = y0_soa.shape[1]
    dev = y0_soa.device
    y_final = torch.empty((D, n), dtype=torch.float64, device=dev)
    n_acc = torch.empty(n, dtype=torch.int32, device=dev)
    n_rej = torch.empty(n, dtype=torch.int32, device=dev)
    status = torch.empty(n, dtype=torch.int32, device=dev)
    N.call("gsx_kuramoto_fwd", n, D, _ptr(K), _ptr(y0_soa), _ptr(om_soa), float(t0),
           float(t1), float(rtol), float(atol), int(max_steps), None, 0, _ptr(y_final),
           _ptr(n_acc), _ptr(n_rej), _ptr(status), None, 0, None, None, None,
           N.FP32 if fp32 else 0, _stream(stream))
    return y_final, n_acc, n_rej, status


def estimate_cap(K, y0_soa, om_soa, *, t0, t1, rtol, atol, sample=2048, seed=0,
                 fp32=False) -> int:
    """Tape rows for the ensemble: 1.25 x the largest accepted-step count on a random
    subsample, plus 16 (a trajectory that still does not fit grows the tape)."""
    n = y0_soa.shape[1]
    m = min(n, sample)
    g = torch.Generator(device="cpu").manual_seed(seed)
    idx = torch.randperm(n, generator=g)[:m].to(y0_soa.device)
    _, na, _, _ = forward_only(K, y0_soa[:, idx].contiguous(), om_soa[:, idx].contiguous(),
                               t0=t0, t1=t1, rtol=rtol, atol=atol, fp32=fp32)
    return int(1.25 * int(na.max().item())) + 16 if m else 16


def record(K, y0_soa, om_soa, *, t0, t1, rtol, atol, ts=None, max_steps=100000,
           tape_cap=None, grow=True, stream=None, fp32=False) -> KuramotoRecord:
    """Forward + record (tape, states, save brackets). ``ts`` is a device float64 tensor of
    validated save times or None. Grows the tape on status 3. ``fp32`` selects the FP32-MFMA
    variant (FP32 stages and coupling, compensated state, FP64 t, h, norm and controller)."""
    flags = N.FP32 if fp32 else 0
    n = y0_soa.shape[1]
    dev = y0_soa.device
    if tape_cap is None:
        tape_cap = estimate_cap(K, y0_soa, om_soa, t0=t0, t1=t1, rtol=rtol, atol=atol,
                                fp32=fp32)
    cap = int(max(1, min(tape_cap, max_steps)))
    k = 0 if ts is None else int(ts.numel())
    while True:
        y_final = torch.empty((D, n), dtype=torch.float64, device=dev)
        n_acc = torch.empty(n, dtype=torch.int32, device=dev)
        n_rej = torch.empty(n, dtype=torch.int32, device=dev)
        status = torch.empty(n, dtype=torch.int32, device=dev)
        tape = torch.empty((cap, n), dtype=torch.float64, device=dev)
        states = torch.empty((cap, n, D), dtype=torch.float64, device=dev)
        save_step = torch.empty((k, n), dtype=torch.int32, device=dev) if k else None
        save_t = torch.empty((k, n), dtype=torch.float64, device=dev) if k else None
        N.call("gsx_kuramoto_fwd", n, D, _ptr(K), _ptr(y0_soa), _ptr(om_soa), float(t0),
               float(t1), float(rtol), float(atol), int(max_steps), _ptr(ts), k,
               _ptr(y_final), _ptr(n_acc), _ptr(n_rej), _ptr(status), _ptr(tape), cap,
               _ptr(states), _ptr(save_step), _ptr(save_t), flags, _stream(stream))
        rec = KuramotoRecord(n, K, y0_soa, om_soa, float(t0), float(t1), ts, y_final, n_acc,
                             n_rej, status, tape, cap, states, save_step, save_t, flags)
        if not grow or cap >= max_steps or n == 0:
            return rec
        if not bool((status == 3).any().item()):
            return rec
        del states, tape
        cap = int(min(2 * cap, max_steps))


def replay(rec: KuramotoRecord, y0_soa, om_soa, stream=None):
    """The frozen mesh from new (y0, omega): (y_final [32, n], states)."""
    y_final = torch.empty_like(rec.y_final)
    states = torch.empty_like(rec.states)
    N.call("gsx_kuramoto_replay", rec.n, D, _ptr(rec.K), _ptr(y0_soa), _ptr(om_soa),
           _ptr(rec.tape), _ptr(rec.n_acc), _ptr(y_final), _ptr(states), rec.flags,
           _stream(stream))
    return y_final, states


def saveat(rec: KuramotoRecord, y0_soa=None, om_soa=None, y_final=None, states=None,
           out=None, stream=None) -> torch.Tensor:
    """ys [n, k, 32] on the recorded mesh (from the record's or replayed inputs)."""
    y0_soa = rec.y0 if y0_soa is None else y0_soa
    om_soa = rec.omega if om_soa is None else om_soa
    y_final = rec.y_final if y_final is None else y_final
    states = rec.states if states is None else states
    k = rec.nsave
    ys = out if out is not None else torch.empty((rec.n, k, D), dtype=torch.float64,
                                                 device=rec.y0.device)
    N.call("gsx_kuramoto_saveat", rec.n, D, _ptr(rec.K), _ptr(y0_soa), _ptr(om_soa),
           _ptr(y_final), _ptr(states), _ptr(rec.ts), k, _ptr(rec.save_step),
           _ptr(rec.save_t), _ptr(ys), rec.flags, _stream(stream))
    return ys


def adjoint(rec: KuramotoRecord, *, ysbar=None, yfbar=None, om_soa=None, states=None,
            mu=None, out=None, stream=None):
    """(y0bar [32, n], omega_bar [32, n]) for cotangents ysbar [n, k, 32] and/or
    yfbar [32, n]."""
    om_soa = rec.omega if om_soa is None else om_soa
    states = rec.states if states is None else states
    dev = rec.y0.device
    if ysbar is not None:
        ysbar = ysbar.to(torch.float64).contiguous()
        if mu is None:
            mu = torch.empty_like(ysbar)
    if yfbar is not None:
        yfbar = yfbar.to(torch.float64).contiguous()
    if out is None:
        y0bar = torch.empty((D, rec.n), dtype=torch.float64, device=dev)
        ombar = torch.empty((D, rec.n), dtype=torch.float64, device=dev)
    else:
        y0bar, ombar = out
    N.call("gsx_kuramoto_adjoint", rec.n, D, _ptr(rec.K), _ptr(om_soa), _ptr(rec.tape),
           _ptr(rec.n_acc), _ptr(states), _ptr(rec.ts), rec.nsave, _ptr(rec.save_step),
           _ptr(rec.save_t), _ptr(ysbar), _ptr(yfbar), _ptr(mu), _ptr(y0bar), _ptr(ombar),
           rec.flags, _stream(stream))
    return y0bar, ombar
