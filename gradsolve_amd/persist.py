"""Host side of the persistent work-queue Tsit5 engine (``gsx_tsit5_persist_fwd`` /
``gsx_tsit5_persist_adjoint``): buffer sizing, the forward+adjoint pair, and capture of the
pair in a HIP graph.

Layout. tape[lane_cap/16][lanes][16] (a lane's steps 16k..16k+15 contiguous),
ckpt[ckpt_cap][lanes][rows], order[traj_cap][lanes].

Capacities. Each lane appends to its own tape, checkpoint and trajectory-list streams, so
the per-lane capacities follow from the expected number of accepted steps per lane:
``lane_cap = margin * total_steps / lanes + longest_trajectory`` (the longest trajectory
bounds what one lane can add after crossing the average). ``estimate_steps`` measures the
step-count distribution on a random subsample with the per-trajectory kernel. A trajectory
that still does not fit gets status 3, and ``run_forward`` doubles the capacities and runs
again, as the reference's device recorder doubles its tape (gradsolve/solvers/record_jax.py:
192-199).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _native as N
from . import hip_tsit5 as H


TAPE_BLOCK = 16  # accepted steps of one lane stored contiguously in the tape


def _ptr(t):
    return None if t is None else t.data_ptr()


@dataclass
class StepEstimate:
    mean_steps: float
    max_steps: int


def estimate_steps(field, y0_soa, p_soa, *, t0, t1, t1_per, rtol, atol, sample=4096,
                   seed=0, aux=None, fp32=False) -> StepEstimate:
    """Accepted-step statistics on a random subsample (per-trajectory forward kernel)."""
    n = y0_soa.shape[1]
    m = min(n, sample)
    g = torch.Generator(device="cpu").manual_seed(seed)
    idx = torch.randperm(n, generator=g)[:m].to(y0_soa.device)
    t1s = None if t1_per is None else t1_per[idx].contiguous()
    yf, na, nr, st = H.forward_only(field, y0_soa[:, idx].contiguous(),
                                    p_soa[:, idx].contiguous(), t0=t0, t1=t1, t1_per=t1s,
                                    rtol=rtol, atol=atol, max_steps=10 ** 7, aux=aux)
    na = na.to(torch.float64)
    return StepEstimate(float(na.mean().item()), int(na.max().item()))


class PersistentTsit5:
    """Workspace and launches of the persistent forward + adjoint for a fixed ensemble."""

    def __init__(self, field: int, D: int, P: int, n: int, *, est: StepEstimate,
                 waves_per_simd: int = 4, ckpt_interval: int = 8, fp32: bool = False,
                 margin: float = 1.3, device=None, aux: Optional[torch.Tensor] = None):
        self.field, self.D, self.P, self.n = field, D, P, n
        self.C = int(ckpt_interval)
        self.fp32 = bool(fp32)
        self.aux = aux
        self.device = device
        self.waves_per_simd = int(waves_per_simd)
        self.lanes = int(N.call("gsx_persist_lanes", self.waves_per_simd))
        self.rows = int(N.call("gsx_ckpt_rows", field, D))
        L = self.lanes
        self.lane_cap = int(math.ceil(margin * est.mean_steps * n / L)) + 2 * est.max_steps + 8
        self.traj_cap = int(math.ceil(margin * n / L)) + 8
        self.queue_order: Optional[torch.Tensor] = None  # see plan_longest_first
        self._alloc(margin)

    def _alloc(self, margin):
        L, n, dev = self.lanes, self.n, self.device
        # the tape stores a lane's steps in blocks of TAPE_BLOCK (one 128-byte line)
        self.lane_cap = -(-self.lane_cap // TAPE_BLOCK) * TAPE_BLOCK
        self.ckpt_cap = self.lane_cap // self.C + self.traj_cap + 8
        f64 = dict(dtype=torch.float64, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        self.tape = torch.empty((self.lane_cap // TAPE_BLOCK, L, TAPE_BLOCK), **f64)
        self.ckpt = torch.empty((self.ckpt_cap, L, self.rows), **f64)
        self.order = torch.empty((self.traj_cap, L), **i32)
        self.lane_count = torch.empty(L, **i32)
        self.lane_of = torch.empty(n, **i32)
        self.tape_start = torch.empty(n, **i32)
        self.ckpt_start = torch.empty(n, **i32)
        self.queue = torch.zeros(1, dtype=torch.int64, device=dev)
        self.y_final = torch.empty((self.D, n), **f64)
        self.n_acc = torch.empty(n, **i32)
        self.n_rej = torch.empty(n, **i32)
        self.status = torch.empty(n, **i32)
        self.ybar = torch.empty((self.D, n), **f64)
        self.y0bar = torch.empty((self.D, n), **f64)
        self.pbar = torch.empty((self.P, n), **f64)

    @property
    def flags(self):
        return N.FP32 if self.fp32 else 0

    def plan_longest_first(self, t1_per: Optional[torch.Tensor], t0: float = 0.0) -> None:
        """Have the queue hand out trajectories in order of decreasing integration length
        (t1_per - t0), a proxy for their step counts, so that the lanes' last trajectories
        are short and few lanes idle at the end (measured on 2^22 uneven Lorenz: forward
        11.3 -> 10.4 ms, forward + adjoint 29.9 -> 28.1 ms at two waves per SIMD). Any
        permutation gives the same results; only the schedule changes. No-op without
        per-trajectory end times."""
        if t1_per is None:
            self.queue_order = None
            return
        if t1_per.numel() != self.n:
            raise ValueError(f"t1_per has {t1_per.numel()} entries for {self.n} trajectories")
        order = torch.argsort(t1_per.to(torch.float64) - float(t0), descending=True,
                              stable=True)
        self.queue_order = order.to(torch.int32).contiguous()

    def bytes(self) -> int:
        return sum(t.numel() * t.element_size() for t in
                   (self.tape, self.ckpt, self.order, self.lane_of, self.tape_start,
                    self.ckpt_start))

    def forward(self, y0_soa, p_soa, *, t0, t1, rtol, atol, max_steps, t1_per=None,
                stream=None):
        N.call("gsx_tsit5_persist_fwd", self.field, self.n, self.D, self.P, _ptr(self.aux),
               _ptr(y0_soa), _ptr(p_soa), float(t0), float(t1), _ptr(t1_per), float(rtol),
               float(atol), int(max_steps), _ptr(self.y_final), _ptr(self.n_acc),
               _ptr(self.n_rej), _ptr(self.status), self.lanes, _ptr(self.tape), self.lane_cap,
               _ptr(self.ckpt), self.ckpt_cap, self.C, _ptr(self.order), self.traj_cap,
               _ptr(self.lane_count), _ptr(self.lane_of), _ptr(self.tape_start),
               _ptr(self.ckpt_start), _ptr(self.queue), _ptr(self.queue_order), self.flags,
               H._stream_handle(stream))

    def adjoint(self, p_soa, ybar_soa, *, t0, stream=None):
        N.call("gsx_tsit5_persist_adjoint", self.field, self.n, self.D, self.P,
               _ptr(self.aux), _ptr(p_soa), float(t0), _ptr(self.n_acc), self.lanes,
               _ptr(self.tape), _ptr(self.ckpt), self.C, _ptr(self.order),
               _ptr(self.lane_count), _ptr(self.tape_start), _ptr(self.ckpt_start),
               _ptr(ybar_soa), _ptr(self.y0bar), _ptr(self.pbar), self.flags,
               H._stream_handle(stream))
        return self.y0bar, self.pbar

    def run_forward(self, y0_soa, p_soa, *, t0, t1, rtol, atol, max_steps, t1_per=None,
                    max_grow=4):
        """Forward with capacity growth on status 3 (synchronises to read the statuses)."""
        for _ in range(max_grow + 1):
            self.forward(y0_soa, p_soa, t0=t0, t1=t1, rtol=rtol, atol=atol,
                         max_steps=max_steps, t1_per=t1_per)
            if not bool((self.status == 3).any().item()):
                return
            self.lane_cap *= 2
            self.traj_cap *= 2
            self._alloc(None)
        raise RuntimeError("persistent record: tape capacity still exceeded after growth")

    def padded_tape(self) -> torch.Tensor:
        """The reference's zero-padded [n, S] tape, gathered from the lane streams."""
        S = int(self.n_acc.max().item()) if self.n else 0
        s = torch.arange(S, device=self.tape.device, dtype=torch.int64)[None, :]
        pos = self.tape_start.to(torch.int64)[:, None] + s
        lane = self.lane_of.to(torch.int64)[:, None]
        valid = s < self.n_acc.to(torch.int64)[:, None]
        pos = pos.clamp(max=self.lane_cap - 1)
        flat = ((pos // TAPE_BLOCK) * self.lanes + lane) * TAPE_BLOCK + pos % TAPE_BLOCK
        out = self.tape.reshape(-1)[flat.reshape(-1)].reshape(self.n, S)
        return torch.where(valid, out, torch.zeros((), dtype=out.dtype, device=out.device))


class CapturedStep:
    """Forward + cotangent + adjoint captured once in a HIP graph and replayed.

    ``cotangent(y_final_soa, out_soa)`` writes dL/dy_final; it runs inside the graph, so it
    must be a capturable device computation (no host synchronisation, no allocation).
    """

    def __init__(self, eng: PersistentTsit5, y0_soa, p_soa, cotangent, *, t0, t1, rtol, atol,
                 max_steps, t1_per=None):
        self.eng = eng
        kw = dict(t0=t0, t1=t1, rtol=rtol, atol=atol, max_steps=max_steps, t1_per=t1_per)
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):  # warm the allocator and the kernels outside the graph
            eng.forward(y0_soa, p_soa, **kw)
            cotangent(eng.y_final, eng.ybar)
            eng.adjoint(p_soa, eng.ybar, t0=t0)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            eng.forward(y0_soa, p_soa, **kw)
            cotangent(eng.y_final, eng.ybar)
            eng.adjoint(p_soa, eng.ybar, t0=t0)

    def replay(self):
        self.graph.replay()
