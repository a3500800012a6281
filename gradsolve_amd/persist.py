"""Host side of the persistent work-queue Tsit5 engine (``gsx_tsit5_persist_fwd`` /
``gsx_tsit5_persist_adjoint``): buffer sizing, the forward+adjoint pair, and capture of the
pair in a HIP graph.

Layout. rec[rec_cap][lanes][R]: each lane's stream of segment records, record = the state at
the start of a checkpoint segment followed by the step sizes of its C accepted steps
(R = gsx_persist_record_len: rows + C rounded up to 8 doubles); order[traj_cap][lanes].

Capacities. Each lane appends to its own record and trajectory-list streams, so the per-lane
capacities follow from the expected accepted steps per lane: a trajectory of S steps takes
ceil(S / C) records, so ``rec_cap = margin * (mean_steps / C + 1) * n / lanes`` plus the
longest trajectory's records (what one lane can add after crossing the average). Lanes do not
share the work evenly: a lane takes trajectories as fast as its wavefront steps, and on
2^22 uneven Lorenz the busiest lane used 1.56-1.6x the mean records, hence margin 2.
``estimate_steps`` measures the step-count distribution on a random subsample with the
per-trajectory kernel. The trajectory list is 4 bytes per entry and sized generously: with a
longest-first queue the last trajectories every lane draws are the shortest, so the count per
lane spreads more than the steps per lane. A trajectory that still does not fit gets status 3,
and ``run_forward`` doubles the capacities and runs again, as the reference's device recorder
doubles its tape (gradsolve/solvers/record_jax.py:192-199).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _native as N
from . import hip_tsit5 as H


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
                 margin: float = 2.0, device=None, aux: Optional[torch.Tensor] = None):
        self.field, self.D, self.P, self.n = field, D, P, n
        self.C = int(ckpt_interval)
        self.fp32 = bool(fp32)
        self.aux = aux
        self.device = device
        self.waves_per_simd = int(waves_per_simd)
        self.lanes = int(N.call("gsx_persist_lanes", self.waves_per_simd))
        self.rows = int(N.call("gsx_ckpt_rows", field, D))
        self.R = int(N.call("gsx_persist_record_len", field, D, self.C))
        L, C = self.lanes, self.C
        per_lane = n / L
        self.rec_cap = int(math.ceil(margin * (est.mean_steps / C + 1.0) * per_lane)) + \
            2 * (est.max_steps // C + 1) + 8
        self.traj_cap = int(math.ceil(4.0 * per_lane)) + 64
        self.queue_order: Optional[torch.Tensor] = None  # see plan_longest_first
        self.frozen = False  # set while a captured HIP graph holds the buffers
        self._alloc()

    def _alloc(self):
        if self.frozen:
            raise RuntimeError(
                "persistent engine: its buffers are held by a captured HIP graph and cannot be "
                "reallocated (capacity growth); build a new engine and capture again")
        L, n, dev = self.lanes, self.n, self.device
        f64 = dict(dtype=torch.float64, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        self.rec = torch.empty((self.rec_cap, L, self.R), **f64)
        self.order = torch.empty((self.traj_cap, L), **i32)
        self.lane_count = torch.empty(L, **i32)
        self.lane_of = torch.empty(n, **i32)
        self.rec_start = torch.empty(n, **i32)
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
                   (self.rec, self.order, self.lane_of, self.rec_start))

    def forward(self, y0_soa, p_soa, *, t0, t1, rtol, atol, max_steps, t1_per=None,
                stream=None):
        N.call("gsx_tsit5_persist_fwd", self.field, self.n, self.D, self.P, _ptr(self.aux),
               _ptr(y0_soa), _ptr(p_soa), float(t0), float(t1), _ptr(t1_per), float(rtol),
               float(atol), int(max_steps), _ptr(self.y_final), _ptr(self.n_acc),
               _ptr(self.n_rej), _ptr(self.status), self.lanes, _ptr(self.rec), self.rec_cap,
               self.C, _ptr(self.order), self.traj_cap, _ptr(self.lane_count),
               _ptr(self.lane_of), _ptr(self.rec_start), _ptr(self.queue),
               _ptr(self.queue_order), self.flags, H._stream_handle(stream))

    def adjoint(self, p_soa, ybar_soa, *, t0, stream=None):
        N.call("gsx_tsit5_persist_adjoint", self.field, self.n, self.D, self.P,
               _ptr(self.aux), _ptr(p_soa), float(t0), _ptr(self.n_acc), self.lanes,
               _ptr(self.rec), self.C, _ptr(self.order), _ptr(self.lane_count),
               _ptr(self.rec_start), _ptr(ybar_soa), _ptr(self.y0bar), _ptr(self.pbar),
               self.flags, H._stream_handle(stream))
        return self.y0bar, self.pbar

    def run_forward(self, y0_soa, p_soa, *, t0, t1, rtol, atol, max_steps, t1_per=None,
                    max_grow=4):
        """Forward with capacity growth on status 3 (synchronises to read the statuses)."""
        for _ in range(max_grow + 1):
            self.forward(y0_soa, p_soa, t0=t0, t1=t1, rtol=rtol, atol=atol,
                         max_steps=max_steps, t1_per=t1_per)
            if not bool((self.status == 3).any().item()):
                return
            self.rec_cap *= 2
            self.traj_cap *= 2
            self._alloc()
        raise RuntimeError("persistent record: capacity still exceeded after growth")

    def padded_tape(self, rows: Optional[torch.Tensor] = None) -> torch.Tensor:
        """The reference's zero-padded [n, S] tape, gathered from the lane records; with
        ``rows`` (trajectory indices), only those trajectories, padded to their own longest."""
        n_acc = self.n_acc.to(torch.int64)
        lane_of, rec_start = self.lane_of, self.rec_start
        if rows is not None:
            rows = rows.to(device=self.rec.device, dtype=torch.int64)
            n_acc, lane_of, rec_start = n_acc[rows], lane_of[rows], rec_start[rows]
        m = n_acc.numel()
        S = int(n_acc.max().item()) if m else 0
        s = torch.arange(S, device=self.rec.device, dtype=torch.int64)[None, :]
        valid = s < n_acc[:, None]
        r = (rec_start.to(torch.int64)[:, None] + s // self.C).clamp(max=self.rec_cap - 1)
        lane = lane_of.to(torch.int64)[:, None]
        flat = (r * self.lanes + lane) * self.R + self.rows + s % self.C
        flat = torch.where(valid, flat, torch.zeros((), dtype=flat.dtype, device=flat.device))
        out = self.rec.reshape(-1)[flat.reshape(-1)].reshape(m, S)
        return torch.where(valid, out, torch.zeros((), dtype=out.dtype, device=out.device))


class CapturedStep:
    """Forward + cotangent + adjoint captured once in a HIP graph and replayed.

    ``cotangent(y_final_soa, out_soa)`` writes dL/dy_final; it runs inside the graph, so it
    must be a capturable device computation (no host synchronisation, no allocation).

    Capacities are grown (``run_forward``) on the capture inputs before capture; from then on
    the engine's buffers are frozen, since the graph holds their addresses. A replay cannot
    grow them: a trajectory that no longer fits (inputs changed in place between replays) ends
    with status 3 and no gradient, and one that exhausts ``max_steps`` with status 1, so
    ``replay`` checks the statuses (one synchronisation) and raises on either unless
    ``check=False``; ``check()`` does the same later.
    """

    def __init__(self, eng: PersistentTsit5, y0_soa, p_soa, cotangent, *, t0, t1, rtol, atol,
                 max_steps, t1_per=None):
        self.eng = eng
        kw = dict(t0=t0, t1=t1, rtol=rtol, atol=atol, max_steps=max_steps, t1_per=t1_per)
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):  # size the buffers, warm the kernels outside the graph
            eng.run_forward(y0_soa, p_soa, **kw)
            cotangent(eng.y_final, eng.ybar)
            eng.adjoint(p_soa, eng.ybar, t0=t0)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.check()
        eng.frozen = True
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            eng.forward(y0_soa, p_soa, **kw)
            cotangent(eng.y_final, eng.ybar)
            eng.adjoint(p_soa, eng.ybar, t0=t0)

    def check(self) -> None:
        """Raise if a trajectory of the last run did not finish (status 1 or 3)."""
        st = self.eng.status
        full = int((st == 3).sum().item())
        exhausted = int((st == 1).sum().item())
        if full:
            raise RuntimeError(
                f"persistent record: {full} trajectories did not fit the captured capacities "
                "(status 3); their gradients were not computed. Capture again on the new "
                "inputs (capacities are grown before capture)")
        if exhausted:
            raise RuntimeError(f"{exhausted} trajectories exhausted max_steps before t1")

    def replay(self, check: bool = True) -> None:
        self.graph.replay()
        if check:
            self.check()
