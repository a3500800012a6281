"""Device-side driver of the fused HIP Tsit5 engine: record, replay and adjoint.

This module owns the device buffers of one recording (``Tsit5Record``) and calls the C ABI
(``include/gradsolve_amd.h``). It mirrors the reference's record step
(``gradsolve/warp/warp_replay.py:121-183`` driving ``warp_ode._launch``,
``gradsolve/warp/warp_ode.py:215-327``) with three differences of layout, not of semantics:

* device arrays are structure-of-arrays ``[D, n]`` (the layout of the reference's CUDA FFI
  handler, ``gradsolve/cuda/kernels/tsit5_template.cu.in:146-151``);
* the tape is step-major ``[cap, n]`` and only rows ``< n_acc[j]`` of column ``j`` are
  meaningful; ``padded_tape()`` produces the reference's zero-padded ``[n, S]``;
* the forward also writes state checkpoints every ``ckpt_interval`` accepted steps, which
  the adjoint kernel re-integrates from instead of storing stages.

Capacity overflow (status 3) doubles the tape capacity and records again, as
``gradsolve/solvers/record_jax.py:192-199`` does for its device recorder.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import torch
import torch.utils.dlpack

from . import _native as N

DEFAULT_TAPE_CAP = 256
MAX_CKPT_INTERVAL = 32


def _stream_handle(stream: Optional[torch.cuda.Stream]) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else t.data_ptr()


def flags_for(strict: bool = False, fp32: bool = False) -> int:
    """Kernel-variant flags: contraction-off FP64 (strict) or FP32 stages with compensated
    state (fp32); the two are exclusive."""
    if strict and fp32:
        raise ValueError("strict and fp32 are exclusive kernel variants")
    return N.STRICT if strict else (N.FP32 if fp32 else 0)


def max_segment_interval(rows: int) -> int:
    """Largest checkpoint interval whose segment adjoint fits its LDS ring. The segment
    kernels are built for at most 1, 4, 8, 16 or 32 steps per segment and hold that many
    checkpoint rows per lane: CMAX * rows * 64 lanes * 8 B <= 64 KiB (C = 2 runs the pair
    adjoint, which needs no LDS)."""
    best = 2
    for cmax in (4, 8, 16, 32):
        if cmax * rows * 64 * 8 <= 65536:
            best = cmax
    return best


def default_ckpt_interval(expected_steps: int) -> int:
    """sqrt(S)-spaced checkpoints: the interval that balances checkpoint storage against
    the re-integration a segment needs."""
    c = int(round(math.sqrt(max(1, expected_steps))))
    return max(1, min(MAX_CKPT_INTERVAL, c))


@dataclass
class Tsit5Record:
    """Device buffers of one adaptive Tsit5 forward + record."""

    field: int
    n: int
    D: int
    P: int
    t0: float
    t1: float
    y0: torch.Tensor        # [D, n] f64
    p: torch.Tensor         # [P, n] f64
    t1_per: Optional[torch.Tensor]
    aux: Optional[torch.Tensor]
    y_final: torch.Tensor   # [D, n]
    n_acc: torch.Tensor     # [n] int32
    n_rej: torch.Tensor     # [n] int32
    status: torch.Tensor    # [n] int32
    tape: torch.Tensor      # [cap, n] f64
    cap: int
    ckpt: torch.Tensor      # [slots, DCK, n]
    ckpt_interval: int

    def max_steps_taken(self) -> int:
        return int(self.n_acc.max().item()) if self.n else 0

    def trimmed_tape(self) -> torch.Tensor:
        """The reference's trimmed step-major record ``[S, n]`` (S = max n_acc, zero past each
        trajectory's n_acc), as a zero-copy view of this record's tape: hand it to JAX with
        ``jax.dlpack.from_dlpack(t)`` or to any DLPack consumer via ``t.__dlpack__()``
        (gradsolve/warp/warp_ode.py:164-212)."""
        S = self.max_steps_taken()
        N.call("gsx_tape_zero_tail", self.n, _ptr(self.tape), self.cap, _ptr(self.n_acc), S,
               _stream_handle(None))
        return self.tape[:S]

    def padded_tape(self, S: Optional[int] = None) -> torch.Tensor:
        """The reference's zero-padded trajectory-major tape ``[n, S]``, S = max n_acc."""
        if S is None:
            S = self.max_steps_taken()
        out = torch.empty((self.n, S), dtype=torch.float64, device=self.tape.device)
        N.call("gsx_tape_to_padded", self.n, _ptr(self.tape), self.cap, _ptr(self.n_acc), S,
               _ptr(out), _stream_handle(None))
        return out


def _alloc_ckpt(field: int, D: int, n: int, cap: int, C: int, device) -> torch.Tensor:
    rows = N.call("gsx_ckpt_rows", field, D)
    slots = N.call("gsx_ckpt_slots", cap, C)
    return torch.empty((slots, rows, n), dtype=torch.float64, device=device)


def record(field: int, y0_soa: torch.Tensor, p_soa: torch.Tensor, *, t0: float, t1: float,
           rtol: float, atol: float, max_steps: int = 4096,
           t1_per: Optional[torch.Tensor] = None, aux: Optional[torch.Tensor] = None,
           tape_cap: Optional[int] = None, ckpt_interval: Optional[int] = None,
           stream: Optional[torch.cuda.Stream] = None, grow: bool = True,
           strict: bool = False, entry: str = "gsx_tsit5_fwd",
           fp32: bool = False) -> Tsit5Record:
    """Adaptive forward with the accepted-h record and checkpoints (device-resident).

    ``y0_soa [D, n]`` and ``p_soa [P, n]`` are float64 device tensors. Returns a record whose
    ``status`` the caller inspects (0 reached t1, 1 exhausted max_steps, 3 tape full).
    ``entry`` selects the method's forward entry point (``gsx_tsit5_fwd`` or ``gsx_rb23_fwd``;
    both share the tape and checkpoint contract).
    """
    D, n = y0_soa.shape
    P = p_soa.shape[0]
    dev = y0_soa.device
    cap = int(min(max_steps, tape_cap if tape_cap else DEFAULT_TAPE_CAP))
    cap = max(cap, 1)
    rows = int(N.call("gsx_ckpt_rows", field, D))
    while True:
        C = ckpt_interval if ckpt_interval else default_ckpt_interval(cap // 2)
        C = max(1, min(MAX_CKPT_INTERVAL, int(C)))
        if not ckpt_interval:
            C = min(C, max(1, max_segment_interval(rows)))
        y_final = torch.empty((D, n), dtype=torch.float64, device=dev)
        n_acc = torch.empty(n, dtype=torch.int32, device=dev)
        n_rej = torch.empty(n, dtype=torch.int32, device=dev)
        status = torch.empty(n, dtype=torch.int32, device=dev)
        tape = torch.empty((cap, n), dtype=torch.float64, device=dev)
        ckpt = _alloc_ckpt(field, D, n, cap, C, dev)
        N.call(entry, field, n, D, P, _ptr(aux), _ptr(y0_soa), _ptr(p_soa),
               float(t0), float(t1), _ptr(t1_per), float(rtol), float(atol), int(max_steps),
               _ptr(y_final), _ptr(n_acc), _ptr(n_rej), _ptr(status), _ptr(tape), cap,
               _ptr(ckpt), C, flags_for(strict, fp32), _stream_handle(stream))
        rec = Tsit5Record(field, n, D, P, float(t0), float(t1), y0_soa, p_soa, t1_per, aux,
                          y_final, n_acc, n_rej, status, tape, cap, ckpt, C)
        if not grow or cap >= max_steps or n == 0:
            return rec
        if not bool((status == 3).any().item()):
            return rec
        cap = int(min(2 * cap, max_steps))


def forward_only(field: int, y0_soa: torch.Tensor, p_soa: torch.Tensor, *, t0: float,
                 t1: float, rtol: float, atol: float, max_steps: int = 4096,
                 t1_per: Optional[torch.Tensor] = None, aux: Optional[torch.Tensor] = None,
                 stream: Optional[torch.cuda.Stream] = None, strict: bool = False,
                 entry: str = "gsx_tsit5_fwd", fp32: bool = False):
    """The forward-only lane (no tape, no checkpoints): (y_final, n_acc, n_rej, status)."""
    D, n = y0_soa.shape
    P = p_soa.shape[0]
    dev = y0_soa.device
    y_final = torch.empty((D, n), dtype=torch.float64, device=dev)
    n_acc = torch.empty(n, dtype=torch.int32, device=dev)
    n_rej = torch.empty(n, dtype=torch.int32, device=dev)
    status = torch.empty(n, dtype=torch.int32, device=dev)
    N.call(entry, field, n, D, P, _ptr(aux), _ptr(y0_soa), _ptr(p_soa), float(t0),
           float(t1), _ptr(t1_per), float(rtol), float(atol), int(max_steps), _ptr(y_final),
           _ptr(n_acc), _ptr(n_rej), _ptr(status), None, 0, None, 1,
           flags_for(strict, fp32), _stream_handle(stream))
    return y_final, n_acc, n_rej, status


def adjoint(rec: Tsit5Record, ybar_soa: torch.Tensor,
            stream: Optional[torch.cuda.Stream] = None, strict: bool = False,
            fp32: bool = False):
    """dL/dy0 [D, n] and dL/dp [P, n] for the cotangent ``ybar_soa = dL/dy_final [D, n]``."""
    ybar_soa = ybar_soa.to(dtype=torch.float64).contiguous()
    y0bar = torch.empty((rec.D, rec.n), dtype=torch.float64, device=rec.y0.device)
    pbar = torch.empty((rec.P, rec.n), dtype=torch.float64, device=rec.y0.device)
    N.call("gsx_tsit5_adjoint", rec.field, rec.n, rec.D, rec.P, _ptr(rec.aux), _ptr(rec.y0),
           _ptr(rec.p), rec.t0, _ptr(rec.tape), rec.cap, _ptr(rec.n_acc), _ptr(rec.ckpt),
           rec.ckpt_interval, _ptr(ybar_soa), _ptr(y0bar), _ptr(pbar),
           flags_for(strict, fp32), _stream_handle(stream))
    return y0bar, pbar


def record_from_trimmed(field: int, y0_soa: torch.Tensor, p_soa: torch.Tensor, tape_sn,
                        *, t0: float, t1: float, ckpt_interval: int = 8,
                        aux: Optional[torch.Tensor] = None, strict: bool = False,
                        entry: str = "gsx_tsit5_replay") -> Tsit5Record:
    """A record built from a trimmed ``[S, n]`` step-major tape made elsewhere (the
    reference's fused recorder hands one over DLPack; any DLPack producer or array works):
    n_acc from its nonzero rows, checkpoints and the final state from one replay, ready for
    the adjoint."""
    if not isinstance(tape_sn, torch.Tensor):
        tape_sn = torch.utils.dlpack.from_dlpack(tape_sn) if hasattr(tape_sn, "__dlpack__") \
            else torch.as_tensor(tape_sn)
    D, n = y0_soa.shape
    P = p_soa.shape[0]
    dev = y0_soa.device
    S = int(tape_sn.shape[0])
    if tape_sn.dim() != 2 or int(tape_sn.shape[1]) != n:
        raise ValueError(f"tape must be [S, n={n}]; got {tuple(tape_sn.shape)}")
    cap = max(S, 1)
    tape = torch.zeros((cap, n), dtype=torch.float64, device=dev)
    tape[:S] = tape_sn.to(device=dev, dtype=torch.float64)
    n_acc = torch.empty(n, dtype=torch.int32, device=dev)
    N.call("gsx_tape_count", n, _ptr(tape), S, _ptr(n_acc), _stream_handle(None))
    C = max(1, min(MAX_CKPT_INTERVAL, int(ckpt_interval)))
    y_final = torch.empty((D, n), dtype=torch.float64, device=dev)
    ckpt = _alloc_ckpt(field, D, n, cap, C, dev)
    N.call(entry, field, n, D, P, _ptr(aux), _ptr(y0_soa), _ptr(p_soa), float(t0), _ptr(tape),
           _ptr(n_acc), _ptr(y_final), _ptr(ckpt), C, flags_for(strict),
           _stream_handle(None))
    zeros = torch.zeros(n, dtype=torch.int32, device=dev)
    return Tsit5Record(field, n, D, P, float(t0), float(t1), y0_soa, p_soa, None, aux,
                       y_final, n_acc, zeros.clone(), zeros, tape, cap, ckpt, C)


def record_from_padded(field: int, y0_soa: torch.Tensor, p_soa: torch.Tensor,
                       padded: torch.Tensor, *, t0: float, t1: float,
                       aux: Optional[torch.Tensor] = None):
    """Import a reference-format tape ``[n, S]`` into the step-major layout; returns
    ``(tape [cap, n], n_acc [n])`` with cap = S."""
    n, S = padded.shape
    padded = padded.to(dtype=torch.float64).contiguous()
    tape = torch.empty((max(S, 1), n), dtype=torch.float64, device=padded.device)
    n_acc = torch.empty(n, dtype=torch.int32, device=padded.device)
    N.call("gsx_tape_from_padded", n, _ptr(padded), S, _ptr(tape), max(S, 1), _ptr(n_acc),
           _stream_handle(None))
    return tape, n_acc


class Workspace:
    """Preallocated caller-owned buffers for repeated forward+adjoint calls on a fixed
    ensemble shape (no allocation between calls; every call is stream-ordered)."""

    def __init__(self, field: int, D: int, P: int, n: int, cap: int, ckpt_interval: int,
                 device, aux: Optional[torch.Tensor] = None, checkpoints: bool = True,
                 strict: bool = False):
        self.field, self.D, self.P, self.n = field, D, P, n
        self.cap, self.C, self.aux = int(cap), int(ckpt_interval), aux
        self.flags = N.STRICT if strict else 0
        f64 = dict(dtype=torch.float64, device=device)
        i32 = dict(dtype=torch.int32, device=device)
        self.y_final = torch.empty((D, n), **f64)
        self.n_acc = torch.empty(n, **i32)
        self.n_rej = torch.empty(n, **i32)
        self.status = torch.empty(n, **i32)
        self.tape = torch.empty((self.cap, n), **f64)
        self.ckpt = (_alloc_ckpt(field, D, n, self.cap, self.C, device) if checkpoints
                     else None)
        self.ybar = torch.empty((D, n), **f64)
        self.y0bar = torch.empty((D, n), **f64)
        self.pbar = torch.empty((P, n), **f64)

    def forward(self, y0_soa, p_soa, *, t0, t1, rtol, atol, max_steps, t1_per=None,
                record=True, checkpoints=True, stream=None):
        ck = _ptr(self.ckpt) if (record and checkpoints) else None
        N.call("gsx_tsit5_fwd", self.field, self.n, self.D, self.P, _ptr(self.aux),
               _ptr(y0_soa), _ptr(p_soa), float(t0), float(t1), _ptr(t1_per), float(rtol),
               float(atol), int(max_steps), _ptr(self.y_final), _ptr(self.n_acc),
               _ptr(self.n_rej), _ptr(self.status), _ptr(self.tape) if record else None,
               self.cap if record else 0, ck, self.C, self.flags, _stream_handle(stream))

    def adjoint(self, p_soa, ybar_soa, *, t0, stream=None):
        N.call("gsx_tsit5_adjoint", self.field, self.n, self.D, self.P, _ptr(self.aux), None,
               _ptr(p_soa), float(t0), _ptr(self.tape), self.cap, _ptr(self.n_acc),
               _ptr(self.ckpt), self.C, _ptr(ybar_soa), _ptr(self.y0bar), _ptr(self.pbar),
               self.flags, _stream_handle(stream))
        return self.y0bar, self.pbar


@dataclass
class SaveatResult:
    ys: torch.Tensor          # [nsave, D, n]
    save_step: torch.Tensor   # [nsave, n] int32
    save_t: torch.Tensor      # [nsave, n]
    states: Optional[torch.Tensor]  # [nsave, D, n]


_DENSE_ENTRIES = {"tsit5": ("gsx_tsit5_saveat", "gsx_tsit5_adjoint_saveat"),
                  "rb23": ("gsx_rb23_saveat", "gsx_rb23_adjoint_saveat"),
                  "rodas5p": ("gsx_rodas5p_saveat", "gsx_rodas5p_adjoint_saveat"),
                  "rodas5p_dense": ("gsx_rodas5p_saveat_dense",
                                    "gsx_rodas5p_adjoint_saveat_dense")}


def _dense_entry(method: str, which: int) -> str:
    try:
        return _DENSE_ENTRIES[method][which]
    except KeyError:
        raise ValueError(f"unknown method {method!r}; known: {sorted(_DENSE_ENTRIES)}") from None


def saveat(rec: Tsit5Record, ts: torch.Tensor, *, y0_soa: Optional[torch.Tensor] = None,
           p_soa: Optional[torch.Tensor] = None, y_final: Optional[torch.Tensor] = None,
           keep_states: bool = True, strict: bool = False, method: str = "tsit5",
           stream: Optional[torch.cuda.Stream] = None) -> SaveatResult:
    """Dense output on the record's mesh (gsx_tsit5_saveat, or gsx_rb23_saveat for a
    Rosenbrock23 record with ``method="rb23"``): the states at the save times ``ts``
    (ascending, float64 device tensor), replayed from ``y0_soa`` / ``p_soa`` (default: the
    recording inputs, whose final-time saves then copy the record's final state)."""
    if y0_soa is None and p_soa is None and y_final is None:
        y_final = rec.y_final
    y0_soa = rec.y0 if y0_soa is None else y0_soa
    p_soa = rec.p if p_soa is None else p_soa
    k = int(ts.shape[0])
    dev = rec.y0.device
    ys = torch.empty((k, rec.D, rec.n), dtype=torch.float64, device=dev)
    step = torch.empty((k, rec.n), dtype=torch.int32, device=dev)
    st_t = torch.empty((k, rec.n), dtype=torch.float64, device=dev)
    states = torch.empty((k, rec.D, rec.n), dtype=torch.float64, device=dev) \
        if keep_states else None
    entry = _dense_entry(method, 0)
    args = [rec.field, rec.n, rec.D, rec.P, _ptr(rec.aux), _ptr(y0_soa), _ptr(p_soa), rec.t0,
            _ptr(rec.tape), _ptr(rec.n_acc), _ptr(y_final), _ptr(ts), k, _ptr(ys), _ptr(step),
            _ptr(st_t)]
    if method == "rodas5p_dense":  # the interpolant needs no bracket states
        states = None
    else:
        args.append(_ptr(states))
    N.call(entry, *args, flags_for(strict), _stream_handle(stream))
    return SaveatResult(ys, step, st_t, states)


def adjoint_saveat(rec: Tsit5Record, ts: torch.Tensor, sv: SaveatResult,
                   ysbar: torch.Tensor, ybar: Optional[torch.Tensor] = None, *,
                   ckpt: Optional[torch.Tensor] = None, p_soa: Optional[torch.Tensor] = None,
                   strict: bool = False, method: str = "tsit5",
                   stream: Optional[torch.cuda.Stream] = None):
    """dL/dy0 [D, n] and dL/dp [P, n] for save cotangents ``ysbar [nsave, D, n]`` and an
    optional final-state cotangent ``ybar [D, n]`` (gsx_tsit5_adjoint_saveat, or
    gsx_rb23_adjoint_saveat with ``method="rb23"``)."""
    dev = rec.y0.device
    k = int(ts.shape[0])
    p_soa = rec.p if p_soa is None else p_soa
    ckpt = rec.ckpt if ckpt is None else ckpt
    ysbar = ysbar.to(dtype=torch.float64).contiguous()
    if ybar is not None:
        ybar = ybar.to(dtype=torch.float64).contiguous()
    y0bar = torch.empty((rec.D, rec.n), dtype=torch.float64, device=dev)
    pbar = torch.empty((rec.P, rec.n), dtype=torch.float64, device=dev)
    if method == "rodas5p_dense":  # one reverse sweep, no re-step VJPs
        N.call(_dense_entry(method, 1), rec.field, rec.n, rec.D, rec.P, _ptr(rec.aux),
               _ptr(rec.y0), _ptr(p_soa), rec.t0, _ptr(rec.tape), rec.cap, _ptr(rec.n_acc),
               _ptr(ckpt), rec.ckpt_interval, _ptr(ts), k, _ptr(sv.save_step),
               _ptr(sv.save_t), _ptr(ysbar), _ptr(ybar), _ptr(y0bar), _ptr(pbar),
               flags_for(strict), _stream_handle(stream))
        return y0bar, pbar
    mu = torch.empty((k, rec.D, rec.n), dtype=torch.float64, device=dev)
    N.call(_dense_entry(method, 1), rec.field, rec.n, rec.D, rec.P, _ptr(rec.aux),
           _ptr(rec.y0), _ptr(p_soa), rec.t0, _ptr(rec.tape), rec.cap, _ptr(rec.n_acc),
           _ptr(ckpt), rec.ckpt_interval, _ptr(ts), k, _ptr(sv.save_step), _ptr(sv.save_t),
           _ptr(sv.states), _ptr(ysbar), _ptr(ybar), _ptr(mu), _ptr(y0bar), _ptr(pbar),
           flags_for(strict), _stream_handle(stream))
    return y0bar, pbar
