"""Record-once gradient closures on the native engines, differentiable with torch.autograd.

``ReplayClosure`` records the accepted-step mesh once, at the ``(y0, params)`` it is built
with, exactly as the reference's closure factories do (gradsolve/warp/warp_replay.py:186-237,
gradsolve/solvers/tsit5_replay.py:294-334). Calling it replays that frozen mesh from the
given inputs; the backward pass is the native adjoint kernel. At the recording inputs the
replay is the recording itself (its checkpoints are reused); at other inputs the replay
kernel re-integrates the frozen mesh and writes fresh checkpoints.

The gradient semantics are the reference's: step sizes are data, so the gradient is the
exact discrete adjoint of the frozen-step integration, valid near the recording point.
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from . import _native as N
from . import hip_tsit5 as H


def _as_device(x, dev) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=torch.float64)
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=dev)


class _Lane:
    """Native calls of one record-and-replay lane (Tsit5 or Rosenbrock23)."""

    FWD, REPLAY, ADJ = "gsx_tsit5_fwd", "gsx_tsit5_replay", "gsx_tsit5_adjoint"
    METHOD = "tsit5"  # the dense-output entry points' method (hip_tsit5.saveat)

    def __init__(self, binding, aux, strict, fp32=False):
        self.b, self.aux, self.strict, self.fp32 = binding, aux, strict, fp32
        self.flags = H.flags_for(strict, fp32)

    def record(self, y0_soa, p_soa, *, t0, t1, rtol, atol, max_steps, ckpt_interval):
        return H.record(self.b.field_id, y0_soa, p_soa, t0=t0, t1=t1, rtol=rtol, atol=atol,
                        max_steps=max_steps, aux=self.aux, ckpt_interval=ckpt_interval,
                        strict=self.strict, entry=self.FWD, fp32=self.fp32)

    def replay(self, rec, y0_soa, p_soa):
        y_final = torch.empty_like(rec.y_final)
        ckpt = torch.empty_like(rec.ckpt)
        N.call(self.REPLAY, rec.field, rec.n, rec.D, rec.P, H._ptr(self.aux),
               y0_soa.data_ptr(), p_soa.data_ptr(), rec.t0, rec.tape.data_ptr(),
               rec.n_acc.data_ptr(), y_final.data_ptr(), ckpt.data_ptr(), rec.ckpt_interval,
               self.flags, H._stream_handle(None))
        return y_final, ckpt

    def adjoint(self, rec, ckpt, p_soa, ybar_soa):
        y0bar = torch.empty((rec.D, rec.n), dtype=torch.float64, device=ybar_soa.device)
        pbar = torch.empty((rec.P, rec.n), dtype=torch.float64, device=ybar_soa.device)
        N.call(self.ADJ, rec.field, rec.n, rec.D, rec.P, H._ptr(self.aux), None,
               p_soa.data_ptr(), rec.t0, rec.tape.data_ptr(), rec.cap, rec.n_acc.data_ptr(),
               ckpt.data_ptr(), rec.ckpt_interval, ybar_soa.data_ptr(), y0bar.data_ptr(),
               pbar.data_ptr(), self.flags, H._stream_handle(None))
        return y0bar, pbar


class _Tsit5Lane(_Lane):
    pass


class _Rb23Lane(_Lane):
    FWD, REPLAY, ADJ = "gsx_rb23_fwd", "gsx_rb23_replay", "gsx_rb23_adjoint"
    METHOD = "rb23"


class _Rodas5PLane(_Lane):
    FWD, REPLAY, ADJ = "gsx_rodas5p_fwd", "gsx_rodas5p_replay", "gsx_rodas5p_adjoint"

    def __init__(self, binding, aux, strict, dense=False):
        super().__init__(binding, aux, strict)
        # saveat by the continuous extension (saveat_dense) or by re-stepping
        self.METHOD = "rodas5p_dense" if dense else "rodas5p"


class _ReplayFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, z, p_user, state):
        y0_soa = z.detach().t().contiguous()
        p_soa = state.binding.native_params(p_user.detach()).t().contiguous()
        rec = state.rec
        if state.is_recording_point(y0_soa, p_soa):
            y_final, ckpt = rec.y_final, rec.ckpt
        else:
            y_final, ckpt = state.lane.replay(rec, y0_soa, p_soa)
        ctx.state = state
        ctx.ckpt = ckpt
        ctx.p_soa = p_soa
        ctx.p_cols = p_user.shape[1]
        return y_final.t().contiguous()

    @staticmethod
    def backward(ctx, gy):
        st = ctx.state
        ybar = gy.to(torch.float64).t().contiguous()
        y0bar, pbar = st.lane.adjoint(st.rec, ctx.ckpt, ctx.p_soa, ybar)
        g_p = None
        if ctx.needs_input_grad[1]:
            g_p = torch.zeros((st.rec.n, ctx.p_cols), dtype=torch.float64, device=gy.device)
            cols = st.binding.param_cols
            if cols is None:
                g_p = pbar.t().contiguous()
            else:
                g_p[:, list(cols)] = pbar.t()
        g_z = y0bar.t().contiguous() if ctx.needs_input_grad[0] else None
        return g_z, g_p, None


class ReplayClosure:
    """Callable closure over a frozen mesh; see the module docstring.

    Arguments follow ``wrt``: ``closure(params)``, ``closure(y0)`` or ``closure(y0, params)``,
    each shaped like the corresponding input. torch tensors (any device; computed on the
    recording device) give a differentiable torch result; NumPy inputs give a NumPy result.
    """

    def __init__(self, lane, binding, problem, y0, params, *, wrt, rtol, atol, device,
                 ckpt_interval=None, max_steps=100000):
        self.lane, self.binding, self.problem, self.wrt = lane, binding, problem, wrt
        self.device = device
        y0_t = _as_device(y0, device)
        p_t = _as_device(params, device)
        if y0_t.ndim == 1:
            y0_t = y0_t[None, :]
        if p_t.ndim == 1:
            p_t = p_t[None, :]
        self._y0 = y0_t
        self._params = p_t
        self._y0_soa = y0_t.t().contiguous()
        self._p_soa = binding.native_params(p_t).t().contiguous()
        self.rec = lane.record(self._y0_soa, self._p_soa, t0=problem.t0, t1=problem.t1,
                               rtol=rtol, atol=atol, max_steps=max_steps,
                               ckpt_interval=ckpt_interval)
        st = self.rec.status
        bad = torch.nonzero(st != 0).flatten()
        if bad.numel():
            i = int(bad[0].item())
            code = int(st[i].item())
            if code == 2:
                raise RuntimeError(
                    f"record: trajectory {i} reached the step-size underflow floor before "
                    f"t1={problem.t1}; the recorded dt sequence is degenerate and a frozen "
                    "replay of it would be silently wrong.")
            raise RuntimeError(
                f"record: {bad.numel()} trajectories (first: {i}) exhausted max_steps="
                f"{max_steps} before t1={problem.t1} (status {code}); the recorded dt "
                "sequence is incomplete and a frozen replay of it would be silently wrong. "
                "Increase max_steps, loosen the tol, or shorten the horizon.")

    def is_recording_point(self, y0_soa, p_soa) -> bool:
        return (y0_soa.shape == self._y0_soa.shape and p_soa.shape == self._p_soa.shape
                and bool(torch.equal(y0_soa, self._y0_soa))
                and bool(torch.equal(p_soa, self._p_soa)))

    # -- public ------------------------------------------------------------------------
    @property
    def n_acc(self) -> np.ndarray:
        return self.rec.n_acc.cpu().numpy().astype(np.int64)

    @property
    def n_rej(self) -> np.ndarray:
        return self.rec.n_rej.cpu().numpy().astype(np.int64)

    def dts_padded(self) -> np.ndarray:
        """The recorded mesh in the reference's zero-padded ``[n, S]`` form."""
        return self.rec.padded_tape().cpu().numpy()

    def __call__(self, *args):
        if self.wrt == ("params",):
            (p,), z = args, None
        elif self.wrt == ("y0",):
            (z,), p = args, None
        else:
            z, p = args
        numpy_in = not any(isinstance(a, torch.Tensor) for a in args)
        zt = self._y0 if z is None else self._coerce(z)
        pt = self._params if p is None else self._coerce(p)
        out = _ReplayFn.apply(zt, pt, self)
        if numpy_in:
            return out.detach().cpu().numpy()
        return out

    def _coerce(self, x):
        if isinstance(x, torch.Tensor):
            if x.device != self.device or x.dtype != torch.float64:
                x = x.to(device=self.device, dtype=torch.float64)
            return x
        return _as_device(x, self.device)


def make_tsit5_closure(binding, problem, y0, params, *, wrt, rtol, atol, device,
                       ckpt_interval=None, strict=False, aux=None,
                       max_steps=100000) -> ReplayClosure:
    lane = _Tsit5Lane(binding, aux, strict)
    return ReplayClosure(lane, binding, problem, y0, params, wrt=wrt, rtol=rtol, atol=atol,
                         device=device, ckpt_interval=ckpt_interval, max_steps=max_steps)


def make_tsit5_f32_closure(binding, problem, y0, params, *, wrt, rtol, atol, device,
                           ckpt_interval=None, aux=None, max_steps=100000) -> ReplayClosure:
    """The FP32 lane: FP32 stages with a compensated state, FP64 time, error norm and
    controller; the tape stays FP64 and the adjoint runs the FP32 step VJP with compensated
    accumulators."""
    lane = _Tsit5Lane(binding, aux, False, fp32=True)
    return ReplayClosure(lane, binding, problem, y0, params, wrt=wrt, rtol=rtol, atol=atol,
                         device=device, ckpt_interval=ckpt_interval, max_steps=max_steps)


def make_rb23_closure(binding, problem, y0, params, *, wrt, rtol, atol, device,
                      ckpt_interval=None, strict=False, aux=None, max_steps=50000,
                      fp32=False) -> ReplayClosure:
    """The Rosenbrock23 lane (``fp32``: FP32 stages on a compensated state, FP64 W, LU,
    solves, error norm, controller and tape)."""
    lane = _Rb23Lane(binding, aux, strict, fp32=fp32)
    return ReplayClosure(lane, binding, problem, y0, params, wrt=wrt, rtol=rtol, atol=atol,
                         device=device, ckpt_interval=ckpt_interval, max_steps=max_steps)


class _SaveatFn(torch.autograd.Function):
    """The saveat series of a record-once closure (Tsit5 or Rosenbrock23 lane): forward replays
    the frozen mesh from the given inputs and re-steps to the save times; backward runs the
    re-steps' VJPs and the adjoint sweep with the save cotangents injected at their bracketing
    steps."""

    @staticmethod
    def forward(ctx, z, p_user, state):
        y0_soa = z.detach().t().contiguous()
        p_soa = state.binding.native_params(p_user.detach()).t().contiguous()
        rec = state.rec
        if state.is_recording_point(y0_soa, p_soa):
            y_final, ckpt = rec.y_final, rec.ckpt
        else:
            y_final, ckpt = state.lane.replay(rec, y0_soa, p_soa)
        sv = H.saveat(rec, state.ts, y0_soa=y0_soa, p_soa=p_soa, y_final=y_final,
                      strict=state.lane.strict, method=state.lane.METHOD)
        ctx.state, ctx.ckpt, ctx.p_soa, ctx.sv = state, ckpt, p_soa, sv
        ctx.p_cols = p_user.shape[1]
        return sv.ys.permute(2, 0, 1).contiguous()  # (n, k, D)

    @staticmethod
    def backward(ctx, g):
        st = ctx.state
        ysbar = g.to(torch.float64).permute(1, 2, 0).contiguous()  # [k, D, n]
        y0bar, pbar = H.adjoint_saveat(st.rec, st.ts, ctx.sv, ysbar, ckpt=ctx.ckpt,
                                       p_soa=ctx.p_soa, strict=st.lane.strict,
                                       method=st.lane.METHOD)
        g_p = None
        if ctx.needs_input_grad[1]:
            cols = st.binding.param_cols
            if cols is None:
                g_p = pbar.t()[:, :ctx.p_cols].contiguous()
            else:
                g_p = torch.zeros((st.rec.n, ctx.p_cols), dtype=torch.float64,
                                  device=g.device)
                g_p[:, list(cols)] = pbar.t()
        g_z = y0bar.t().contiguous() if ctx.needs_input_grad[0] else None
        return g_z, g_p, None


class Tsit5SaveatClosure(ReplayClosure):
    """Record-once closure returning the series at the save times ``(n, k, D)``
    (gradsolve/api.py:750-753 with gradsolve/solvers/dense.py:318-335); with a Rosenbrock23
    lane it is the stiff series of make_rosenbrock_replay_kernel(saveat=...)
    (gradsolve/warp/warp_replay.py:364-382)."""

    def __init__(self, lane, binding, problem, y0, params, *, wrt, rtol, atol, device, ts,
                 ckpt_interval=None, max_steps=100000):
        super().__init__(lane, binding, problem, y0, params, wrt=wrt, rtol=rtol, atol=atol,
                         device=device, ckpt_interval=ckpt_interval, max_steps=max_steps)
        self.ts = torch.as_tensor(np.asarray(ts, np.float64), device=device)

    def __call__(self, *args):
        if self.wrt == ("params",):
            (p,), z = args, None
        elif self.wrt == ("y0",):
            (z,), p = args, None
        else:
            z, p = args
        numpy_in = not any(isinstance(a, torch.Tensor) for a in args)
        zt = self._y0 if z is None else self._coerce(z)
        pt = self._params if p is None else self._coerce(p)
        out = _SaveatFn.apply(zt, pt, self)
        return out.detach().cpu().numpy() if numpy_in else out


def make_tsit5_saveat_closure(binding, problem, y0, params, *, wrt, rtol, atol, device, ts,
                              ckpt_interval=None, strict=False, aux=None,
                              max_steps=100000) -> Tsit5SaveatClosure:
    lane = _Tsit5Lane(binding, aux, strict)
    return Tsit5SaveatClosure(lane, binding, problem, y0, params, wrt=wrt, rtol=rtol,
                              atol=atol, device=device, ts=ts, ckpt_interval=ckpt_interval,
                              max_steps=max_steps)


def make_rodas5p_closure(binding, problem, y0, params, *, wrt, rtol, atol, device,
                         ckpt_interval=None, strict=False, aux=None,
                         max_steps=50000) -> ReplayClosure:
    lane = _Rodas5PLane(binding, aux, strict)
    return ReplayClosure(lane, binding, problem, y0, params, wrt=wrt, rtol=rtol, atol=atol,
                         device=device, ckpt_interval=ckpt_interval, max_steps=max_steps)


def make_rodas5p_saveat_closure(binding, problem, y0, params, *, wrt, rtol, atol, device, ts,
                                dense=False, ckpt_interval=None, strict=False, aux=None,
                                max_steps=50000) -> "Tsit5SaveatClosure":
    """The Rodas5P series closure: saves by re-stepping, or with ``dense`` by the method's
    order-4 continuous extension (the reference's ``saveat_dense=True``,
    gradsolve/solvers/rodas5p_replay.py:179-228)."""
    lane = _Rodas5PLane(binding, aux, strict, dense=dense)
    return Tsit5SaveatClosure(lane, binding, problem, y0, params, wrt=wrt, rtol=rtol,
                              atol=atol, device=device, ts=ts, ckpt_interval=ckpt_interval,
                              max_steps=max_steps)


def make_rb23_saveat_closure(binding, problem, y0, params, *, wrt, rtol, atol, device, ts,
                             ckpt_interval=None, strict=False, aux=None,
                             max_steps=50000) -> Tsit5SaveatClosure:
    lane = _Rb23Lane(binding, aux, strict)
    return Tsit5SaveatClosure(lane, binding, problem, y0, params, wrt=wrt, rtol=rtol,
                              atol=atol, device=device, ts=ts, ckpt_interval=ckpt_interval,
                              max_steps=max_steps)


class _KuramotoFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, z, om_user, state):
        from . import kuramoto as KU
        rec = state.rec
        y0_soa = z.detach().t().contiguous()
        om_soa = om_user.detach().t().contiguous()
        if state.is_recording_point(y0_soa, om_soa):
            y_final, states = rec.y_final, rec.states
        else:
            y_final, states = KU.replay(rec, y0_soa, om_soa)
        ctx.state, ctx.states, ctx.om_soa = state, states, om_soa
        if state.dense:
            return KU.saveat(rec, y0_soa, om_soa, y_final, states)
        return y_final.t().contiguous()

    @staticmethod
    def backward(ctx, g):
        from . import kuramoto as KU
        st = ctx.state
        g = g.to(torch.float64)
        if st.dense:
            y0bar, ombar = KU.adjoint(st.rec, ysbar=g.contiguous(), om_soa=ctx.om_soa,
                                      states=ctx.states)
        else:
            y0bar, ombar = KU.adjoint(st.rec, yfbar=g.t().contiguous(), om_soa=ctx.om_soa,
                                      states=ctx.states)
        g_z = y0bar.t().contiguous() if ctx.needs_input_grad[0] else None
        g_p = ombar.t().contiguous() if ctx.needs_input_grad[1] else None
        return g_z, g_p, None


class KuramotoClosure(ReplayClosure):
    """Record-once closure of the Kuramoto engine: ``closure(omega)``, ``closure(theta0)`` or
    ``closure(theta0, omega)`` returns the final state ``(n, 32)``, or with save times the
    series ``(n, k, 32)`` (gradsolve/api.py:750-753); the backward is the native adjoint with
    the save-time cotangents."""

    def __init__(self, binding, problem, y0, params, *, wrt, rtol, atol, device, aux,
                 ts=None, max_steps=100000, fp32=False):
        from . import kuramoto as KU
        self.binding, self.problem, self.wrt, self.device = binding, problem, wrt, device
        y0_t = _as_device(y0, device)
        p_t = _as_device(params, device)
        self._y0, self._params = y0_t, p_t
        self._y0_soa = y0_t.t().contiguous()
        self._p_soa = p_t.t().contiguous()
        self.dense = ts is not None
        self.ts = None if ts is None else torch.as_tensor(ts, dtype=torch.float64,
                                                          device=device)
        self.rec = KU.record(aux, self._y0_soa, self._p_soa, t0=problem.t0, t1=problem.t1,
                             rtol=rtol, atol=atol, ts=self.ts, max_steps=max_steps,
                             fp32=fp32)
        st = self.rec.status
        bad = torch.nonzero(st != 0).flatten()
        if bad.numel():
            i = int(bad[0].item())
            raise RuntimeError(
                f"record: {bad.numel()} trajectories (first: {i}) exhausted max_steps="
                f"{max_steps} before t1={problem.t1} (status {int(st[i].item())}); the "
                "recorded dt sequence is incomplete and a frozen replay of it would be "
                "silently wrong. Increase max_steps, loosen the tol, or shorten the horizon.")

    def __call__(self, *args):
        if self.wrt == ("params",):
            (p,), z = args, None
        elif self.wrt == ("y0",):
            (z,), p = args, None
        else:
            z, p = args
        numpy_in = not any(isinstance(a, torch.Tensor) for a in args)
        zt = self._y0 if z is None else self._coerce(z)
        pt = self._params if p is None else self._coerce(p)
        out = _KuramotoFn.apply(zt, pt, self)
        return out.detach().cpu().numpy() if numpy_in else out
