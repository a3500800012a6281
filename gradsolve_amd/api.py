"""Public API: ``solve`` and ``grad_closure`` with the reference's signatures
(gradsolve/api.py:291-400 and :792-1030), routed to the native engines.

Differences from the reference, all reported rather than hidden:

* engines run on the GPU (``device`` "cuda"/"gpu"); ``device="cpu"`` raises, because this
  build has no CPU engine;
* the reference's Rodas5P engine (``rodas5p_replay``) is the native ``hip_rodas5p`` lane; its
  other general engines (``tsit5_replay``, ``vern7_replay``, ``fixed_step_*``, ``diffrax``)
  are not part of this build: a request that needs one raises ``ValueError`` naming it;
* closures are differentiated with torch.autograd (a ``torch.autograd.Function`` whose
  backward is the native adjoint kernel) -- JAX is not installed here.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

from . import dispatch as D
from . import fields as F
from .base import Route, SolveResult

_VALID_WRT = "'params', 'y0', or ('y0', 'params')"
_NOT_BUILT = {"vern7_replay", "fixed_step_tsit5", "fixed_step_imex", "diffrax",
              "fused_rosenbrock_backward"}
_RODAS5P_NAMES = ("rodas5p_replay", "hip_rodas5p")


@dataclass(frozen=True)
class EngineSpec:
    """An engine: name, ``supports(problem)``, forward ``solve`` and optional ``reverse``
    closure factory (gradsolve/api.py:48-71)."""

    name: str
    supports: Callable
    solve: Callable
    reverse: Optional[Callable] = None


def _normalize_wrt(wrt) -> tuple:
    if wrt == "params":
        return ("params",)
    if wrt == "y0":
        return ("y0",)
    if isinstance(wrt, (tuple, list, set, frozenset)) and set(wrt) == {"y0", "params"}:
        return ("y0", "params")
    raise ValueError(f"invalid wrt={wrt!r}; expected {_VALID_WRT}")


def _device(device):
    import torch
    d = str(device)
    if d in ("cuda", "gpu"):
        d = "cuda:0"
    if not d.startswith("cuda"):
        raise RuntimeError(
            f"device={device!r}: this build has no CPU engine; the native engines run on the "
            "GPU (pass device='cuda').")
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU is visible (torch.cuda.is_available() is False); the "
                           "native engines need an MI355X")
    return torch.device(d)


def _check_shapes(problem, y0, params):
    y0 = np.asarray(y0) if not hasattr(y0, "shape") else y0
    if len(y0.shape) != 2 or y0.shape[1] != problem.dim:
        raise ValueError(f"y0 must have shape (n, {problem.dim}); got {tuple(y0.shape)}")
    if len(params.shape) != 2 or params.shape[0] != y0.shape[0]:
        raise ValueError(f"params must have shape (n, P) with n={y0.shape[0]}; got "
                         f"{tuple(params.shape)}")


def _native_ok(binding, problem, params) -> bool:
    if binding is None:
        return False
    if binding.param_cols is not None and params.shape[1] <= max(binding.param_cols):
        return False
    if binding.param_cols is None and params.shape[1] != binding.n_params:
        return False
    return True


def _supports_tsit5(problem) -> bool:
    b = F.binding_for(problem)
    return b is not None and not b.stiff and not problem.is_stiff


def _supports_wide(problem) -> bool:
    """The wide lane serves non-stiff fields compiled at run time (any D up to 4096)."""
    b = F.binding_for(problem)
    return (b is not None and b.source is not None and not b.stiff
            and not problem.is_stiff)


def _solve_wide(problem, y0, params, *, rtol, atol, device, **kw) -> SolveResult:
    from . import hip_wide as HW
    return HW.solve(problem, y0, params, rtol=rtol, atol=atol, device=device)


def _supports_rb23(problem) -> bool:
    b = F.binding_for(problem)
    return b is not None and b.stiff and problem.is_stiff


def _solve_tsit5(problem, y0, params, *, rtol, atol, device, strict=False,
                 max_steps=100000) -> SolveResult:
    import torch

    from . import hip_tsit5 as H
    b = F.binding_for(problem)
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
                                    aux=aux, strict=strict)
    y_final = yf.t().cpu().numpy().copy()
    status = st.cpu().numpy()
    y_final[status != 0] = np.nan  # failed rows NaN'd by the driver (warp_ode.py:319)
    return SolveResult(y_final=y_final, accepted_steps=na.cpu().numpy().astype(np.int64),
                       rejected_steps=nr.cpu().numpy().astype(np.int64), solver=D.HIP_TSIT5)


def _solve_rb23(problem, y0, params, *, rtol, atol, device, strict=False,
                **kw) -> SolveResult:
    from . import hip_rosenbrock23 as R
    return R.solve(problem, y0, params, rtol=rtol, atol=atol, device=device, strict=strict)


def _supports_rodas5p(problem) -> bool:
    """A field with an analytic Jacobian (a stiff binding); one that reads t also carries its
    generated df/dt (gradsolve/solvers/rodas5p_step.py:119-146)."""
    b = F.binding_for(problem)
    return b is not None and b.stiff


def _solve_rodas5p(problem, y0, params, *, rtol, atol, device, strict=False,
                   **kw) -> SolveResult:
    from . import hip_rodas5p as R5
    return R5.solve(problem, y0, params, rtol=rtol, atol=atol, device=device, strict=strict)


def _reject_incapable_saveat_dense(name: str) -> None:
    """saveat_dense needs the method's published continuous extension: only the Rodas5P lane
    has one (gradsolve/api.py:425-435)."""
    if name != D.HIP_RODAS5P:
        raise ValueError(
            f"saveat_dense=True is implemented only on the 'rodas5p_replay' lane, which has a "
            f"published continuous extension; engine {name!r} has none, and a generic "
            "interpolant is not offered as a substitute. Use engine='rodas5p_replay', or drop "
            "saveat_dense for the re-step path.")


def _is_kuramoto(b) -> bool:
    from . import _native as N
    return b is not None and b.field_id == N.FIELD_KURAMOTO


def _supports_kuramoto(problem) -> bool:
    return _is_kuramoto(F.binding_for(problem)) and not problem.is_stiff


def _solve_kuramoto(problem, y0, params, *, rtol, atol, device, max_steps=100000,
                    fp32=False, **kw) -> SolveResult:
    import torch

    from . import kuramoto as KU
    b = F.binding_for(problem)
    dev = _device(device)
    y0_soa = torch.as_tensor(np.asarray(y0, np.float64), device=dev).t().contiguous()
    om_soa = torch.as_tensor(np.asarray(params, np.float64), device=dev).t().contiguous()
    K = torch.as_tensor(b.aux_of(problem), device=dev)
    yf, na, nr, st = KU.forward_only(K, y0_soa, om_soa, t0=problem.t0, t1=problem.t1,
                                     rtol=rtol, atol=atol, max_steps=max_steps, fp32=fp32)
    y_final = yf.t().cpu().numpy().copy()
    y_final[st.cpu().numpy() != 0] = np.nan
    return SolveResult(y_final=y_final, accepted_steps=na.cpu().numpy().astype(np.int64),
                       rejected_steps=nr.cpu().numpy().astype(np.int64), solver=D.HIP_KURAMOTO)


ENGINE_REGISTRY = {
    D.HIP_TSIT5: EngineSpec(D.HIP_TSIT5, _supports_tsit5, _solve_tsit5, None),
    D.HIP_REPLAY: EngineSpec(D.HIP_REPLAY, _supports_tsit5, _solve_tsit5, "tsit5"),
    D.HIP_ROSENBROCK23: EngineSpec(D.HIP_ROSENBROCK23, _supports_rb23, _solve_rb23, "rb23"),
    D.HIP_KURAMOTO: EngineSpec(D.HIP_KURAMOTO, _supports_kuramoto, _solve_kuramoto,
                               "kuramoto"),
    D.HIP_RODAS5P: EngineSpec(D.HIP_RODAS5P, _supports_rodas5p, _solve_rodas5p, "rodas5p"),
    D.HIP_TSIT5_WIDE: EngineSpec(D.HIP_TSIT5_WIDE, _supports_wide, _solve_wide, "wide"),
}


def _resolve(engine, problem, need_grad, batch_n, accuracy_target, pre=None):
    reasons = list(pre or [])
    if engine == "auto":
        name = D.choose_engine(problem.dim, problem.is_stiff, need_grad, batch_n=batch_n,
                               accuracy_target=accuracy_target)
        if name == D.RODAS5P_REPLAY:
            name = D.HIP_RODAS5P
            reasons.append("alias:rodas5p_replay")
        elif name == D.DIFFRAX and problem.is_stiff:
            # the reference substitutes its Rodas5P lane when diffrax is not installed
            name = D.HIP_RODAS5P
            reasons.append("diffrax-not-installed")
        elif name == D.DIFFRAX:
            # ... and its general Tsit5 record-and-replay lane for a non-stiff problem
            name = D.HIP_TSIT5_WIDE
            reasons.append("diffrax-not-installed")
        elif name == D.TSIT5_REPLAY:
            name = D.HIP_TSIT5_WIDE
            reasons.append("alias:tsit5_replay")
    elif engine == D.TSIT5_REPLAY:
        # the reference's general lane: the wide lane for fields compiled at run time, the
        # register lane for a built-in field
        b = F.binding_for(problem)
        name = D.HIP_TSIT5_WIDE if (b is None or b.source is not None) else D.HIP_REPLAY
        reasons.append(f"alias:{engine}")
    elif engine in ENGINE_REGISTRY:
        name = engine
    elif engine in D.ALIASES:
        name = D.ALIASES[engine]
        if need_grad and name == D.HIP_TSIT5:
            name = D.HIP_REPLAY
        reasons.append(f"alias:{engine}")
    elif engine in _NOT_BUILT:
        raise ValueError(f"engine {engine!r} is one of the reference's general engines, which "
                         "this build does not include; use engine='auto' with a registered "
                         f"field ({F.registered_names()})")
    else:
        raise ValueError(f"unknown engine {engine!r}; known: {sorted(ENGINE_REGISTRY)} + "
                         f"aliases {sorted(D.ALIASES)}")
    if name in (D.HIP_TSIT5, D.HIP_REPLAY) and _supports_kuramoto(problem):
        name = D.HIP_KURAMOTO
        reasons.append("kuramoto-field: wavefront-cooperative MFMA engine")
    if name in _NOT_BUILT:
        raise ValueError(
            f"problem {problem.name!r} (dim {problem.dim}, stiff={problem.is_stiff}) routes to "
            f"the reference's general engine {name!r}, which this build does not include")
    spec = ENGINE_REGISTRY[name]
    if not spec.supports(problem):
        why = "".join(f" ({r})" for r in reasons if r.startswith("fused translation"))
        raise ValueError(
            f"no native field serves problem {problem.name!r}{why}: the reference would fall "
            "back to its general-RHS engines, which this build does not include. Register the "
            "field with register_jax_field / register_field (or pass fused='auto' with a "
            "traceable f_jax) or use one of " + str(F.registered_names()))
    return name, reasons


def _translate(problem, params, fused, reasons, engine="auto") -> None:
    """``fused='auto'|True``: a problem with no registered field whose ``f_jax`` traces gets
    one compiled on first use, as the reference translates on first use
    (gradsolve/api.py:291-400 with gradsolve/warp/jax_field.py). ``fused=True`` re-raises an
    untraceable right-hand side; ``'auto'`` records why and lets routing report that no
    native field serves the problem (this build has no general-RHS engine to fall back to)."""
    if fused is False or F.binding_for(problem) is not None:
        return
    f = getattr(problem, "f_jax", None)
    if f is None:
        return
    n_params = int(np.asarray(params).shape[1]) if np.ndim(params) == 2 else 0
    try:
        F.register_jax_field(problem.name, f, int(problem.dim), n_params,
                             stiff=bool(problem.is_stiff) or engine in _RODAS5P_NAMES)
        reasons.append("fused: f_jax compiled into the native kernels")
    except F.UnsupportedRHS as exc:
        if fused is True:
            raise
        reasons.append(f"fused translation fell back ({exc.primitive_name}): {exc}")


def solve(problem, y0, params, *, saveat=None, saveat_dense: bool = False,
          engine: str = "auto", fused="auto", rtol: float = 1e-6, atol: float = 1e-9,
          device: str = "cuda", batch_n: Optional[int] = None,
          accuracy_target: Optional[float] = None, precision: str = "float64",
          strict: bool = False) -> SolveResult:
    """Route and run a batched ensemble solve (gradsolve/api.py:291-400).

    Returns host NumPy arrays; rows whose trajectory did not reach ``t1`` are NaN. ``strict``
    selects the contraction-off kernels (rounding like the IEEE-sequential reference loop).
    """
    if fused not in ("auto", True, False):
        raise ValueError(f"fused must be True, False, or 'auto'; got {fused!r}")
    if saveat_dense and saveat is None:
        raise ValueError("saveat_dense=True requires saveat=...; it is a saveat option.")
    if precision not in ("float64", "float32"):
        raise ValueError(f"unknown precision {precision!r}; known: ['float64', 'float32']")
    _check_shapes(problem, y0, params)
    pre = []
    _translate(problem, params, fused, pre, engine)
    name, reasons = _resolve(engine, problem, need_grad=saveat is not None, batch_n=batch_n,
                             accuracy_target=accuracy_target, pre=pre)
    if saveat_dense:
        _reject_incapable_saveat_dense(name)
    b = F.binding_for(problem)
    if not _native_ok(b, problem, params):
        raise ValueError(f"params shape {tuple(params.shape)} does not match the field "
                         f"{b.key!r} ({b.n_params} parameters)")
    if saveat is not None:
        from . import dense
        res = dense.solve_saveat(name, b, problem, y0, params, saveat, rtol=rtol, atol=atol,
                                 device=device, precision=precision, dense=saveat_dense)
    elif precision == "float32":
        from . import fp32
        res = fp32.solve(b, problem, y0, params, rtol=rtol, atol=atol, device=device,
                         name=name)
    else:
        res = ENGINE_REGISTRY[name].solve(problem, y0, params, rtol=rtol, atol=atol,
                                          device=device, strict=strict)
    res.solver = name
    res.route = Route(requested=engine, actual=name,
                      reason="; ".join(reasons) if reasons else None)
    return res


def grad_closure(problem, y0, params, *, wrt="params", saveat=None, saveat_dense: bool = False,
                 precision: str = "float64", engine: str = "auto", fused="auto",
                 rtol: float = 1e-6, atol: float = 1e-9, device: str = "cuda",
                 batch_n: Optional[int] = None, accuracy_target: Optional[float] = None,
                 remat: Optional[bool] = None, ckpt_interval: Optional[int] = None,
                 strict: bool = False):
    """A differentiable closure over a mesh recorded once at ``(y0, params)``
    (gradsolve/api.py:792-1030).

    ``wrt`` selects the closure's arguments: ``params -> y_final``, ``y0 -> y_final`` or
    ``(y0, params) -> y_final``. Call it with torch tensors and differentiate with
    torch.autograd (``.backward()`` / ``torch.autograd.grad``); the backward is the native
    adjoint kernel. ``remat`` maps onto the checkpoint interval (``False``: a checkpoint at
    every accepted step; ``True``/``None``: sqrt(S)-spaced); ``ckpt_interval`` sets it
    directly. The closure carries ``.route``.
    """
    wrt_t = _normalize_wrt(wrt)
    if fused not in ("auto", True, False):
        raise ValueError(f"fused must be True, False, or 'auto'; got {fused!r}")
    if saveat_dense and saveat is None:
        raise ValueError(
            "dense=True is a saveat option (it replaces the per-save re-step with the step's own "
            "continuous extension); pass saveat=..., or drop dense.")
    if precision not in ("float64", "float32"):
        raise ValueError(f"unknown precision {precision!r}; known: ['float64', 'float32']")
    _check_shapes(problem, y0, params)
    pre = []
    _translate(problem, params, fused, pre, engine)
    name, reasons = _resolve(engine, problem, need_grad=True, batch_n=batch_n,
                             accuracy_target=accuracy_target, pre=pre)
    if name == D.HIP_TSIT5:
        name = D.HIP_REPLAY
    if saveat_dense:
        _reject_incapable_saveat_dense(name)
    b = F.binding_for(problem)
    if precision == "float32":
        from . import fp32
        fp32.reject_incapable(name)
    if not _native_ok(b, problem, params):
        raise ValueError(f"params shape {tuple(params.shape)} does not match the field "
                         f"{b.key!r} ({b.n_params} parameters)")
    dev = _device(device)
    if ckpt_interval is None and remat is False:
        ckpt_interval = 1
    aux = None
    if b.aux_of is not None:
        import torch
        aux = torch.as_tensor(b.aux_of(problem), device=dev)
    if saveat is not None:
        from . import dense
        closure = dense.make_saveat_closure(name, b, problem, y0, params, saveat, wrt=wrt_t,
                                            rtol=rtol, atol=atol, device=dev, aux=aux,
                                            precision=precision, ckpt_interval=ckpt_interval,
                                            strict=strict, dense=saveat_dense)
    elif name == D.HIP_KURAMOTO and precision == "float64":
        from .closure import KuramotoClosure
        closure = KuramotoClosure(b, problem, y0, params, wrt=wrt_t, rtol=rtol, atol=atol,
                                  device=dev, aux=aux)
    elif precision == "float32":
        from . import fp32
        closure = fp32.make_closure(name, b, problem, y0, params, wrt=wrt_t, rtol=rtol,
                                    atol=atol, device=dev, aux=aux,
                                    ckpt_interval=ckpt_interval)
    elif name == D.HIP_REPLAY:
        from .closure import make_tsit5_closure
        closure = make_tsit5_closure(b, problem, y0, params, wrt=wrt_t, rtol=rtol, atol=atol,
                                     device=dev, ckpt_interval=ckpt_interval, strict=strict,
                                     aux=aux)
    elif name == D.HIP_TSIT5_WIDE:
        from . import hip_wide as HW
        closure = HW.make_closure(b, problem, y0, params, wrt=wrt_t, rtol=rtol, atol=atol,
                                  device=dev, ckpt_interval=ckpt_interval, aux=aux)
    elif name == D.HIP_RODAS5P:
        from . import hip_rodas5p as R5
        closure = R5.make_closure(b, problem, y0, params, wrt=wrt_t, rtol=rtol, atol=atol,
                                  device=dev, ckpt_interval=ckpt_interval, strict=strict,
                                  aux=aux)
    else:
        from . import hip_rosenbrock23 as R
        closure = R.make_closure(b, problem, y0, params, wrt=wrt_t, rtol=rtol, atol=atol,
                                 device=dev, ckpt_interval=ckpt_interval, strict=strict)
    closure.route = Route(requested=engine, actual=name,
                          reason="; ".join(reasons) if reasons else None)
    return closure
