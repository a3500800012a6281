"""Registry of problems that route to the native fields, and ``register_jax_field``.

A binding maps a problem (by ``name``) to a native field of the C ABI, and the user's
``params (n, P_user)`` columns to the field's parameters. It replaces the reference's field
routing (``warp_ode._field_for`` / ``_param_of``, gradsolve/warp/warp_ode.py:463-549, and
``warp_rosenbrock._field_for``) and the registry ``register_jax_field`` fills
(gradsolve/warp/jax_field.py:516-564).
"""
from __future__ import annotations

import threading
from dataclasses import dataclass
from typing import Callable, Dict, Optional, Tuple

import numpy as np

from . import _native as N


class UnsupportedRHS(ValueError):
    """The right-hand side cannot drive a native field (gradsolve/warp/jax_field.py:36-49)."""

    def __init__(self, primitive_name: str, msg: str):
        super().__init__(msg)
        self.primitive_name = primitive_name


@dataclass(frozen=True)
class FieldBinding:
    key: str
    field_id: int
    dim: int
    n_params: int                       # parameters the native field takes
    stiff: bool
    param_cols: Optional[Tuple[int, ...]]  # user columns feeding the field (None: all)
    numpy_f: Callable                    # (y (dim,), p (n_params,), aux) -> dy, for probing
    aux_of: Optional[Callable] = None   # problem -> shared field data (K for Kuramoto)
    source: Optional[str] = None        # generated C++ of a field compiled at run time
    time_dependent: bool = False

    def native_params(self, params):
        """(n, P_user) -> (n, n_params) columns the field reads (numpy or torch)."""
        if self.param_cols is not None:
            return params[:, list(self.param_cols)]
        if params.shape[1] < self.n_params:  # a field without parameters takes one unused
            pad = self.n_params - params.shape[1]
            if isinstance(params, np.ndarray):
                return np.concatenate([params, np.zeros((params.shape[0], pad))], axis=1)
            import torch
            return torch.cat([params, params.new_zeros((params.shape[0], pad))], dim=1)
        return params


def _lorenz1(y, p, aux):
    return np.array([10.0 * (y[1] - y[0]), p[0] * y[0] - y[1] - y[0] * y[2],
                     y[0] * y[1] - (8.0 / 3.0) * y[2]])


def _lorenz3(y, p, aux):
    return np.array([p[0] * (y[1] - y[0]), p[2] * y[0] - y[1] - y[0] * y[2],
                     y[0] * y[1] - p[1] * y[2]])


def _vdp(y, p, aux):
    return np.array([y[1], p[0] * ((1.0 - y[0] * y[0]) * y[1] - y[0])])


def _robertson(y, p, aux):
    return np.array([-p[0] * y[0] + p[2] * y[1] * y[2],
                     p[0] * y[0] - p[2] * y[1] * y[2] - p[1] * y[1] * y[1],
                     p[1] * y[1] * y[1]])


def _decay(y, p, aux):
    return -p[0] * y


def _kuramoto(y, p, aux):
    D = y.shape[0]
    s, c = np.sin(y), np.cos(y)
    return p + (1.0 / D) * ((aux @ s) * c - (aux @ c) * s)


_BUILTIN: Dict[str, FieldBinding] = {
    "diffeqgpu_lorenz": FieldBinding("diffeqgpu_lorenz", N.FIELD_LORENZ, 3, 1, False, (0,),
                                     _lorenz1),
    "lorenz": FieldBinding("lorenz", N.FIELD_LORENZ3, 3, 3, False, (0, 1, 2), _lorenz3),
    "vdp": FieldBinding("vdp", N.FIELD_VDP, 2, 1, False, (0,), _vdp),
    "robertson": FieldBinding("robertson", N.FIELD_ROBERTSON, 3, 3, True, (0, 1, 2),
                              _robertson),
}
for _d in (1, 2, 3):
    _BUILTIN[f"linear_decay_{_d}"] = FieldBinding(f"linear_decay_{_d}", N.FIELD_LINEAR_DECAY,
                                                  _d, 1, False, (0,), _decay)

_REGISTERED: Dict[str, FieldBinding] = {}
_lock = threading.Lock()


def _kuramoto_binding(problem) -> Optional[FieldBinding]:
    K = getattr(problem, "K", None)
    if K is None or int(problem.dim) != 32 or problem.is_stiff:
        return None
    return FieldBinding(problem.name, N.FIELD_KURAMOTO, 32, 32, False, None, _kuramoto,
                        aux_of=lambda pr: np.ascontiguousarray(pr.K, dtype=np.float64))


def binding_for(problem) -> Optional[FieldBinding]:
    """The native binding serving ``problem``, or None."""
    name = getattr(problem, "name", None)
    b = _REGISTERED.get(name) or _BUILTIN.get(name)
    if b is not None:
        if int(problem.dim) != b.dim:
            return None
        return b
    if isinstance(name, str) and name.startswith("kuramoto"):
        return _kuramoto_binding(problem)
    return None


def registered_names():
    return sorted(set(_BUILTIN) | set(_REGISTERED))


def _generate(name: str, f, dim: int, n_params: int, stiff: bool) -> FieldBinding:
    """Trace ``f``, emit its C++ field with generated derivatives, compile it with hiprtc."""
    from . import fieldgen as FG
    tr = FG.trace(f, int(dim), int(n_params))
    src = FG.emit_field(tr, stiff=bool(stiff))
    flags = (N.FIELD_STIFF if stiff else 0) | (N.FIELD_TIME_DEPENDENT if tr.time_dependent
                                               else 0)
    P = max(int(n_params), 1)
    fid = N.field_register(src, int(dim), P, flags)

    def numpy_f(y, p, aux, _f=f, _n=int(n_params)):
        return np.asarray(_f(0.0, np.asarray(y, np.float64), np.asarray(p, np.float64)[:_n]),
                          dtype=np.float64)

    return FieldBinding(name, fid, int(dim), P, bool(stiff), None, numpy_f, source=src,
                        time_dependent=tr.time_dependent)


def register_field(name: str, f, dim: int, n_params: int, *, stiff: bool = False):
    """Compile the right-hand side ``f(t, y, p)`` into the fused kernels and route problems
    named ``name`` to it.

    ``f`` is traced once on symbolic NumPy arrays (``gradsolve_amd.fieldgen``); its C++ form,
    with reverse-mode generated VJP (and, for ``stiff=True``, the analytic Jacobian and its
    VJP), is compiled for gfx950 through ``gsx_field_register``. This is what the reference's
    ``register_jax_field`` does with a jaxpr (gradsolve/warp/jax_field.py:516-564). Raises
    ``UnsupportedRHS`` for right-hand sides outside the traceable subset and ``ValueError``
    (with the compiler log) if the emitted source does not compile. Idempotent per name.
    """
    with _lock:
        if name in _REGISTERED:
            return _REGISTERED[name]
        b = _generate(name, f, dim, n_params, stiff)
        _REGISTERED[name] = b
        return b


def register_jax_field(name: str, f_jax, dim: int, n_params: int, *, stiff: bool = False):
    """Route problems named ``name`` to a native field computing ``f_jax``.

    The reference translates the jaxpr of ``f_jax`` into a fused-kernel field
    (gradsolve/__init__.py:31-42, gradsolve/warp/jax_field.py:516-564). JAX is absent here, so
    ``f_jax`` is called on NumPy values: first it is matched numerically against the
    hand-written native fields of the same dimension, parameter count and stiffness (random
    states and parameters); on a match the name routes to that field. Otherwise it is traced
    and compiled as a new field (``register_field``). A right-hand side outside the traceable
    subset raises ``UnsupportedRHS``. Idempotent: an already-registered name is left untouched.
    """
    with _lock:
        if name in _REGISTERED:
            return _REGISTERED[name]
        rng = np.random.default_rng(1234)
        cands = [b for b in _BUILTIN.values()
                 if b.dim == int(dim) and b.stiff == bool(stiff)
                 and (b.param_cols is None or max(b.param_cols) < int(n_params))]
        for b in cands:
            ok = True
            for _ in range(6):
                y = rng.uniform(-2.0, 2.0, int(dim))
                p = rng.uniform(0.5, 3.0, int(n_params))
                want = b.numpy_f(y, np.asarray(b.native_params(p[None, :])[0]), None)
                try:
                    got = np.asarray(f_jax(0.0, y, p), dtype=np.float64)
                except Exception as exc:  # an f_jax that NumPy cannot evaluate
                    raise UnsupportedRHS(type(exc).__name__,
                                         f"f_jax for {name!r} cannot be evaluated on NumPy "
                                         f"arrays: {exc}") from exc
                if got.shape != want.shape or not np.allclose(got, want, rtol=1e-12,
                                                              atol=1e-12):
                    ok = False
                    break
            if ok:
                nb = FieldBinding(name, b.field_id, b.dim, b.n_params, b.stiff, b.param_cols,
                                  b.numpy_f, b.aux_of)
                _REGISTERED[name] = nb
                return nb
        nb = _generate(name, f_jax, dim, n_params, stiff)
        _REGISTERED[name] = nb
        return nb
