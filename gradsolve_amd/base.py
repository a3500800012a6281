"""The problem contract and result types (gradsolve/base.py:16-216).

A problem is any object exposing six members -- ``name``, ``dim``, ``t0``, ``t1``,
``is_stiff`` and ``f_jax(t, y, params)`` -- written per trajectory (``y`` is one ``(dim,)``
state, ``params`` one ``(P,)`` parameter vector). Nothing else on a problem object is read.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Optional, Protocol

import numpy as np


class Problem(Protocol):
    """Duck-typed problem contract (gradsolve/base.py:16-68)."""

    name: str
    dim: int
    t0: float
    t1: float

    @property
    def is_stiff(self) -> bool: ...

    def f_jax(self, t: float, y: Any, params: Any) -> Any: ...


@dataclass(frozen=True)
class Route:
    """Where a ``solve``/``grad_closure`` request went (gradsolve/base.py:71-95): the
    requested engine, the engine that ran, and why they differ (``None`` if they do not;
    several reasons are joined with ``'; '``)."""

    requested: str
    actual: str
    reason: Optional[str] = None


@dataclass
class SolveResult:
    """Result of one ensemble solve (gradsolve/base.py:98-152), host NumPy arrays:
    ``y_final (n, dim)``, per-trajectory ``accepted_steps``/``rejected_steps`` ``(n,)`` int64,
    the engine name in ``solver``, optional dense output ``y_saved (n, k, dim)`` at
    ``ts_saved (k,)``, and the routing record ``route``."""

    y_final: np.ndarray
    accepted_steps: np.ndarray = field(default_factory=lambda: np.empty(0))
    rejected_steps: np.ndarray = field(default_factory=lambda: np.empty(0))
    solver: str = ""
    y_saved: Optional[np.ndarray] = None
    ts_saved: Optional[np.ndarray] = None
    route: Optional[Route] = None


class Backend(Protocol):
    """Engine contract (gradsolve/base.py:155-216): ``name``, ``supports(problem)`` and a
    blocking ``solve(problem, y0, params, *, rtol, atol, device) -> SolveResult``."""

    name: str

    def supports(self, problem: Problem) -> bool: ...

    def solve(self, problem: Problem, y0: np.ndarray, params: np.ndarray, *, rtol: float,
              atol: float, device: str) -> SolveResult: ...
