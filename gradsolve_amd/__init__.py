"""gradsolve_amd: the ensemble ODE record-and-replay hot path of gradsolve on MI355X.

Public surface (gradsolve/__init__.py:11-45): ``solve``, ``grad_closure``,
``register_jax_field``, the ``Problem`` contract and the ``Route`` / ``SolveResult`` types.
The compute path is the gfx950 HIP library behind the C ABI in include/gradsolve_amd.h.
"""
from .api import ENGINE_REGISTRY, EngineSpec, grad_closure, solve
from .base import Backend, Problem, Route, SolveResult
from .dispatch import DECISION_MAP, choose_engine, choose_remat
from .fields import UnsupportedRHS, register_field, register_jax_field
from . import problems

__version__ = "0.1.0"

__all__ = [
    "solve", "grad_closure", "register_jax_field", "register_field", "Problem", "Route",
    "SolveResult", "Backend", "ENGINE_REGISTRY", "EngineSpec", "choose_engine", "choose_remat",
    "DECISION_MAP", "UnsupportedRHS", "problems", "__version__",
]
