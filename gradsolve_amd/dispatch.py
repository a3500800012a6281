"""Pure router: ``(dim, stiff, need_grad) -> engine name`` (gradsolve/dispatch.py:178-293).

The codomain names this build's native engines:

* ``hip_tsit5``        -- fused adaptive Tsit5 forward (replaces ``cuda_tsit5``/``warp_ode``);
* ``hip_replay``       -- fused Tsit5 forward + accepted-h record, then the hand-written
                         adjoint kernel (replaces ``warp_replay``, the record-and-replay
                         gradient lane);
* ``hip_rosenbrock23`` -- fused Rosenbrock23 forward + record and its adjoint (replaces
                         ``warp_rosenbrock``/``cuda_rosenbrock23``);
* ``hip_kuramoto``     -- Kuramoto-32 forward + record, saveat and adjoint, 16 trajectories
                         per wavefront with the coupling on MFMA (the reference runs this
                         problem on its general ``tsit5_replay`` lane);
* ``hip_rodas5p``      -- Rodas5P forward + record, adjoint and dense output (replaces the
                         reference's ``rodas5p_replay``, which the router names for stiff
                         gradients above the low-NVAR band and which explicit opt-in reaches).

High-dimensional non-stiff problems route to the reference's general engines
(``tsit5_replay`` for gradients, ``diffrax`` forward); the API runs both cells on

* ``hip_tsit5_wide``   -- the wide-state Tsit5 lane for fields compiled at run time (any D up
                         to 4096; state and stages in an HBM workspace), recording the reason
                         on ``route.reason`` (``alias:tsit5_replay`` / ``diffrax-not-installed``).
"""
from __future__ import annotations

from typing import List, TypedDict

#: dim <= NVAR_CEILING is "low NVAR", the fused kernels' design point (gradsolve/dispatch.py:108).
NVAR_CEILING = 64

#: Routing flags, kept with the reference's meaning; each has a keyword override in
#: ``choose_engine`` so both routings stay testable without mutating module state.
STIFF_FUSED_ENABLED = True
HIP_TSIT5_ENABLED = True
HIP_ROSENBROCK23_ENABLED = True

HIP_TSIT5 = "hip_tsit5"
HIP_REPLAY = "hip_replay"
HIP_ROSENBROCK23 = "hip_rosenbrock23"
#: The wavefront-cooperative Kuramoto engine (D = 32, coupling on MFMA). Not a router
#: output: the API substitutes it for the router's Tsit5 engines when the problem's native
#: field is the Kuramoto kernel, and records why on ``route.reason``.
HIP_KURAMOTO = "hip_kuramoto"
HIP_RODAS5P = "hip_rodas5p"
#: The wide-state Tsit5 lane; not a router output: the API substitutes it for the router's
#: general non-stiff engines (tsit5_replay, diffrax) above NVAR_CEILING.
HIP_TSIT5_WIDE = "hip_tsit5_wide"
TSIT5_REPLAY = "tsit5_replay"
RODAS5P_REPLAY = "rodas5p_replay"
DIFFRAX = "diffrax"

#: Reference engine names accepted as aliases of the native engine that replaces them.
ALIASES = {
    "warp_ode": HIP_TSIT5,
    "cuda_tsit5": HIP_TSIT5,
    "warp_replay": HIP_REPLAY,
    "warp_rosenbrock": HIP_ROSENBROCK23,
    "cuda_rosenbrock23": HIP_ROSENBROCK23,
    "rodas5p_replay": HIP_RODAS5P,
}


def choose_engine(dim: int, stiff: bool, need_grad: bool, *, batch_n=None,
                  accuracy_target=None, stiff_fused_enabled=None, hip_tsit5_enabled=None,
                  hip_rosenbrock23_enabled=None) -> str:
    """Engine for the (dim, stiff, need_grad) cell. Total, pure, never raises.

    ``batch_n`` and ``accuracy_target`` are accepted routing hints that do not branch yet
    (as in gradsolve/dispatch.py:195-202).
    """
    del batch_n, accuracy_target
    if stiff_fused_enabled is None:
        stiff_fused_enabled = STIFF_FUSED_ENABLED
    if hip_tsit5_enabled is None:
        hip_tsit5_enabled = HIP_TSIT5_ENABLED
    if hip_rosenbrock23_enabled is None:
        hip_rosenbrock23_enabled = HIP_ROSENBROCK23_ENABLED
    low = dim <= NVAR_CEILING
    if stiff:
        if low and stiff_fused_enabled and hip_rosenbrock23_enabled:
            return HIP_ROSENBROCK23
        return RODAS5P_REPLAY if need_grad else DIFFRAX
    if not low:
        return TSIT5_REPLAY if need_grad else DIFFRAX
    if need_grad:
        return HIP_REPLAY
    return HIP_TSIT5 if hip_tsit5_enabled else HIP_REPLAY


def choose_remat(dim: int, stiff: bool, *, batch_n=None) -> bool:
    """The reference's rematerialisation policy (gradsolve/dispatch.py:268-293), kept for
    API compatibility. The native adjoint always re-integrates from checkpoints; ``remat``
    maps onto its checkpoint interval (True: sqrt(S) spacing; False: every state)."""
    del batch_n
    if not stiff:
        return True
    return dim >= 16


class DecisionRow(TypedDict):
    nvar: str
    stiff: bool
    need_grad: bool
    dim: int
    engine: str
    gated_engine: "str | None"
    evidence: str


DECISION_MAP: List[DecisionRow] = [
    {"nvar": "low", "stiff": False, "need_grad": False, "dim": 3, "engine": HIP_TSIT5,
     "gated_engine": HIP_REPLAY,
     "evidence": "The fused Tsit5 forward holds one trajectory per lane with its stages in "
                 "registers and writes no tape when no gradient is requested."},
    {"nvar": "low", "stiff": False, "need_grad": True, "dim": 3, "engine": HIP_REPLAY,
     "gated_engine": None,
     "evidence": "Forward with the accepted-h tape and state checkpoints, then the "
                 "hand-written adjoint kernel: the exact discrete adjoint of the frozen-step "
                 "replay without a per-step scan on the host."},
    {"nvar": "high", "stiff": False, "need_grad": False, "dim": 1000, "engine": DIFFRAX,
     "gated_engine": TSIT5_REPLAY,
     "evidence": "Above NVAR_CEILING a per-lane state no longer fits the register file; the "
                 "API runs this cell on hip_tsit5_wide (state and stages in an HBM "
                 "workspace, one trajectory per lane) for fields compiled at run time."},
    {"nvar": "high", "stiff": False, "need_grad": True, "dim": 1000, "engine": TSIT5_REPLAY,
     "gated_engine": None,
     "evidence": "General-RHS record-and-replay of the reference; the API runs it on "
                 "hip_tsit5_wide: the wide forward with the accepted-h record and "
                 "checkpoints, then its hand-written adjoint."},
    {"nvar": "low", "stiff": True, "need_grad": False, "dim": 3, "engine": HIP_ROSENBROCK23,
     "gated_engine": DIFFRAX,
     "evidence": "Fused Rosenbrock23: per-lane 3x3 W = I - d h J factored once per step in "
                 "registers and reused by the three stage solves."},
    {"nvar": "low", "stiff": True, "need_grad": True, "dim": 3, "engine": HIP_ROSENBROCK23,
     "gated_engine": RODAS5P_REPLAY,
     "evidence": "Rosenbrock23 record plus its hand-written adjoint through W and J(y, p)."},
    {"nvar": "high", "stiff": True, "need_grad": False, "dim": 1000, "engine": DIFFRAX,
     "gated_engine": RODAS5P_REPLAY,
     "evidence": "Out of the fused kernels' register budget; reference general engines."},
    {"nvar": "high", "stiff": True, "need_grad": True, "dim": 1000, "engine": RODAS5P_REPLAY,
     "gated_engine": None,
     "evidence": "General-RHS stiff record-and-replay of the reference; not part of this "
                 "build."},
]
