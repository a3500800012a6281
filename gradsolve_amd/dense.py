"""``saveat``: dense output on a recorded mesh (gradsolve/solvers/dense.py, gradsolve/api.py:
438-535 for ``solve`` and :750-753 for ``grad_closure``).

Semantics as the reference's re-step path: the forward records, for every save time, the
accepted step that brackets it (``t_s < ts[q] <= t_s + h_s``); the saved state is one Tsit5
step of size ``ts[q] - t_s`` from that step's start state. Times at or before ``t0`` give
``y0``; times at or beyond the final time give ``y_final`` (so the last save equals
``y_final`` bit for bit when ``ts[-1] = t1``). The gradient differentiates the re-steps as well
as the mesh replay, with the step sizes as data.

This build serves saveat on the Kuramoto engine (the problem the reference's dense lane is
benchmarked on), on the Tsit5 lanes of every other non-stiff field, built-in or compiled at
run time (``gsx_tsit5_saveat`` / ``gsx_tsit5_adjoint_saveat``), and on the stiff Rosenbrock23
lane, whose re-step is one ``rosenbrock23_step`` (``gsx_rb23_saveat`` /
``gsx_rb23_adjoint_saveat``; the reference's ``replay_solve_rosenbrock_saveat_jax``,
gradsolve/warp/warp_replay.py:78-84), and on the Rodas5P lane, by re-stepping or, with
``saveat_dense=True``, by the method's order-4 continuous extension (``gsx_rodas5p_saveat``,
``gsx_rodas5p_saveat_dense`` and their adjoints; gradsolve/solvers/dense.py:166-227). FP32
dense output outside the Kuramoto engine raises ``ValueError`` naming what is built.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as N
from . import dispatch as D
from . import kuramoto as KU
from .base import SolveResult


def _check(name, b, precision, dense=False) -> str:
    if b.field_id == N.FIELD_KURAMOTO and name == D.HIP_KURAMOTO:
        return "kuramoto"
    if name in (D.HIP_TSIT5, D.HIP_REPLAY) and not b.stiff:
        if precision == "float32":
            raise ValueError(
                f"saveat with precision='float32' is built for the Kuramoto engine only; field "
                f"{b.key!r} serves dense output in float64")
        return "tsit5"
    if name == D.HIP_RODAS5P:
        if precision == "float32":
            raise ValueError("rodas5p_replay is float64 throughout; precision='float32' is "
                             "not supported on it")
        return "rodas5p_dense" if dense else "rodas5p"
    if name == D.HIP_ROSENBROCK23:
        if precision == "float32":
            raise ValueError(
                f"saveat with precision='float32' is built for the Kuramoto engine only; the "
                f"stiff lane serves field {b.key!r} in float64")
        return "rb23"
    raise ValueError(
        f"saveat is not built for field {b.key!r} on engine {name!r}: this build serves "
        "dense output on the Tsit5 lanes (hip_tsit5 / hip_replay), the Rosenbrock23 and "
        "Rodas5P lanes (hip_rosenbrock23, hip_rodas5p) and the Kuramoto engine (hip_kuramoto). "
        "Drop saveat for a final-state solve or gradient.")


_FWD_ENTRY = {"tsit5": "gsx_tsit5_fwd", "rb23": "gsx_rb23_fwd", "rodas5p": "gsx_rodas5p_fwd",
              "rodas5p_dense": "gsx_rodas5p_fwd"}
_SOLVER = {"tsit5": D.HIP_TSIT5, "rb23": D.HIP_ROSENBROCK23, "rodas5p": D.HIP_RODAS5P,
           "rodas5p_dense": D.HIP_RODAS5P}


def _solve_saveat_lane(lane, b, problem, y0, params, ts, *, rtol, atol, dev, max_steps):
    from . import hip_tsit5 as H
    y0_soa = torch.as_tensor(np.asarray(y0, np.float64), device=dev).t().contiguous()
    p_t = torch.as_tensor(np.asarray(params, np.float64), device=dev)
    p_soa = b.native_params(p_t).t().contiguous()
    aux = None if b.aux_of is None else torch.as_tensor(b.aux_of(problem), device=dev)
    rec = H.record(b.field_id, y0_soa, p_soa, t0=problem.t0, t1=problem.t1, rtol=rtol,
                   atol=atol, max_steps=max_steps, aux=aux, entry=_FWD_ENTRY[lane])
    sv = H.saveat(rec, torch.as_tensor(ts, device=dev), keep_states=False, method=lane)
    status = rec.status.cpu().numpy()
    y_final = rec.y_final.t().cpu().numpy().copy()
    y_saved = sv.ys.permute(2, 0, 1).cpu().numpy()
    y_final[status != 0] = np.nan
    y_saved[status != 0] = np.nan
    return SolveResult(y_final=y_final, accepted_steps=rec.n_acc.cpu().numpy().astype(np.int64),
                       rejected_steps=rec.n_rej.cpu().numpy().astype(np.int64),
                       solver=_SOLVER[lane], y_saved=y_saved, ts_saved=ts)


def solve_saveat(name, b, problem, y0, params, saveat, *, rtol, atol, device,
                 precision="float64", max_steps=100000, dense=False) -> SolveResult:
    from .api import _device
    lane = _check(name, b, precision, dense)
    ts = KU.validate_saveat(saveat, problem.t0, problem.t1)
    dev = _device(device)
    if lane in ("tsit5", "rb23", "rodas5p", "rodas5p_dense"):
        if lane == "rb23":
            from .hip_rosenbrock23 import MAX_STEPS as max_steps  # record_rosenbrock's 50000
        elif lane.startswith("rodas5p"):
            from .hip_rodas5p import MAX_STEPS as max_steps  # record_rodas5p's 50000
        return _solve_saveat_lane(lane, b, problem, y0, params, ts, rtol=rtol, atol=atol,
                                  dev=dev, max_steps=max_steps)
    y0_soa = torch.as_tensor(np.asarray(y0, np.float64), device=dev).t().contiguous()
    om_soa = torch.as_tensor(np.asarray(params, np.float64), device=dev).t().contiguous()
    K = torch.as_tensor(b.aux_of(problem), device=dev)
    ts_d = torch.as_tensor(ts, device=dev)
    rec = KU.record(K, y0_soa, om_soa, t0=problem.t0, t1=problem.t1, rtol=rtol, atol=atol,
                    ts=ts_d, max_steps=max_steps, fp32=precision == "float32")
    ys = KU.saveat(rec)
    status = rec.status.cpu().numpy()
    y_final = rec.y_final.t().cpu().numpy().copy()
    y_saved = ys.cpu().numpy()
    y_final[status != 0] = np.nan
    y_saved[status != 0] = np.nan
    return SolveResult(y_final=y_final, accepted_steps=rec.n_acc.cpu().numpy().astype(np.int64),
                       rejected_steps=rec.n_rej.cpu().numpy().astype(np.int64), solver=name,
                       y_saved=y_saved, ts_saved=ts)


def make_saveat_closure(name, b, problem, y0, params, saveat, *, wrt, rtol, atol, device,
                        aux=None, precision="float64", max_steps=100000, ckpt_interval=None,
                        strict=False, dense=False):
    from .closure import (KuramotoClosure, make_rb23_saveat_closure,
                          make_rodas5p_saveat_closure, make_tsit5_saveat_closure)
    lane = _check(name, b, precision, dense)
    ts = KU.validate_saveat(saveat, problem.t0, problem.t1)
    if lane.startswith("rodas5p"):
        from .hip_rodas5p import MAX_STEPS
        return make_rodas5p_saveat_closure(b, problem, y0, params, wrt=wrt, rtol=rtol,
                                           atol=atol, device=device, ts=ts,
                                           dense=lane == "rodas5p_dense",
                                           ckpt_interval=ckpt_interval, strict=strict, aux=aux,
                                           max_steps=MAX_STEPS)
    if lane == "rb23":
        from .hip_rosenbrock23 import MAX_STEPS
        return make_rb23_saveat_closure(b, problem, y0, params, wrt=wrt, rtol=rtol, atol=atol,
                                        device=device, ts=ts, ckpt_interval=ckpt_interval,
                                        strict=strict, aux=aux, max_steps=MAX_STEPS)
    if lane == "tsit5":
        return make_tsit5_saveat_closure(b, problem, y0, params, wrt=wrt, rtol=rtol, atol=atol,
                                         device=device, ts=ts, ckpt_interval=ckpt_interval,
                                         strict=strict, aux=aux, max_steps=max_steps)
    if aux is None:
        aux = torch.as_tensor(b.aux_of(problem), device=device)
    return KuramotoClosure(b, problem, y0, params, wrt=wrt, rtol=rtol, atol=atol,
                           device=device, aux=aux, ts=ts, max_steps=max_steps,
                           fp32=precision == "float32")
