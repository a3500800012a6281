"""Measured selection of the persistent engine's launch settings.

The candidates are the waves kept resident per SIMD (the persistent grid size: more waves hide
latency, fewer leave more registers and LDS per wave) and the checkpoint interval C (storage
and checkpoint traffic against re-integration in the adjoint, and the adjoint's LDS ring of C
states per lane against its occupancy). Each candidate runs the full forward + adjoint on a
random subsample of the ensemble, twice after a warm-up; the fastest is kept. Results are
cached per (field, precision, ensemble size, subsample) within the process.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Dict, Optional, Sequence, Tuple

import torch

from . import persist as PS

_CACHE: Dict[Tuple, "Choice"] = {}


@dataclass
class Choice:
    waves_per_simd: int
    ckpt_interval: int
    ms: float            # forward + adjoint time on the subsample
    table: dict          # every candidate's time, ms


def tune(field, y0_soa, p_soa, *, t0, t1, rtol, atol, max_steps, t1_per=None, fp32=False,
         sample: int = 1 << 18, waves: Sequence[int] = (2, 3, 4),
         intervals: Sequence[int] = (2, 4, 8, 16), aux: Optional[torch.Tensor] = None,
         seed: int = 0) -> Choice:
    D, n = y0_soa.shape
    P = p_soa.shape[0]
    key = (field, bool(fp32), n, min(n, sample), tuple(waves), tuple(intervals), rtol, atol)
    if key in _CACHE:
        return _CACHE[key]
    m = min(n, sample)
    g = torch.Generator(device="cpu").manual_seed(seed)
    idx = torch.randperm(n, generator=g)[:m].to(y0_soa.device)
    ys = y0_soa[:, idx].contiguous()
    ps = p_soa[:, idx].contiguous()
    t1s = None if t1_per is None else t1_per[idx].contiguous()
    est = PS.estimate_steps(field, ys, ps, t0=t0, t1=t1, t1_per=t1s, rtol=rtol, atol=atol,
                            aux=aux)
    table = {}
    best = None
    for w in waves:
        for c in intervals:
            eng = PS.PersistentTsit5(field, D, P, m, est=est, waves_per_simd=w,
                                     ckpt_interval=c, fp32=fp32, device=y0_soa.device,
                                     aux=aux)
            eng.plan_longest_first(t1s, t0=t0)
            kw = dict(t0=t0, t1=t1, rtol=rtol, atol=atol, max_steps=max_steps, t1_per=t1s)
            eng.run_forward(ys, ps, **kw)
            torch.mul(eng.y_final, 2.0, out=eng.ybar)
            eng.adjoint(ps, eng.ybar, t0=t0)
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(2):
                eng.forward(ys, ps, **kw)
                torch.mul(eng.y_final, 2.0, out=eng.ybar)
                eng.adjoint(ps, eng.ybar, t0=t0)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t) * 1e3 / 2
            table[f"w{w}_c{c}"] = round(ms, 3)
            if best is None or ms < best[0]:
                best = (ms, w, c)
            del eng
    ch = Choice(best[1], best[2], best[0], table)
    _CACHE[key] = ch
    return ch
