"""The "plain drop-in" comparison arm: gradients by differentiating a step-by-step replay.

This restates the reference's gradient path -- ``jax.grad`` through ``replay_solve_jax``
(gradsolve/warp/warp_replay.py:30-55): vmap over trajectories, a scan over the recorded
step sizes of ``tsit5_step`` (gradsolve/solvers/tsit5_step.py:80-102) with every step
rematerialised (``jax.checkpoint``, ``choose_remat(3, False) is True``,
gradsolve/dispatch.py:268-293) -- with PyTorch autograd on the GPU in place of JAX, which is
not installed. The forward record is the fused HIP kernel (the reference's ``warp_ode``
record), so only the gradient differs from the fused path. It is a baseline for
measurement, not part of the solve/grad path.
"""
from __future__ import annotations

import torch
from torch.utils.checkpoint import checkpoint

from .tableau import tsit5 as _tableau


def _lorenz(y, rho):
    y0, y1, y2 = y
    return (10.0 * (y1 - y0), rho * y0 - y1 - y0 * y2, y0 * y1 - (8.0 / 3.0) * y2)


def _comb(y, h, coefs, ks):
    out = []
    for d in range(3):
        acc = coefs[0] * ks[0][d]
        for c, k in zip(coefs[1:], ks[1:]):
            acc = acc + c * k[d]
        out.append(y[d] + h * acc)
    return tuple(out)


def tsit5_step_lorenz(y0, y1, y2, h, rho):
    """One tsit5_step on component tensors [n] (all 7 stages, B7*k7 kept)."""
    y = (y0, y1, y2)
    tab = _tableau()
    _B = tab["b"]
    a21, a31, a32, a41, a42, a43, a51, a52, a53, a54, a61, a62, a63, a64, a65 = tab["a"]
    k1 = _lorenz(y, rho)
    k2 = _lorenz(_comb(y, h, [a21], [k1]), rho)
    k3 = _lorenz(_comb(y, h, [a31, a32], [k1, k2]), rho)
    k4 = _lorenz(_comb(y, h, [a41, a42, a43], [k1, k2, k3]), rho)
    k5 = _lorenz(_comb(y, h, [a51, a52, a53, a54], [k1, k2, k3, k4]), rho)
    k6 = _lorenz(_comb(y, h, [a61, a62, a63, a64, a65], [k1, k2, k3, k4, k5]), rho)
    k7 = _lorenz(_comb(y, h, list(_B), [k1, k2, k3, k4, k5, k6]), rho)
    return _comb(y, h, list(_B) + [0.0], [k1, k2, k3, k4, k5, k6, k7])


def replay_lorenz(y0_soa: torch.Tensor, rho: torch.Tensor, tape_steps: torch.Tensor,
                  remat: bool = True) -> torch.Tensor:
    """y_final [3, n] from y0 [3, n], rho [n] and the zero-padded step-major tape [S, n]."""
    y = (y0_soa[0], y0_soa[1], y0_soa[2])
    for s in range(tape_steps.shape[0]):
        h = tape_steps[s]
        if remat:
            y = checkpoint(tsit5_step_lorenz, *y, h, rho, use_reentrant=False)
        else:
            y = tsit5_step_lorenz(*y, h, rho)
    return torch.stack(y)


def grad_lorenz(y0_soa: torch.Tensor, rho: torch.Tensor, tape_steps: torch.Tensor,
                remat: bool = True):
    """Gradients of sum(y_final^2) w.r.t. y0 [3, n] and rho [n] through the replay."""
    y0 = y0_soa.detach().clone().requires_grad_(True)
    r = rho.detach().clone().requires_grad_(True)
    yf = replay_lorenz(y0, r, tape_steps, remat=remat)
    loss = (yf * yf).sum()
    gy0, gr = torch.autograd.grad(loss, (y0, r))
    return yf.detach(), gy0, gr
