"""Synthetic ensemble inputs for the benchmark workloads and parity cases.

Every generator draws from ``numpy.random.default_rng(seed)`` in a fixed order (parameters
first, then initial states, then anything else) so that a seed names one input set.
Arrays are host float64 in the reference's layout: ``y0 [n, D]``, ``params [n, P]``.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np


@dataclass
class Ensemble:
    name: str
    y0: np.ndarray
    params: np.ndarray
    t0: float
    t1: float
    rtol: float
    atol: float
    t1_per: Optional[np.ndarray] = None
    aux: Optional[np.ndarray] = None


def lorenz(n: int, *, seed: int = 0, noise: float = 0.1, rho_lo: float = 20.0,
           rho_hi: float = 30.0, t1: float = 1.0, tol: float = 1e-8) -> Ensemble:
    """Lorenz ensemble (sigma=10, beta=8/3): rho ~ U(rho_lo, rho_hi), then
    y0 = [1, 0, 0] + noise * N(0, 1) per component."""
    rng = np.random.default_rng(seed)
    rho = rng.uniform(rho_lo, rho_hi, n)
    y0 = np.array([1.0, 0.0, 0.0]) + noise * rng.standard_normal((n, 3))
    return Ensemble(f"lorenz_n{n}", y0, rho[:, None].copy(), 0.0, float(t1), tol, tol)


def lorenz_small(seed: int = 0) -> Ensemble:
    """N=1024 Lorenz, y0 noise 0.01, rho ~ U(20, 30), t in [0, 1], rtol = atol = 1e-8."""
    return lorenz(1024, seed=seed, noise=0.01)


def lorenz_1m(seed: int = 0, n: int = 1 << 20) -> Ensemble:
    """N=1,048,576 Lorenz, y0 noise 0.1, rho ~ U(20, 30), t in [0, 1], rtol = atol = 1e-8."""
    return lorenz(n, seed=seed, noise=0.1)


def lorenz_uneven(n: int = 1 << 22, *, seed: int = 0, tol: float = 1e-6) -> Ensemble:
    """Uneven lengths: rho ~ U(0, 60), t1 ~ U(0.5, 10) per trajectory, y0 noise 0.5."""
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.0, 60.0, n)
    y0 = np.array([1.0, 0.0, 0.0]) + 0.5 * rng.standard_normal((n, 3))
    t1 = rng.uniform(0.5, 10.0, n)
    return Ensemble(f"lorenz_uneven_n{n}", y0, rho[:, None].copy(), 0.0, float(t1.max()),
                    tol, tol, t1_per=t1)


def robertson(n: int = 1 << 18, *, seed: int = 0) -> Ensemble:
    """Robertson kinetics: k = [0.04, 3e7, 1e4] * exp(U(log 0.8, log 1.25)) per entry,
    y0 = [1, 0, 0], t in [0, 100], rtol = 1e-6, atol = 1e-10."""
    rng = np.random.default_rng(seed)
    k = np.array([0.04, 3.0e7, 1.0e4]) * np.exp(
        rng.uniform(np.log(0.8), np.log(1.25), (n, 3)))
    y0 = np.tile(np.array([1.0, 0.0, 0.0]), (n, 1))
    return Ensemble(f"robertson_n{n}", y0, k, 0.0, 100.0, 1e-6, 1e-10)


def kuramoto32(n: int = 1 << 17, *, seed: int = 0, D: int = 32) -> Ensemble:
    """Coupled oscillators: K ~ U(0, 2) (D x D, zero diagonal, shared), omega ~ N(0, 1),
    theta0 ~ U(0, 2 pi), t in [0, 10], rtol = atol = 1e-8."""
    rng = np.random.default_rng(seed)
    K = rng.uniform(0.0, 2.0, (D, D))
    np.fill_diagonal(K, 0.0)
    omega = rng.standard_normal((n, D))
    theta0 = rng.uniform(0.0, 2.0 * np.pi, (n, D))
    return Ensemble(f"kuramoto{D}_n{n}", theta0, omega, 0.0, 10.0, 1e-8, 1e-8, aux=K)


def allen_cahn_rhs(t, y, p):
    """Periodic Allen-Cahn lattice: a * (discrete Laplacian) + y - y^3 + b, with p = [a, b]
    (a NumPy right-hand side that the field tracer compiles; any D)."""
    lap = np.roll(y, -1) - 2.0 * y + np.roll(y, 1)
    return p[0] * lap + y - y * y * y + p[1]


def allen_cahn(n: int = 1 << 14, *, D: int = 256, seed: int = 0, t1: float = 2.0,
               tol: float = 1e-8) -> Ensemble:
    """A wide non-stiff ensemble for the high-NVAR cells: a ~ U(0.3, 0.6), b ~ U(-0.1, 0.1),
    y0 ~ U(-1, 1)^D, t in [0, t1], rtol = atol = tol."""
    rng = np.random.default_rng(seed)
    p = np.stack([rng.uniform(0.3, 0.6, n), rng.uniform(-0.1, 0.1, n)], axis=1)
    y0 = rng.uniform(-1.0, 1.0, (n, D))
    return Ensemble(f"allen_cahn{D}_n{n}", y0, p, 0.0, float(t1), tol, tol)
