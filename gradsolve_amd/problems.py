"""Built-in problems: plain objects satisfying the six-member problem contract.

Each ``f_jax(t, y, params)`` is written per trajectory with the same association as the
native field it routes to, and is array-library agnostic (NumPy, torch or jax arrays).
"""
from __future__ import annotations

from typing import Optional

import numpy as np


def _stack(parts, like):
    mod = type(like).__module__
    if mod.startswith("torch"):
        import torch
        return torch.stack(list(parts))
    if mod.startswith("jax") or mod.startswith("jaxlib"):
        import jax.numpy as jnp
        return jnp.stack(list(parts))
    return np.array(parts, dtype=np.float64)


class DiffEqGPULorenz:
    """Lorenz with sigma=10, beta=8/3 and one parameter column rho (params (n, 1)); the name
    routes it to the fused engines (gradsolve/warp/warp_ode.py:463-498)."""

    name = "diffeqgpu_lorenz"
    dim = 3
    is_stiff = False

    def __init__(self, t0: float = 0.0, t1: float = 1.0):
        self.t0, self.t1 = float(t0), float(t1)

    def f_jax(self, t, y, params):
        sigma, beta, rho = 10.0, 8.0 / 3.0, params[0]
        return _stack([sigma * (y[1] - y[0]), rho * y[0] - y[1] - y[0] * y[2],
                       y[0] * y[1] - beta * y[2]], y)


class Lorenz:
    """Lorenz with parameter columns [sigma, beta, rho] (params (n, 3))."""

    name = "lorenz"
    dim = 3
    is_stiff = False

    def __init__(self, t0: float = 0.0, t1: float = 1.0):
        self.t0, self.t1 = float(t0), float(t1)

    def f_jax(self, t, y, params):
        sigma, beta, rho = params[0], params[1], params[2]
        return _stack([sigma * (y[1] - y[0]), rho * y[0] - y[1] - y[0] * y[2],
                       y[0] * y[1] - beta * y[2]], y)


class VanDerPol:
    """Van der Pol, params (n, 1) = mu (gradsolve/warp/_warp_kernel.py:158-174)."""

    name = "vdp"
    dim = 2
    is_stiff = False

    def __init__(self, t0: float = 0.0, t1: float = 1.0):
        self.t0, self.t1 = float(t0), float(t1)

    def f_jax(self, t, y, params):
        x, v = y[0], y[1]
        return _stack([v, params[0] * ((1.0 - x * x) * v - x)], y)


class Robertson:
    """Robertson kinetics, params (n, 3) = [k1, k2, k3], stiff
    (gradsolve/warp/_warp_rosenbrock.py:96-139)."""

    name = "robertson"
    dim = 3
    is_stiff = True

    def __init__(self, t0: float = 0.0, t1: float = 100.0):
        self.t0, self.t1 = float(t0), float(t1)

    def f_jax(self, t, y, params):
        k1, k2, k3 = params[0], params[1], params[2]
        return _stack([-k1 * y[0] + k3 * y[1] * y[2],
                       k1 * y[0] - k3 * y[1] * y[2] - k2 * y[1] * y[1],
                       k2 * y[1] * y[1]], y)


class Kuramoto:
    """Coupled oscillators with a shared D x D coupling K (params (n, D) = omega):
    f = omega + (1/D) * ((K sin theta) * cos theta - (K cos theta) * sin theta)."""

    is_stiff = False

    def __init__(self, K: np.ndarray, t0: float = 0.0, t1: float = 10.0,
                 name: Optional[str] = None):
        self.K = np.ascontiguousarray(np.asarray(K, dtype=np.float64))
        self.dim = int(self.K.shape[0])
        self.name = name or f"kuramoto{self.dim}"
        self.t0, self.t1 = float(t0), float(t1)

    def f_jax(self, t, y, params):
        mod = type(y).__module__
        if mod.startswith("torch"):
            import torch
            K = torch.as_tensor(self.K, dtype=y.dtype, device=y.device)
            s, c = torch.sin(y), torch.cos(y)
        else:
            xp = np
            if mod.startswith("jax") or mod.startswith("jaxlib"):
                import jax.numpy as xp  # noqa: F811
            K = xp.asarray(self.K)
            s, c = xp.sin(y), xp.cos(y)
        return params + (1.0 / self.dim) * ((K @ s) * c - (K @ c) * s)


class LinearDecay:
    """y' = -k y, params (n, 1) = k (closed form y0 * exp(-k (t - t0)))."""

    is_stiff = False

    def __init__(self, dim: int = 2, t0: float = 0.0, t1: float = 1.0, stiff: bool = False):
        self.dim = int(dim)
        self.name = f"linear_decay_{self.dim}"
        self.t0, self.t1 = float(t0), float(t1)
        self.is_stiff = bool(stiff)

    def f_jax(self, t, y, params):
        return -params[0] * y
