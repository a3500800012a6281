"""The Tsit5 tableau as compiled into the native kernels (read through the C ABI)."""
from __future__ import annotations

import ctypes
import functools

import numpy as np

from . import _native as N


@functools.lru_cache(maxsize=None)
def tsit5() -> dict:
    """{"c": c2..c6, "a": a21..a65 (row order), "b": b1..b6, "e": e1..e7} as tuples."""
    c, a, b, e = np.zeros(5), np.zeros(15), np.zeros(6), np.zeros(7)
    vp = ctypes.c_void_p
    N.call("gsx_tsit5_tableau", vp(c.ctypes.data), vp(a.ctypes.data), vp(b.ctypes.data),
           vp(e.ctypes.data))
    return {"c": tuple(c), "a": tuple(a), "b": tuple(b), "e": tuple(e)}
