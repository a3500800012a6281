"""Loader for the gfx950 C-ABI library ``_lib/libgradsolve_amd.so``.

The library is the only compute path of this package. If it is missing or cannot be loaded
the calls below raise ``RuntimeError`` -- there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libgradsolve_amd.so")

GSX_OK = 0
GSX_E_INVALID = -1
GSX_E_HIP = -2
GSX_E_UNSUPPORTED = -3

STRICT = 1  # GSX_STRICT: no fused multiply-add contraction
FP32 = 2    # GSX_FP32: FP32 stages, compensated state, FP64 controller

FIELD_LORENZ = 1
FIELD_LINEAR_DECAY = 2
FIELD_ROBERTSON = 3
FIELD_KURAMOTO = 4
FIELD_VDP = 5
FIELD_LORENZ3 = 6
FIELD_FORCED_STIFF = 8
FIELD_RUNTIME_BASE = 1000  # ids of fields registered at run time start here
FIELD_STIFF = 1            # gsx_field_register flags
FIELD_TIME_DEPENDENT = 2

_lock = threading.Lock()
_lib = None

_vp = ctypes.c_void_p
_i = ctypes.c_int
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_d = ctypes.c_double

# name -> (restype, argtypes); every symbol include/gradsolve_amd.h declares.
SIGNATURES = {
    "gsx_version": (_i, []),
    "gsx_last_error": (ctypes.c_char_p, []),
    "gsx_tsit5_tableau": (_i, [_vp, _vp, _vp, _vp]),
    "gsx_ckpt_rows": (_i, [_i, _i]),
    "gsx_ckpt_slots": (_i64, [_i64, _i32]),
    "gsx_tsit5_fwd": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _d, _vp, _d, _d, _i64,
                           _vp, _vp, _vp, _vp, _vp, _i64, _vp, _i32, _i, _vp]),
    "gsx_tsit5_replay": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _vp, _vp, _vp, _vp, _i32,
                              _i, _vp]),
    "gsx_tsit5_adjoint": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _vp, _i64, _vp, _vp,
                               _i32, _vp, _vp, _vp, _i, _vp]),
    "gsx_rb23_fwd": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _d, _vp, _d, _d, _i64,
                          _vp, _vp, _vp, _vp, _vp, _i64, _vp, _i32, _i, _vp]),
    "gsx_rb23_replay": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _vp, _vp, _vp, _vp, _i32,
                             _i, _vp]),
    "gsx_rb23_adjoint": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _vp, _i64, _vp, _vp,
                              _i32, _vp, _vp, _vp, _i, _vp]),
    "gsx_persist_lanes": (_i64, [_i]),
    "gsx_persist_record_len": (_i32, [_i, _i, _i32]),
    "gsx_tsit5_persist_fwd": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _d, _vp, _d, _d, _i64,
                                   _vp, _vp, _vp, _vp, _i64, _vp, _i64, _i32, _vp, _i64, _vp,
                                   _vp, _vp, _vp, _vp, _i, _vp]),
    "gsx_tsit5_persist_adjoint": (_i, [_i, _i64, _i, _i, _vp, _vp, _d, _vp, _i64, _vp, _i32,
                                       _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp]),
    "gsx_kuramoto_fwd": (_i, [_i64, _i, _vp, _vp, _vp, _d, _d, _d, _d, _i64, _vp, _i64, _vp,
                              _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _i, _vp]),
    "gsx_kuramoto_replay": (_i, [_i64, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp]),
    "gsx_kuramoto_saveat": (_i, [_i64, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp,
                                 _i, _vp]),
    "gsx_kuramoto_adjoint": (_i, [_i64, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp,
                                  _vp, _vp, _vp, _vp, _i, _vp]),
    "gsx_plain_tsit5_fwd": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _d, _d, _d, _i64, _vp,
                                 _vp, _vp, _vp, _vp, _vp, _i64, _vp]),
    "gsx_plain_tsit5_adjoint": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                     _vp, _vp]),
    "gsx_tape_to_padded": (_i, [_i64, _vp, _i64, _vp, _i64, _vp, _vp]),
    "gsx_tape_from_padded": (_i, [_i64, _vp, _i64, _vp, _i64, _vp, _vp]),
    "gsx_tape_zero_tail": (_i, [_i64, _vp, _i64, _vp, _i64, _vp]),
    "gsx_tape_count": (_i, [_i64, _vp, _i64, _vp, _vp]),
    "gsx_tsit5_saveat": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _vp, _vp, _vp, _vp, _i64,
                              _vp, _vp, _vp, _vp, _i, _vp]),
    "gsx_tsit5_adjoint_saveat": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _vp, _i64, _vp, _vp,
                                      _i32, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                      _i, _vp]),
    "gsx_tsit5_wide_ws_rows": (_i64, [_i, _i32]),
    "gsx_tsit5_wide_fwd": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _d, _vp, _d, _d, _i64,
                                _vp, _vp, _vp, _vp, _vp, _i64, _vp, _i32, _vp, _vp]),
    "gsx_tsit5_wide_replay": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _vp, _vp, _vp, _vp,
                                   _i32, _vp, _vp]),
    "gsx_tsit5_wide_adjoint": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _vp, _i64, _vp, _vp,
                                    _i32, _vp, _vp, _vp, _vp, _vp]),
    "gsx_tsit5_adjoint_saveat_noalloc": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _vp, _i64,
                                              _vp, _vp, _i32, _vp, _i64, _vp, _vp, _vp, _vp,
                                              _vp, _vp, _vp, _i, _vp]),
    "gsx_rb23_saveat": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _vp, _vp, _vp, _vp, _i64,
                             _vp, _vp, _vp, _vp, _i, _vp]),
    "gsx_rb23_adjoint_saveat": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _vp, _i64, _vp, _vp,
                                     _i32, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                     _i, _vp]),
    "gsx_rodas5p_fwd": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _d, _vp, _d, _d, _i64, _vp,
                             _vp, _vp, _vp, _vp, _i64, _vp, _i32, _i, _vp]),
    "gsx_rodas5p_replay": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _vp, _vp, _vp, _vp, _i32,
                                _i, _vp]),
    "gsx_rodas5p_adjoint": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _vp, _i64, _vp, _vp, _i32,
                                 _vp, _vp, _vp, _i, _vp]),
    "gsx_rodas5p_saveat": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _vp, _vp, _vp, _vp, _i64,
                                _vp, _vp, _vp, _vp, _i, _vp]),
    "gsx_rodas5p_adjoint_saveat": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _vp, _i64, _vp,
                                        _vp, _i32, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp,
                                        _vp, _vp, _i, _vp]),
    "gsx_rodas5p_saveat_dense": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _vp, _vp, _vp, _vp,
                                      _i64, _vp, _vp, _vp, _i, _vp]),
    "gsx_rodas5p_adjoint_saveat_dense": (_i, [_i, _i64, _i, _i, _vp, _vp, _vp, _d, _vp, _i64,
                                              _vp, _vp, _i32, _vp, _i64, _vp, _vp, _vp, _vp,
                                              _vp, _vp, _i, _vp]),
    "gsx_field_register": (_i, [ctypes.c_char_p, _i, _i, _i, ctypes.POINTER(_i)]),
    "gsx_field_release": (_i, [_i]),
}


class NativeError(RuntimeError):
    """A gsx_* call returned an error code."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed (code {code}): {msg}")
        self.code = code


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    """The loaded library; raises RuntimeError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"gradsolve_amd native library not found at {LIB_PATH}: build it with "
                    "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950). "
                    "There is no non-native fallback.")
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def call(name: str, *args) -> int:
    """Call a gsx_* entry point and raise NativeError on a negative return."""
    rc = getattr(lib(), name)(*args)
    if isinstance(rc, int) and rc < 0:
        msg = lib().gsx_last_error().decode(errors="replace")
        if rc == GSX_E_INVALID:
            raise ValueError(f"{name}: {msg}")
        raise NativeError(name, rc, msg)
    return rc


def field_register(source: str, dim: int, n_params: int, flags: int) -> int:
    """Compile a field's C++ source into the fused kernels; returns its field id. A compile
    error raises ValueError carrying the compiler's log."""
    fid = _i(-1)
    call("gsx_field_register", source.encode(), int(dim), int(n_params), int(flags),
         ctypes.byref(fid))
    return int(fid.value)
