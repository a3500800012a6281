"""ctypes wrapper of the CPU oracle (``oracle/libgs_oracle.so``).

TEST INFRASTRUCTURE ONLY. This is the checker the HIP path is compared against; only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import it.
It restates gradsolve's FP64 record-and-replay algorithms (see ``gs_oracle.cpp`` for the
reference file:line each function follows). Arrays use the reference's host layout:
``y0 [n, D]``, ``params [n, P]``, ``tape [n, S]`` zero-padded past ``n_acc``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libgs_oracle.so")

LORENZ = 1
LINEAR_DECAY = 2
ROBERTSON = 3
KURAMOTO = 4
VDP = 5
LORENZ3 = 6
LINEAR_LADDER = 7
FORCED_STIFF = 8

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++, contraction off)."""
    if force and os.path.exists(_LIB_PATH):
        os.remove(_LIB_PATH)
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)


def _declare(L):
    i, i64, d = ctypes.c_int, ctypes.c_int64, ctypes.c_double
    L.gso_tsit5_tableau.argtypes = [_dp, _dp, _dp, _dp]
    rec = [i, i64, i, i, _dp, _dp, _dp, d, _dp, _dp, d, d, i64, _dp, _i64p, _i64p, _i32p,
           _dp, i64, i]
    L.gso_tsit5_record.argtypes = rec
    L.gso_rb23_record.argtypes = rec
    rep = [i, i64, i, i, _dp, _dp, _dp, d, _dp, i64, _dp, i]
    L.gso_tsit5_replay.argtypes = rep
    L.gso_rb23_replay.argtypes = rep
    adj = [i, i64, i, i, _dp, _dp, _dp, d, _dp, i64, _dp, _dp, _dp, i]
    L.gso_tsit5_adjoint.argtypes = adj
    L.gso_rb23_adjoint.argtypes = adj
    rsv = [i, i64, i, i, _dp, _dp, _dp, d, _dp, i64, _dp, i64, _dp, _dp, i]
    asv = [i, i64, i, i, _dp, _dp, _dp, d, _dp, i64, _dp, i64, _dp, _dp, _dp, _dp, i]
    L.gso_tsit5_replay_saveat.argtypes = rsv
    L.gso_rb23_replay_saveat.argtypes = rsv
    L.gso_tsit5_adjoint_saveat.argtypes = asv
    L.gso_rb23_adjoint_saveat.argtypes = asv
    L.gso_rodas5p_record.argtypes = rec
    L.gso_rodas5p_replay.argtypes = rep
    L.gso_rodas5p_adjoint.argtypes = adj
    L.gso_rodas5p_replay_saveat.argtypes = rsv
    L.gso_rodas5p_adjoint_saveat.argtypes = asv
    L.gso_rodas5p_replay_saveat_dense.argtypes = rsv
    L.gso_rodas5p_adjoint_saveat_dense.argtypes = asv
    L.gso_rodas5p_tableau.argtypes = [_dp] * 8
    L.gso_rodas5p_stages.argtypes = [i, i, i, _dp, d, _dp, d, _dp, _dp]
    L.gso_tsit5_teacher_forced.argtypes = [i, i64, i, i, _dp, _dp, _dp, d, _dp, d, d, _dp, i64,
                                           _i64p, _dp, _i64p, _i64p, i]
    L.gso_tsit5_teacher_forced.restype = ctypes.c_int
    L.gso_field.argtypes = [i, i, i, _dp, d, _dp, _dp, _dp]
    L.gso_field_vjp.argtypes = [i, i, i, _dp, d, _dp, _dp, _dp, _dp, _dp]
    L.gso_count_ops.argtypes = [i, i, i, i, _dp, _dp, _dp]
    for name in ("gso_tsit5_tableau", "gso_tsit5_record", "gso_rb23_record", "gso_tsit5_replay",
                 "gso_rb23_replay", "gso_tsit5_adjoint", "gso_rb23_adjoint",
                 "gso_tsit5_replay_saveat", "gso_tsit5_adjoint_saveat", "gso_rb23_replay_saveat",
                 "gso_rb23_adjoint_saveat", "gso_rodas5p_record", "gso_rodas5p_replay",
                 "gso_rodas5p_adjoint", "gso_rodas5p_replay_saveat", "gso_rodas5p_adjoint_saveat",
                 "gso_rodas5p_replay_saveat_dense", "gso_rodas5p_adjoint_saveat_dense",
                 "gso_rodas5p_tableau", "gso_rodas5p_stages", "gso_field",
                 "gso_field_vjp", "gso_count_ops"):
        getattr(L, name).restype = ctypes.c_int


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _check(rc, what):
    if rc != 0:
        raise ValueError(f"oracle {what} rejected its arguments (rc={rc})")


def _threads(nthreads):
    return int(nthreads if nthreads else (os.cpu_count() or 1))


def tableau():
    c, a, b, e = np.zeros(5), np.zeros(15), np.zeros(6), np.zeros(7)
    _check(lib().gso_tsit5_tableau(_ptr(c), _ptr(a), _ptr(b), _ptr(e)), "tableau")
    return c, a, b, e


def _record(fn, field, y0, params, t0, t1, rtol, atol, max_steps, cap, aux, dt0, nthreads,
            keep_tape=True):
    y0 = _f64(np.atleast_2d(y0))
    params = _f64(np.atleast_2d(params))
    n, D = y0.shape
    P = params.shape[1]
    t1a = _f64(np.broadcast_to(np.asarray(t1, dtype=np.float64), (n,)))
    dt0a = None if dt0 is None else _f64(np.broadcast_to(np.asarray(dt0, np.float64), (n,)))
    auxa = None if aux is None else _f64(aux)
    yf = np.zeros((n, D))
    nacc = np.zeros(n, np.int64)
    nrej = np.zeros(n, np.int64)
    st = np.zeros(n, np.int32)
    cap = int(min(cap, max_steps)) if keep_tape else 0
    tape = np.zeros((n, cap)) if keep_tape else None
    rc = fn(field, n, D, P, _ptr(auxa), _ptr(y0), _ptr(params), float(t0), _ptr(t1a),
            _ptr(dt0a), float(rtol), float(atol), int(max_steps), _ptr(yf),
            nacc.ctypes.data_as(_i64p), nrej.ctypes.data_as(_i64p), st.ctypes.data_as(_i32p),
            _ptr(tape), cap, _threads(nthreads))
    _check(rc, "record")
    if not keep_tape:
        return yf, None, nacc, nrej, st
    S = int(nacc.max()) if n else 0
    return yf, np.ascontiguousarray(tape[:, :S]), nacc, nrej, st


def tsit5_record(field, y0, params, t0, t1, *, rtol, atol, max_steps=50000, cap=4096,
                 aux=None, dt0=None, nthreads=None, keep_tape=True):
    """Adaptive Tsit5 record -> (y_final[n,D], tape[n,S] or None, n_acc, n_rej, status)."""
    return _record(lib().gso_tsit5_record, field, y0, params, t0, t1, rtol, atol, max_steps,
                   cap, aux, dt0, nthreads, keep_tape)


def rb23_record(field, y0, params, t0, t1, *, rtol, atol, max_steps=50000, cap=8192,
                aux=None, dt0=None, nthreads=None):
    """Adaptive Rosenbrock23 record -> (y_final[n,D], tape[n,S], n_acc, n_rej, status)."""
    return _record(lib().gso_rb23_record, field, y0, params, t0, t1, rtol, atol, max_steps,
                   cap, aux, dt0, nthreads)


def _replay(fn, field, y0, params, t0, tape, aux, nthreads):
    y0 = _f64(np.atleast_2d(y0))
    params = _f64(np.atleast_2d(params))
    tape = _f64(np.atleast_2d(tape))
    n, D = y0.shape
    yf = np.zeros((n, D))
    auxa = None if aux is None else _f64(aux)
    _check(fn(field, n, D, params.shape[1], _ptr(auxa), _ptr(y0), _ptr(params), float(t0),
              _ptr(tape), tape.shape[1], _ptr(yf), _threads(nthreads)), "replay")
    return yf


def tsit5_teacher_forced(field, y0, params, t0, t1, tape, n_acc, *, rtol, atol, aux=None,
                         nthreads=None):
    """Oracle step sizes given the device tape's history
    -> (h_oracle[n,S], n_rej_oracle[n], rejected attempts before each entry [n,S])."""
    y0 = _f64(np.atleast_2d(y0))
    params = _f64(np.atleast_2d(params))
    tape = _f64(np.atleast_2d(tape))
    n, D = y0.shape
    t1a = _f64(np.broadcast_to(np.asarray(t1, dtype=np.float64), (n,)))
    nacc = np.ascontiguousarray(np.asarray(n_acc, dtype=np.int64))
    ho = np.zeros_like(tape)
    nrej = np.zeros(n, np.int64)
    rej_at = np.zeros(tape.shape, np.int64)
    auxa = None if aux is None else _f64(aux)
    _check(lib().gso_tsit5_teacher_forced(field, n, D, params.shape[1], _ptr(auxa), _ptr(y0),
                                          _ptr(params), float(t0), _ptr(t1a), float(rtol),
                                          float(atol), _ptr(tape), tape.shape[1],
                                          nacc.ctypes.data_as(_i64p), _ptr(ho),
                                          nrej.ctypes.data_as(_i64p),
                                          rej_at.ctypes.data_as(_i64p), _threads(nthreads)),
           "teacher_forced")
    return ho, nrej, rej_at


def rodas5p_record(field, y0, params, t0, t1, *, rtol, atol, max_steps=50000, cap=8192,
                   aux=None, dt0=None, nthreads=None):
    """Rodas5P record with the device recorder's semantics (record_jax.py:83-132)."""
    return _record(lib().gso_rodas5p_record, field, y0, params, t0, t1, rtol, atol, max_steps,
                   cap, aux, dt0, nthreads)


def rodas5p_tableau():
    """(alpha [8,8], Gamma strictly lower [8,8], c, d, m, m - mhat, hd [3,8], gamma)."""
    a, g = np.zeros(64), np.zeros(64)
    c, d, m, e = np.zeros(8), np.zeros(8), np.zeros(8), np.zeros(8)
    hd = np.zeros(24)
    gam = np.zeros(1)
    _check(lib().gso_rodas5p_tableau(_ptr(a), _ptr(g), _ptr(c), _ptr(d), _ptr(m), _ptr(e),
                                     _ptr(hd), _ptr(gam)), "rodas5p_tableau")
    return a.reshape(8, 8), g.reshape(8, 8), c, d, m, e, hd.reshape(3, 8), float(gam[0])


def rodas5p_stages(field, t, y, h, params, *, aux=None):
    """The stage vectors k [8, D] of one step of size h from (t, y)."""
    y = np.ascontiguousarray(y, np.float64)
    p = np.ascontiguousarray(params, np.float64)
    k = np.zeros((8, y.shape[0]))
    a = None if aux is None else np.ascontiguousarray(aux, np.float64)
    _check(lib().gso_rodas5p_stages(field, y.shape[0], p.shape[0], _ptr(a), float(t), _ptr(y),
                                    float(h), _ptr(p), _ptr(k)), "rodas5p_stages")
    return k


def rodas5p_replay(field, y0, params, t0, tape, *, aux=None, nthreads=None):
    return _replay(lib().gso_rodas5p_replay, field, y0, params, t0, tape, aux, nthreads)


def rodas5p_adjoint(field, y0, params, t0, tape, ybar, *, aux=None, nthreads=None):
    return _adjoint(lib().gso_rodas5p_adjoint, field, y0, params, t0, tape, ybar, aux, nthreads)


def rodas5p_replay_saveat(field, y0, params, t0, tape, ts, *, dense=False, aux=None,
                          nthreads=None):
    fn = lib().gso_rodas5p_replay_saveat_dense if dense else lib().gso_rodas5p_replay_saveat
    return _replay_saveat(fn, field, y0, params, t0, tape, ts, aux, nthreads)


def rodas5p_adjoint_saveat(field, y0, params, t0, tape, ts, ysbar, yfbar=None, *, dense=False,
                           aux=None, nthreads=None):
    fn = lib().gso_rodas5p_adjoint_saveat_dense if dense else lib().gso_rodas5p_adjoint_saveat
    return _adjoint_saveat(fn, field, y0, params, t0, tape, ts, ysbar, yfbar, aux, nthreads)


def tsit5_replay(field, y0, params, t0, tape, *, aux=None, nthreads=None):
    return _replay(lib().gso_tsit5_replay, field, y0, params, t0, tape, aux, nthreads)


def rb23_replay(field, y0, params, t0, tape, *, aux=None, nthreads=None):
    return _replay(lib().gso_rb23_replay, field, y0, params, t0, tape, aux, nthreads)


def _adjoint(fn, field, y0, params, t0, tape, ybar, aux, nthreads):
    y0 = _f64(np.atleast_2d(y0))
    params = _f64(np.atleast_2d(params))
    tape = _f64(np.atleast_2d(tape))
    ybar = _f64(np.atleast_2d(ybar))
    n, D = y0.shape
    P = params.shape[1]
    y0b = np.zeros((n, D))
    pb = np.zeros((n, P))
    auxa = None if aux is None else _f64(aux)
    _check(fn(field, n, D, P, _ptr(auxa), _ptr(y0), _ptr(params), float(t0), _ptr(tape),
              tape.shape[1], _ptr(ybar), _ptr(y0b), _ptr(pb), _threads(nthreads)), "adjoint")
    return y0b, pb


def tsit5_adjoint(field, y0, params, t0, tape, ybar, *, aux=None, nthreads=None):
    """dL/dy0, dL/dp of the frozen-tape Tsit5 replay for cotangent ybar = dL/dy_final."""
    return _adjoint(lib().gso_tsit5_adjoint, field, y0, params, t0, tape, ybar, aux, nthreads)


def rb23_adjoint(field, y0, params, t0, tape, ybar, *, aux=None, nthreads=None):
    return _adjoint(lib().gso_rb23_adjoint, field, y0, params, t0, tape, ybar, aux, nthreads)


def tsit5_replay_saveat(field, y0, params, t0, tape, ts, *, aux=None, nthreads=None):
    return _replay_saveat(lib().gso_tsit5_replay_saveat, field, y0, params, t0, tape, ts, aux,
                          nthreads)


def rb23_replay_saveat(field, y0, params, t0, tape, ts, *, aux=None, nthreads=None):
    """Rosenbrock23 bracket-and-re-step (gradsolve/warp/warp_replay.py:78-84)."""
    return _replay_saveat(lib().gso_rb23_replay_saveat, field, y0, params, t0, tape, ts, aux,
                          nthreads)


def _replay_saveat(fn, field, y0, params, t0, tape, ts, aux, nthreads):
    y0 = _f64(np.atleast_2d(y0))
    params = _f64(np.atleast_2d(params))
    tape = _f64(np.atleast_2d(tape))
    ts = _f64(ts)
    n, D = y0.shape
    k = ts.shape[0]
    yf = np.zeros((n, D))
    ys = np.zeros((n, k, D))
    auxa = None if aux is None else _f64(aux)
    _check(fn(field, n, D, params.shape[1], _ptr(auxa), _ptr(y0), _ptr(params), float(t0),
              _ptr(tape), tape.shape[1], _ptr(ts), k, _ptr(yf), _ptr(ys), _threads(nthreads)),
           "replay_saveat")
    return yf, ys


def tsit5_adjoint_saveat(field, y0, params, t0, tape, ts, ysbar, yfbar=None, *, aux=None,
                         nthreads=None):
    return _adjoint_saveat(lib().gso_tsit5_adjoint_saveat, field, y0, params, t0, tape, ts,
                           ysbar, yfbar, aux, nthreads)


def rb23_adjoint_saveat(field, y0, params, t0, tape, ts, ysbar, yfbar=None, *, aux=None,
                        nthreads=None):
    return _adjoint_saveat(lib().gso_rb23_adjoint_saveat, field, y0, params, t0, tape, ts,
                           ysbar, yfbar, aux, nthreads)


def _adjoint_saveat(fn, field, y0, params, t0, tape, ts, ysbar, yfbar, aux, nthreads):
    y0 = _f64(np.atleast_2d(y0))
    params = _f64(np.atleast_2d(params))
    tape = _f64(np.atleast_2d(tape))
    ts = _f64(ts)
    n, D = y0.shape
    P = params.shape[1]
    ysbar = _f64(ysbar)
    yfb = None if yfbar is None else _f64(yfbar)
    y0b = np.zeros((n, D))
    pb = np.zeros((n, P))
    auxa = None if aux is None else _f64(aux)
    _check(fn(field, n, D, P, _ptr(auxa), _ptr(y0), _ptr(params), float(t0), _ptr(tape),
              tape.shape[1], _ptr(ts), ts.shape[0], _ptr(ysbar), _ptr(yfb), _ptr(y0b), _ptr(pb),
              _threads(nthreads)), "adjoint_saveat")
    return y0b, pb


def field(fid, t, y, p, aux=None):
    y = _f64(y)
    p = _f64(p)
    dy = np.zeros_like(y)
    auxa = None if aux is None else _f64(aux)
    _check(lib().gso_field(fid, y.shape[0], p.shape[0], _ptr(auxa), float(t), _ptr(y), _ptr(p),
                           _ptr(dy)), "field")
    return dy


def field_vjp(fid, t, y, p, v, aux=None):
    y, p, v = _f64(y), _f64(p), _f64(v)
    yb = np.zeros_like(y)
    pb = np.zeros_like(p)
    auxa = None if aux is None else _f64(aux)
    _check(lib().gso_field_vjp(fid, y.shape[0], p.shape[0], _ptr(auxa), float(t), _ptr(y),
                               _ptr(p), _ptr(v), _ptr(yb), _ptr(pb)), "field_vjp")
    return yb, pb


def count_ops(which, fid, D, P, aux=None):
    """(flops, pows) of one step; which: 0 trial, 1 replay step, 2 adjoint step,
    3 Rosenbrock23 trial, 4 Rosenbrock23 adjoint step."""
    f = ctypes.c_double()
    pw = ctypes.c_double()
    auxa = None if aux is None else _f64(aux)
    _check(lib().gso_count_ops(which, fid, D, P, _ptr(auxa), ctypes.byref(f), ctypes.byref(pw)),
           "count_ops")
    return f.value, pw.value
