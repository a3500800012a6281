"""The per-lane kernel bodies (gradsolve_amd/csrc/gsx_core.h), compiled for the CPU with
contraction off, against the oracle: bitwise forward/tape/counts, adjoint to 1e-13, and
guard regions around every output buffer (sized exactly as the C ABI sizes them) intact."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from gradsolve_amd import workloads as W
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "native_cpu", "liblane_harness.so")
G = 512  # guard elements either side


@pytest.fixture(scope="module")
def lh():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "native_cpu")], check=True)
    L = ctypes.CDLL(LIB)
    vp, i64, i32, d, i = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double, ctypes.c_int
    L.lh_tsit5_fwd.argtypes = [i, i64, vp, vp, d, d, vp, d, d, i64, vp, vp, vp, vp, vp, i64, vp, i32]
    L.lh_tsit5_replay.argtypes = [i, i64, vp, vp, d, vp, vp, vp, vp, i32]
    L.lh_tsit5_adjoint.argtypes = [i, i64, vp, d, vp, i64, vp, vp, i32, vp, vp, vp]
    return L


class Guarded:
    def __init__(self, shape, dtype=np.float64):
        size = int(np.prod(shape))
        fill = np.nan if dtype == np.float64 else -7
        self.buf = np.full(size + 2 * G, fill, dtype=dtype)
        self.a = self.buf[G:G + size].reshape(shape)
        self.dtype = dtype

    @property
    def ptr(self):
        return self.a.ctypes.data

    def intact(self):
        g = np.concatenate([self.buf[:G], self.buf[-G:]])
        return bool(np.all(np.isnan(g))) if self.dtype == np.float64 else bool(np.all(g == -7))


@pytest.mark.parametrize("C", [1, 3, 8, 32])
def test_lane_code_matches_oracle(lh, C):
    e = W.lorenz(200, seed=11, noise=0.1)
    n = 200
    cap = 200
    y0s = np.ascontiguousarray(e.y0.T)
    ps = np.ascontiguousarray(e.params.T)
    yf, na, nr, st = Guarded((3, n)), Guarded((n,), np.int32), Guarded((n,), np.int32), \
        Guarded((n,), np.int32)
    tape = Guarded((cap, n))
    ck = Guarded((cap // C + 1, 3, n))
    lh.lh_tsit5_fwd(1, n, y0s.ctypes.data, ps.ctypes.data, 0.0, 1.0, None, 1e-8, 1e-8, 100000,
                    yf.ptr, na.ptr, nr.ptr, st.ptr, tape.ptr, cap, ck.ptr, C)
    assert all(g.intact() for g in (yf, na, nr, st, tape, ck))
    oyf, otape, ona, onr, ost = O.tsit5_record(O.LORENZ, e.y0, e.params, 0.0, 1.0, rtol=1e-8,
                                               atol=1e-8)
    assert np.array_equal(na.a, ona) and np.array_equal(nr.a, onr) and np.all(st.a == 0)
    assert np.array_equal(yf.a.T, oyf)
    padded = np.zeros_like(otape)
    for j in range(n):
        padded[j, :na.a[j]] = tape.a[:na.a[j], j]
    assert np.array_equal(padded, otape)
    ry, rck = Guarded((3, n)), Guarded((cap // C + 1, 3, n))
    lh.lh_tsit5_replay(1, n, y0s.ctypes.data, ps.ctypes.data, 0.0, tape.ptr, na.ptr, ry.ptr,
                       rck.ptr, C)
    assert ry.intact() and rck.intact()
    assert np.array_equal(ry.a, yf.a)
    ybar = np.ascontiguousarray(2.0 * yf.a)
    gy, gp = Guarded((3, n)), Guarded((1, n))
    lh.lh_tsit5_adjoint(1, n, ps.ctypes.data, 0.0, tape.ptr, cap, na.ptr, ck.ptr, C,
                        ybar.ctypes.data, gy.ptr, gp.ptr)
    assert gy.intact() and gp.intact()
    og0, ogp = O.tsit5_adjoint(O.LORENZ, e.y0, e.params, 0.0, otape, 2.0 * oyf)
    assert np.max(np.abs(gp.a[0] - ogp[:, 0]) / np.abs(ogp[:, 0])) < 1e-13
    assert np.max(np.abs(gy.a.T - og0) / np.abs(og0).max(axis=1, keepdims=True)) < 1e-13


def test_lane_tape_full_and_exhausted(lh):
    e = W.lorenz(8, seed=12)
    n, cap = 8, 20
    y0s = np.ascontiguousarray(e.y0.T)
    ps = np.ascontiguousarray(e.params.T)
    yf, na, nr, st = Guarded((3, n)), Guarded((n,), np.int32), Guarded((n,), np.int32), \
        Guarded((n,), np.int32)
    tape, ck = Guarded((cap, n)), Guarded((cap // 4 + 1, 3, n))
    lh.lh_tsit5_fwd(1, n, y0s.ctypes.data, ps.ctypes.data, 0.0, 1.0, None, 1e-8, 1e-8, 100000,
                    yf.ptr, na.ptr, nr.ptr, st.ptr, tape.ptr, cap, ck.ptr, 4)
    assert all(g.intact() for g in (yf, na, nr, st, tape, ck))
    assert np.all(st.a == 3) and np.all(na.a == cap)
    lh.lh_tsit5_fwd(1, n, y0s.ctypes.data, ps.ctypes.data, 0.0, 1.0, None, 1e-8, 1e-8, 7,
                    yf.ptr, na.ptr, nr.ptr, st.ptr, tape.ptr, cap, ck.ptr, 4)
    assert np.all(st.a == 1) and np.all(na.a + nr.a == 7)


def test_controller_power_fast_variant_accuracy():
    """The default variant's controller power (hardware f32 seed + one second-order step of
    the binomial series in f64) is within 1 ulp of the C library's pow, and the clamped step
    factor within 2 ulp."""
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "native_cpu")], check=True)
    F = ctypes.CDLL(os.path.join(HERE, "native_cpu", "liblane_harness_fast.so"))
    F.lh_ctrl_pow.restype = ctypes.c_double
    F.lh_ctrl_pow.argtypes = [ctypes.c_int, ctypes.c_double]
    F.lh_ctrl_fac.restype = ctypes.c_double
    F.lh_ctrl_fac.argtypes = [ctypes.c_int, ctypes.c_double]
    rng = np.random.default_rng(0)
    xs = np.exp(rng.uniform(np.log(1e-5), np.log(1e4), 20000))
    xx = np.exp(rng.uniform(np.log(1e-14), np.log(1e12), 20000))
    for which, e in ((0, -0.2), (1, -1.0 / 3.0)):
        got = np.array([F.lh_ctrl_pow(which, x) for x in xs])
        ref = xs ** e
        assert np.max(np.abs(got - ref) / np.spacing(ref)) <= 1.0
        fac = np.array([F.lh_ctrl_fac(which, x) for x in xx])
        rf = np.minimum(np.maximum(0.9 * np.maximum(xx, 1e-10) ** e, 0.2), 5.0)
        assert np.max(np.abs(fac - rf) / np.spacing(rf)) <= 2.0


def test_squared_norm_factor_and_mean():
    """The default variant's squared-norm controller: m^(-0.1) against err^(-0.2) with
    err = sqrt(m) evaluated in extended precision (within 2 ulp after the 0.9 factor and
    the [0.2, 5] clamp), and the mean of three squares with one Markstein correction against
    the IEEE quotient (within 1 ulp)."""
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "native_cpu")], check=True)
    F = ctypes.CDLL(os.path.join(HERE, "native_cpu", "liblane_harness_fast.so"))
    F.lh_ctrl_fac_sq.restype = ctypes.c_double
    F.lh_ctrl_fac_sq.argtypes = [ctypes.c_double]
    F.lh_mean_sq3.restype = ctypes.c_double
    F.lh_mean_sq3.argtypes = [ctypes.c_void_p]
    rng = np.random.default_rng(3)
    ms = np.exp(rng.uniform(np.log(1e-24), np.log(1e20), 20000))
    got = np.array([F.lh_ctrl_fac_sq(m) for m in ms])
    ld = np.longdouble
    err = np.sqrt(ms.astype(ld))
    ref = np.clip(ld(0.9) * np.maximum(err, ld(1e-10)) ** ld(-0.2), ld(0.2), ld(5.0))
    ref = ref.astype(np.float64)
    assert np.max(np.abs(got - ref) / np.spacing(ref)) <= 2.0
    # The accept tests m <= 1 and sqrt(m) <= 1 part only at the double just above 1.
    above = np.nextafter(1.0, 2.0)
    assert np.sqrt(above) == 1.0 and not above <= 1.0
    assert np.sqrt(np.nextafter(above, 2.0)) > 1.0 and np.sqrt(np.nextafter(1.0, 0.0)) < 1.0
    # The error-norm quotient: refined reciprocal and one Markstein correction, against the
    # IEEE quotient (within 1 ulp), over the scales the norm sees.
    F.lh_norm_div.restype = ctypes.c_double
    F.lh_norm_div.argtypes = [ctypes.c_double, ctypes.c_double]
    a = rng.standard_normal(20000) * np.exp(rng.uniform(-30, 10, 20000))
    b = np.exp(rng.uniform(np.log(1e-12), np.log(1e6), 20000))
    got = np.array([F.lh_norm_div(x, y) for x, y in zip(a, b)])
    assert np.max(np.abs(got - a / b) / np.spacing(np.abs(a / b))) <= 1.0
    qs = rng.standard_normal((20000, 3)) * np.exp(rng.uniform(-20, 20, (20000, 1)))
    mean = np.array([F.lh_mean_sq3(np.ascontiguousarray(q).ctypes.data) for q in qs])
    s = (qs[:, 0] * qs[:, 0] + qs[:, 1] * qs[:, 1]) + qs[:, 2] * qs[:, 2]
    ieee = s / 3.0
    assert np.max(np.abs(mean - ieee) / np.spacing(ieee)) <= 1.0


def test_sincos_within_two_ulp():
    """The device sincos (three-part pi/2 reduction, degree-5 fitted polynomials; host build of
    the same source) against the C library's sin and cos: within 2 ulp over |x| < 2^20, exact
    library values beyond."""
    import math
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "native_cpu")], check=True)
    F = ctypes.CDLL(os.path.join(HERE, "native_cpu", "liblane_harness_fast.so"))
    F.lh_sincos.restype = None
    F.lh_sincos.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    rng = np.random.default_rng(11)
    k = rng.integers(-600000, 600000, 20000).astype(np.float64)
    near = k * (np.pi / 2) + rng.uniform(-1e-6, 1e-6, k.size)  # close to multiples of pi/2
    xs = np.concatenate([rng.uniform(-40.0, 40.0, 200000), rng.uniform(-1e6, 1e6, 50000),
                         near, rng.uniform(-1e-8, 1e-8, 1000), [0.0, -0.0, 1e-300],
                         rng.uniform(2.0 ** 20, 1e9, 1000)])
    xs = np.ascontiguousarray(xs)
    s = np.empty_like(xs)
    c = np.empty_like(xs)
    F.lh_sincos(xs.size, xs.ctypes.data, s.ctypes.data, c.ctypes.data)
    rs = np.array([math.sin(v) for v in xs])
    rc = np.array([math.cos(v) for v in xs])
    us = np.abs(s - rs) / np.spacing(np.abs(rs))
    uc = np.abs(c - rc) / np.spacing(np.abs(rc))
    print(f"sincos: max ulp sin {us.max():.2f}, cos {uc.max():.2f}; exact "
          f"{np.mean(us == 0):.4f} / {np.mean(uc == 0):.4f}")
    assert us.max() <= 2.0 and uc.max() <= 2.0
    big = np.abs(xs) >= 2.0 ** 20
    assert np.array_equal(s[big], rs[big]) and np.array_equal(c[big], rc[big])


def test_sincosf_within_two_ulp():
    """The FP32 device sincos (host build of the same source) against sin and cos computed in
    FP64 and rounded to FP32: within 2 FP32 ulp over |x| < 2^14."""
    import math
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "native_cpu")], check=True)
    F = ctypes.CDLL(os.path.join(HERE, "native_cpu", "liblane_harness_fast.so"))
    F.lh_sincosf.restype = None
    F.lh_sincosf.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    rng = np.random.default_rng(12)
    k = rng.integers(-10000, 10000, 20000).astype(np.float64)
    near = (k * (np.pi / 2) + rng.uniform(-1e-3, 1e-3, k.size))
    xs = np.concatenate([rng.uniform(-40.0, 40.0, 200000), rng.uniform(-16000, 16000, 50000),
                         near, rng.uniform(-1e-6, 1e-6, 1000), [0.0, -0.0]]).astype(np.float32)
    xs = np.ascontiguousarray(xs)
    s = np.empty_like(xs)
    c = np.empty_like(xs)
    F.lh_sincosf(xs.size, xs.ctypes.data, s.ctypes.data, c.ctypes.data)
    rs = np.array([math.sin(float(v)) for v in xs]).astype(np.float32)
    rc = np.array([math.cos(float(v)) for v in xs]).astype(np.float32)
    us = np.abs(s.astype(np.float64) - rs) / np.spacing(np.abs(rs))
    uc = np.abs(c.astype(np.float64) - rc) / np.spacing(np.abs(rc))
    print(f"sincosf: max ulp sin {us.max():.2f}, cos {uc.max():.2f}")
    assert us.max() <= 2.0 and uc.max() <= 2.0


def test_pair_adjoint_equals_generic_adjoint(lh):
    """Checkpoint interval 2: the pair adjoint (each step's stages formed once, the advance's
    kept for the second VJP, a zero-size step padding odd step counts) returns the generic
    segment adjoint's gradients bit for bit, on a tape with odd and even step counts."""
    e = W.lorenz(300, seed=21)
    n, cap = 300, 200
    y0s = np.ascontiguousarray(e.y0.T)
    ps = np.ascontiguousarray(e.params.T)
    yf, na, nr, st = Guarded((3, n)), Guarded((n,), np.int32), Guarded((n,), np.int32), \
        Guarded((n,), np.int32)
    tape, ck = Guarded((cap, n)), Guarded((cap // 2 + 1, 3, n))
    lh.lh_tsit5_fwd(1, n, y0s.ctypes.data, ps.ctypes.data, 0.0, 1.0, None, 1e-8, 1e-8, 100000,
                    yf.ptr, na.ptr, nr.ptr, st.ptr, tape.ptr, cap, ck.ptr, 2)
    assert np.all(st.a == 0) and len(set(na.a % 2)) == 2
    ybar = np.ascontiguousarray(2.0 * yf.a)
    g0, gp = Guarded((3, n)), Guarded((1, n))
    lh.lh_tsit5_adjoint(1, n, ps.ctypes.data, 0.0, tape.ptr, cap, na.ptr, ck.ptr, 2,
                        ybar.ctypes.data, g0.ptr, gp.ptr)
    q0, qp = Guarded((3, n)), Guarded((1, n))
    lh.lh_tsit5_adjoint_pair.argtypes = [ctypes.c_int, ctypes.c_int64] + [ctypes.c_void_p] + \
        [ctypes.c_double] + [ctypes.c_void_p, ctypes.c_int64] + [ctypes.c_void_p] * 5
    lh.lh_tsit5_adjoint_pair(1, n, ps.ctypes.data, 0.0, tape.ptr, cap, na.ptr, ck.ptr,
                             ybar.ctypes.data, q0.ptr, qp.ptr)
    assert q0.intact() and qp.intact()
    assert np.array_equal(q0.a, g0.a) and np.array_equal(qp.a, gp.a)


@pytest.mark.parametrize("C", [2, 8])
def test_saveat_and_its_gradient_match_oracle(lh, C):
    """The dense-output lanes (gsx_tsit5_saveat / gsx_tsit5_adjoint_saveat's per-trajectory
    bodies, strict build) against the oracle's replay_saveat / adjoint_saveat: the saved
    states bit for bit (saves before t0, interior, exactly at t1 and past it), the gradient of
    a random save-series cotangent plus a final-state cotangent to 1e-12, through the pair
    (C = 2) and the segment (C = 8) adjoints."""
    vp, i64, i32, d, i = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double, \
        ctypes.c_int
    lh.lh_tsit5_saveat.argtypes = [i, i64, vp, vp, d, vp, vp, vp, i64, vp, vp, vp, vp]
    lh.lh_tsit5_adjoint_saveat.argtypes = [i, i64, vp, d, vp, i64, vp, vp, i32, vp, i64, vp,
                                           vp, vp, vp, vp, vp, vp, vp]
    e = W.lorenz(64, seed=31, noise=0.1)
    n, cap = 64, 256
    y0s = np.ascontiguousarray(e.y0.T)
    ps = np.ascontiguousarray(e.params.T)
    yf, na, nr, st = Guarded((3, n)), Guarded((n,), np.int32), Guarded((n,), np.int32), \
        Guarded((n,), np.int32)
    tape = Guarded((cap, n))
    ck = Guarded((cap // C + 1, 3, n))
    lh.lh_tsit5_fwd(1, n, y0s.ctypes.data, ps.ctypes.data, 0.0, 1.0, None, 1e-8, 1e-8, 100000,
                    yf.ptr, na.ptr, nr.ptr, st.ptr, tape.ptr, cap, ck.ptr, C)
    ts = np.array([-0.5, 0.0, 1e-4, 0.013, 0.25, 0.5, 0.5, 0.77, 0.999, 1.0, 1.25])
    k = ts.shape[0]
    ys, sstep, st_t, states = Guarded((k, 3, n)), Guarded((k, n), np.int32), \
        Guarded((k, n)), Guarded((k, 3, n))
    lh.lh_tsit5_saveat(1, n, y0s.ctypes.data, ps.ctypes.data, 0.0, tape.ptr, na.ptr,
                       ts.ctypes.data, k, ys.ptr, sstep.ptr, st_t.ptr, states.ptr)
    assert all(g.intact() for g in (ys, sstep, st_t, states))
    padded = np.zeros((n, int(na.a.max())))
    for j in range(n):
        padded[j, :na.a[j]] = tape.a[:na.a[j], j]
    oyf, oys = O.tsit5_replay_saveat(O.LORENZ, e.y0, e.params, 0.0, padded, ts)
    assert np.array_equal(ys.a.transpose(2, 0, 1), oys)
    assert np.array_equal(ys.a[-2].T, yf.a.T) and np.array_equal(ys.a[1].T, e.y0)
    assert np.all(sstep.a[0] == -1) and np.all(sstep.a[-1] == na.a)
    rng = np.random.default_rng(32)
    ysbar = rng.standard_normal((n, k, 3))
    ybar = rng.standard_normal((n, 3))
    mu = Guarded((k, 3, n))
    g0, gp = Guarded((3, n)), Guarded((1, n))
    ysb_soa = np.ascontiguousarray(ysbar.transpose(1, 2, 0))
    ybar_soa = np.ascontiguousarray(ybar.T)
    lh.lh_tsit5_adjoint_saveat(1, n, ps.ctypes.data, 0.0, tape.ptr, cap, na.ptr, ck.ptr, C,
                               ts.ctypes.data, k, sstep.ptr, st_t.ptr, states.ptr,
                               ysb_soa.ctypes.data, ybar_soa.ctypes.data, mu.ptr, g0.ptr, gp.ptr)
    assert all(g.intact() for g in (mu, g0, gp))
    og0, ogp = O.tsit5_adjoint_saveat(O.LORENZ, e.y0, e.params, 0.0, padded, ts, ysbar, ybar)
    np.testing.assert_allclose(gp.a.T, ogp, rtol=1e-12, atol=1e-12 * np.abs(ogp).max())
    np.testing.assert_allclose(g0.a.T, og0, rtol=1e-12, atol=1e-12 * np.abs(og0).max())


def _stiff_case(case, n, seed, t1_forced=5.0):
    """(lane field id, oracle field id, D, checkpoint rows, y0 [n, D], params [n, P], t1) of a
    stiff harness case: Robertson, or the forced stiff field that reads t (its checkpoints
    carry t in an extra row). Rosenbrock23 is the autonomous specialisation (no df/dt term,
    every stage at the step's start time, as the reference's), so on the forced field its
    error estimate keeps an O(h^2) time term and it needs a shorter span."""
    if case == "robertson":
        e = W.robertson(n, seed=seed)
        return 3, O.ROBERTSON, 3, 3, e.y0, e.params, 20.0
    rng = np.random.default_rng(seed)
    y0 = np.column_stack([rng.uniform(0.2, 0.6, n), rng.uniform(0.4, 0.8, n)])
    prm = np.column_stack([rng.uniform(20.0, 500.0, n), rng.uniform(0.5, 0.9, n),
                           rng.uniform(-0.5, 0.5, n)])
    return 8, O.FORCED_STIFF, 2, 3, y0, prm, t1_forced


@pytest.mark.parametrize("case", ["robertson", "forced"])
@pytest.mark.parametrize("C", [1, 2, 4])
def test_rb23_lanes_and_saveat_match_oracle(lh, C, case):
    """The Rosenbrock23 lane bodies (strict build) against the oracle on Robertson: the record
    (counts, tape and final state bit for bit), the replay, the adjoint (1e-12; C = 2 runs the
    pair adjoint, with and without the save injection), the dense output of gsx_rb23_saveat
    bit for bit, and the gradient of gsx_rb23_adjoint_saveat (save cotangents plus a
    final-state cotangent) to 1e-12, with every buffer bounds-guarded."""
    vp, i64, i32, d, i = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double, \
        ctypes.c_int
    lh.lh_rb23_fwd.argtypes = [i, i64, vp, vp, d, d, vp, d, d, i64, vp, vp, vp, vp, vp, i64, vp,
                               i32]
    lh.lh_rb23_replay.argtypes = [i, i64, vp, vp, d, vp, vp, vp, vp, i32]
    lh.lh_rb23_adjoint.argtypes = [i, i64, vp, d, vp, i64, vp, vp, i32, vp, vp, vp]
    lh.lh_rb23_saveat.argtypes = [i, i64, vp, vp, d, vp, vp, vp, i64, vp, vp, vp, vp]
    lh.lh_rb23_adjoint_saveat.argtypes = [i, i64, vp, d, vp, i64, vp, vp, i32, vp, i64, vp,
                                          vp, vp, vp, vp, vp, vp, vp]
    fid, oid, D, R, y0, prm, t1 = _stiff_case(case, 12, 33, t1_forced=0.5)
    P = prm.shape[1]
    n, cap = 12, 1024
    y0s = np.ascontiguousarray(y0.T)
    ps = np.ascontiguousarray(prm.T)
    yf, na, nr, st = Guarded((D, n)), Guarded((n,), np.int32), Guarded((n,), np.int32), \
        Guarded((n,), np.int32)
    tape = Guarded((cap, n))
    ck = Guarded((cap // C + 1, R, n))
    lh.lh_rb23_fwd(fid, n, y0s.ctypes.data, ps.ctypes.data, 0.0, t1, None, 1e-6, 1e-10, 50000,
                   yf.ptr, na.ptr, nr.ptr, st.ptr, tape.ptr, cap, ck.ptr, C)
    assert all(g.intact() for g in (yf, na, nr, st, tape, ck))
    oyf, otape, ona, onr, ost = O.rb23_record(oid, y0, prm, 0.0, t1, rtol=1e-6,
                                              atol=1e-10)
    assert np.all(st.a == 0) and np.array_equal(na.a, ona) and np.array_equal(nr.a, onr)
    assert np.array_equal(yf.a.T, oyf)
    padded = np.zeros((n, int(na.a.max())))
    for j in range(n):
        padded[j, :na.a[j]] = tape.a[:na.a[j], j]
    assert np.array_equal(padded, otape[:, :padded.shape[1]])
    ry, rck = Guarded((D, n)), Guarded((cap // C + 1, R, n))
    lh.lh_rb23_replay(fid, n, y0s.ctypes.data, ps.ctypes.data, 0.0, tape.ptr, na.ptr, ry.ptr,
                      rck.ptr, C)
    assert ry.intact() and rck.intact() and np.array_equal(ry.a, yf.a)
    w = np.ascontiguousarray(np.tile([[1.0], [10.0], [1.0]][:D], (1, n)))
    gy, gp = Guarded((D, n)), Guarded((P, n))
    lh.lh_rb23_adjoint(fid, n, ps.ctypes.data, 0.0, tape.ptr, cap, na.ptr, ck.ptr, C,
                       w.ctypes.data, gy.ptr, gp.ptr)
    assert gy.intact() and gp.intact()
    og0, ogp = O.rb23_adjoint(oid, y0, prm, 0.0, padded, w.T.copy())
    np.testing.assert_allclose(gp.a.T, ogp, rtol=1e-12, atol=0)
    np.testing.assert_allclose(gy.a.T, og0, rtol=1e-12, atol=1e-12 * np.abs(og0).max())
    ts = np.array([-1.0, 0.0, 1e-5, 0.3, 0.3, 2.0, 7.5, 19.99, 20.0, 25.0]) * (t1 / 20.0)
    k = ts.shape[0]
    ys, sstep, st_t, states = Guarded((k, D, n)), Guarded((k, n), np.int32), \
        Guarded((k, n)), Guarded((k, D, n))
    lh.lh_rb23_saveat(fid, n, y0s.ctypes.data, ps.ctypes.data, 0.0, tape.ptr, na.ptr,
                      ts.ctypes.data, k, ys.ptr, sstep.ptr, st_t.ptr, states.ptr)
    assert all(g.intact() for g in (ys, sstep, st_t, states))
    _, oys = O.rb23_replay_saveat(oid, y0, prm, 0.0, padded, ts)
    assert np.array_equal(ys.a.transpose(2, 0, 1), oys)
    assert np.array_equal(ys.a[-2].T, yf.a.T) and np.array_equal(ys.a[1].T, y0)
    rng = np.random.default_rng(34)
    ysbar = rng.standard_normal((n, k, D))
    ybar = rng.standard_normal((n, D))
    mu = Guarded((k, D, n))
    g0, gq = Guarded((D, n)), Guarded((P, n))
    ysb_soa = np.ascontiguousarray(ysbar.transpose(1, 2, 0))  # kept alive across the call
    ybar_soa = np.ascontiguousarray(ybar.T)
    lh.lh_rb23_adjoint_saveat(fid, n, ps.ctypes.data, 0.0, tape.ptr, cap, na.ptr, ck.ptr, C,
                              ts.ctypes.data, k, sstep.ptr, st_t.ptr, states.ptr,
                              ysb_soa.ctypes.data, ybar_soa.ctypes.data, mu.ptr, g0.ptr, gq.ptr)
    assert all(g.intact() for g in (mu, g0, gq))
    sg0, sgp = O.rb23_adjoint_saveat(oid, y0, prm, 0.0, padded, ts, ysbar, ybar)
    np.testing.assert_allclose(gq.a.T, sgp, rtol=1e-12, atol=1e-12 * np.abs(sgp).max())
    np.testing.assert_allclose(g0.a.T, sg0, rtol=1e-12, atol=1e-12 * np.abs(sg0).max())


@pytest.mark.parametrize("case", ["robertson", "forced"])
@pytest.mark.parametrize("C", [1, 4])
def test_rodas5p_lanes_match_oracle(lh, C, case):
    """The Rodas5P lane bodies (gsx_rodas5p_*, strict build) against the oracle on Robertson:
    the record (counts, tape, final state) and the replay bit for bit, the re-step and the
    dense saves bit for bit (saves before t0, on the mesh, at t1 and past it), and the
    gradients of the final state, of a re-step series and of a dense series to 1e-12, every
    buffer bounds-guarded."""
    vp, i64, i32, d, i = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double, \
        ctypes.c_int
    lh.lh_r5_fwd.argtypes = [i, i64, vp, vp, d, d, d, d, i64, vp, vp, vp, vp, vp, i64, vp, i32]
    lh.lh_r5_replay.argtypes = [i, i64, vp, vp, d, vp, vp, vp]
    lh.lh_r5_saveat.argtypes = [i, i, i64, vp, vp, d, vp, vp, vp, i64, vp, vp, vp, vp]
    lh.lh_r5_adjoint.argtypes = [i, i, i64, vp, d, vp, i64, vp, vp, i32, vp, i64, vp, vp, vp,
                                 vp, vp, vp, vp, vp]
    fid, oid, D, R, y0, prm, t1 = _stiff_case(case, 12, 35)
    P = prm.shape[1]
    n, cap = 12, 512
    y0s = np.ascontiguousarray(y0.T)
    ps = np.ascontiguousarray(prm.T)
    yf, na, nr, st = Guarded((D, n)), Guarded((n,), np.int32), Guarded((n,), np.int32), \
        Guarded((n,), np.int32)
    tape, ck = Guarded((cap, n)), Guarded((cap // C + 1, R, n))
    assert lh.lh_r5_fwd(fid, n, y0s.ctypes.data, ps.ctypes.data, 0.0, t1, 1e-6, 1e-10, 50000,
                        yf.ptr, na.ptr, nr.ptr, st.ptr, tape.ptr, cap, ck.ptr, C) == 0
    assert all(g.intact() for g in (yf, na, nr, st, tape, ck))
    oyf, otape, ona, onr, ost = O.rodas5p_record(oid, y0, prm, 0.0, t1,
                                                 rtol=1e-6, atol=1e-10)
    assert np.all(st.a == 0) and np.array_equal(na.a, ona) and np.array_equal(nr.a, onr)
    assert np.array_equal(yf.a.T, oyf)
    S = int(na.a.max())
    padded = np.zeros((n, S))
    for j in range(n):
        padded[j, :na.a[j]] = tape.a[:na.a[j], j]
    assert np.array_equal(padded, otape[:, :S])
    ry = Guarded((D, n))
    lh.lh_r5_replay(fid, n, y0s.ctypes.data, ps.ctypes.data, 0.0, tape.ptr, na.ptr, ry.ptr)
    assert ry.intact() and np.array_equal(ry.a, yf.a)
    rng = np.random.default_rng(36)
    ybar = rng.standard_normal((n, D))
    ybar_soa = np.ascontiguousarray(ybar.T)
    g0, gp = Guarded((D, n)), Guarded((P, n))
    lh.lh_r5_adjoint(fid, 0, n, ps.ctypes.data, 0.0, tape.ptr, cap, na.ptr, ck.ptr, C, None, 0,
                     None, None, None, None, ybar_soa.ctypes.data, None, g0.ptr, gp.ptr)
    assert g0.intact() and gp.intact()
    og0, ogp = O.rodas5p_adjoint(oid, y0, prm, 0.0, padded, ybar)
    np.testing.assert_allclose(gp.a.T, ogp, rtol=1e-12, atol=1e-12 * np.abs(ogp).max())
    np.testing.assert_allclose(g0.a.T, og0, rtol=1e-12, atol=1e-12 * np.abs(og0).max())
    ts = np.array([-1.0, 0.0, 1e-5, 0.3, 0.3, 2.0, 7.5, 19.99, 20.0, 25.0]) * (t1 / 20.0)
    k = ts.shape[0]
    ysbar = rng.standard_normal((n, k, D))
    ysb_soa = np.ascontiguousarray(ysbar.transpose(1, 2, 0))
    for dense in (0, 1):
        ys, sstep, st_t, states = Guarded((k, D, n)), Guarded((k, n), np.int32), \
            Guarded((k, n)), Guarded((k, D, n))
        lh.lh_r5_saveat(fid, dense, n, y0s.ctypes.data, ps.ctypes.data, 0.0, tape.ptr, na.ptr,
                        ts.ctypes.data, k, ys.ptr, sstep.ptr, st_t.ptr, states.ptr)
        assert all(g.intact() for g in (ys, sstep, st_t))
        _, oys = O.rodas5p_replay_saveat(oid, y0, prm, 0.0, padded, ts,
                                         dense=bool(dense))
        assert np.array_equal(ys.a.transpose(2, 0, 1), oys), dense
        assert np.array_equal(ys.a[-2].T, yf.a.T) and np.array_equal(ys.a[1].T, y0)
        mu, h0, hp = Guarded((k, D, n)), Guarded((D, n)), Guarded((P, n))
        lh.lh_r5_adjoint(fid, dense, n, ps.ctypes.data, 0.0, tape.ptr, cap, na.ptr, ck.ptr, C,
                         ts.ctypes.data, k, sstep.ptr, st_t.ptr, states.ptr,
                         ysb_soa.ctypes.data, ybar_soa.ctypes.data, mu.ptr, h0.ptr, hp.ptr)
        assert all(g.intact() for g in (mu, h0, hp))
        sg0, sgp = O.rodas5p_adjoint_saveat(oid, y0, prm, 0.0, padded, ts, ysbar,
                                            ybar, dense=bool(dense))
        np.testing.assert_allclose(hp.a.T, sgp, rtol=1e-12, atol=1e-12 * np.abs(sgp).max())
        np.testing.assert_allclose(h0.a.T, sg0, rtol=1e-12, atol=1e-12 * np.abs(sg0).max())
