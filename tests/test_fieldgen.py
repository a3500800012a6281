"""Fields compiled from Python right-hand sides (gradsolve_amd.fieldgen, gsx_field_register).

CPU tests. The emitted C++ field is compiled for the host (the same source the library hands
to hiprtc) and checked:
  * f against the Python function evaluated on float64 NumPy values: bitwise for arithmetic
    fields (same operation order, contraction off); rel 1e-14 where transcendental functions
    enter (NumPy's and the C library's implementations may differ in the last place, and
    cancellation between terms can magnify that relative to the result);
  * vjp against v^T J with J from complex-step differentiation of the Python function
    (exact to rounding: Im f(y + i h e_j) / h with h = 1e-30), rel 1e-13;
  * for stiff fields, jac against the complex-step Jacobian and jac_vjp against central
    differences of the generated jac (exact to rounding for the polynomial Robertson field);
  * the translatable subset: branches and comparisons on traced values raise UnsupportedRHS
    naming the construct, as the reference's translator does for `cond`
    (gradsolve tests/test_jax_field.py:112-118).
The library side registers fields through hiprtc (no GPU needed to compile): errors carry the
compiler log, and the registered id answers the layout queries. GPU parity of the kernels
built from these fields is in test_gpu_runtime_field.py.
"""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

from gradsolve_amd import _native as N
from gradsolve_amd import fieldgen as FG
from gradsolve_amd.fields import UnsupportedRHS

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "gradsolve_amd", "csrc")


def lorenz(t, y, p):
    sigma, beta, rho = 10.0, 8.0 / 3.0, p[0]
    return np.array([sigma * (y[1] - y[0]), rho * y[0] - y[1] - y[0] * y[2],
                     y[0] * y[1] - beta * y[2]])


def vdp(t, y, p):
    x, v, mu = y[0], y[1], p[0]
    return np.array([v, mu * ((1.0 - x * x) * v - x)])


def forced_pendulum(t, y, p):
    return np.array([y[1], -p[0] * np.sin(y[0]) - p[1] * y[1] + 0.5 * np.cos(t)])


def mixed(t, y, p):
    a, b = p[0], p[1]
    return np.array([np.tanh(a * y[0]) - np.exp(-y[1] * y[1]) + y[2] ** 3,
                     np.sqrt(1.0 + y[0] * y[0]) * b - np.log1p(y[2] * y[2]),
                     y[0] / (1.0 + y[1] ** 2) + np.arctan(y[2]) * a])


def ring(t, y, p):
    a, b = p[0], p[1]
    lap = np.roll(y, -1) - 2.0 * y + np.roll(y, 1)
    return a * lap - y * y * y + b


def robertson(t, y, p):
    k1, k2, k3 = p[0], p[1], p[2]
    return np.stack([-k1 * y[0] + k3 * y[1] * y[2],
                     k1 * y[0] - k3 * y[1] * y[2] - k2 * y[1] * y[1],
                     k2 * y[1] * y[1]])


class HostField:
    """The emitted struct compiled for the host."""

    def __init__(self, f, dim, n_params, stiff=False):
        tr = FG.trace(f, dim, n_params)
        self.src = FG.emit_field(tr, stiff=stiff)
        self.D, self.P = dim, max(n_params, 1)
        self.time_dependent = tr.time_dependent
        self.dir = tempfile.mkdtemp(prefix="fg_")
        cpp = os.path.join(self.dir, "h.cpp")
        with open(cpp, "w") as fh:
            fh.write('#include "gsx_core.h"\nnamespace gsx_user {\nusing namespace ::gsx;\n')
            fh.write(self.src)
            fh.write('}\nextern "C" {\n'
                     'void hf_f(double t, const double* y, const double* p, double* dy) '
                     '{ gsx_user::Field::f<double>(t, y, p, dy, nullptr); }\n'
                     'void hf_vjp(double t, const double* y, const double* p, const double* v, '
                     'double* yb, double* pb) '
                     '{ gsx_user::Field::vjp<double>(t, y, p, v, yb, pb, nullptr); }\n')
            if stiff:
                fh.write('void hf_jac(double t, const double* y, const double* p, double* J) '
                         '{ gsx_user::Field::jac<double>(t, y, p, J, nullptr); }\n'
                         'void hf_jac_vjp(double t, const double* y, const double* p, '
                         'const double* Jb, double* yb, double* pb) '
                         '{ gsx_user::Field::jac_vjp<double>(t, y, p, Jb, yb, pb, nullptr); }\n')
            if stiff and tr.time_dependent:
                fh.write('void hf_dfdt(double t, const double* y, const double* p, double* ft) '
                         '{ gsx_user::Field::dfdt<double>(t, y, p, ft, nullptr); }\n'
                         'void hf_dfdt_vjp(double t, const double* y, const double* p, '
                         'const double* ftb, double* yb, double* pb) '
                         '{ gsx_user::Field::dfdt_vjp<double>(t, y, p, ftb, yb, pb, nullptr); }\n')
            fh.write("}\n")
        so = os.path.join(self.dir, "h.so")
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                        "-fno-builtin-sin", "-fno-builtin-cos", f"-I{CSRC}", "-o", so, cpp],
                       check=True)
        self.L = ctypes.CDLL(so)

    @staticmethod
    def _p(a):
        return np.ascontiguousarray(a, np.float64).ctypes.data_as(ctypes.c_void_p)

    def f(self, t, y, p):
        out = np.zeros(self.D)
        self.L.hf_f(ctypes.c_double(t), self._p(y), self._p(p),
                    out.ctypes.data_as(ctypes.c_void_p))
        return out

    def vjp(self, t, y, p, v):
        yb, pb = np.zeros(self.D), np.zeros(self.P)
        self.L.hf_vjp(ctypes.c_double(t), self._p(y), self._p(p), self._p(v),
                      yb.ctypes.data_as(ctypes.c_void_p), pb.ctypes.data_as(ctypes.c_void_p))
        return yb, pb

    def jac(self, y, p, t=0.0):
        J = np.zeros(self.D * self.D)
        self.L.hf_jac(ctypes.c_double(t), self._p(y), self._p(p),
                      J.ctypes.data_as(ctypes.c_void_p))
        return J.reshape(self.D, self.D)

    def jac_vjp(self, y, p, Jb, t=0.0):
        yb, pb = np.zeros(self.D), np.zeros(self.P)
        self.L.hf_jac_vjp(ctypes.c_double(t), self._p(y), self._p(p), self._p(Jb),
                          yb.ctypes.data_as(ctypes.c_void_p), pb.ctypes.data_as(ctypes.c_void_p))
        return yb, pb

    def dfdt(self, t, y, p):
        ft = np.zeros(self.D)
        self.L.hf_dfdt(ctypes.c_double(t), self._p(y), self._p(p),
                       ft.ctypes.data_as(ctypes.c_void_p))
        return ft

    def dfdt_vjp(self, t, y, p, ftb):
        yb, pb = np.zeros(self.D), np.zeros(self.P)
        self.L.hf_dfdt_vjp(ctypes.c_double(t), self._p(y), self._p(p), self._p(ftb),
                           yb.ctypes.data_as(ctypes.c_void_p),
                           pb.ctypes.data_as(ctypes.c_void_p))
        return yb, pb


def cs_jac(f, t, y, p, wrt="y"):
    """Complex-step Jacobian of the Python right-hand side."""
    h = 1e-30
    x = y if wrt == "y" else p
    cols = []
    for j in range(x.shape[0]):
        xc = x.astype(complex)
        xc[j] += 1j * h
        out = f(t, xc, p.astype(complex)) if wrt == "y" else f(t, y.astype(complex), xc)
        cols.append(np.imag(np.asarray(out, dtype=complex)) / h)
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("f,dim,npar,exact", [(lorenz, 3, 1, True), (vdp, 2, 1, True),
                                               (forced_pendulum, 2, 2, False),
                                               (mixed, 3, 2, False), (ring, 20, 2, True)],
                         ids=["lorenz", "vdp", "pendulum", "mixed", "ring20"])
def test_generated_field_and_vjp(f, dim, npar, exact):
    hf = HostField(f, dim, npar)
    rng = np.random.default_rng(0)
    for _ in range(50):
        t = rng.uniform(0.0, 2.0)
        y = rng.uniform(-2.0, 2.0, dim)
        p = rng.uniform(0.5, 2.0, npar)
        want = np.asarray(f(t, y, p), np.float64)
        got = hf.f(t, y, p)
        if exact:
            assert np.array_equal(got, want), (got, want)
        else:
            np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-15)
        v = rng.standard_normal(dim)
        yb, pb = hf.vjp(t, y, p, v)
        Jy, Jp = cs_jac(f, t, y, p, "y"), cs_jac(f, t, y, p, "p")
        np.testing.assert_allclose(yb, v @ Jy, rtol=1e-13, atol=1e-13 * np.abs(v @ Jy).max())
        np.testing.assert_allclose(pb[:npar], v @ Jp, rtol=1e-13,
                                   atol=1e-13 * np.abs(v @ Jp).max())
    assert hf.time_dependent == (f is forced_pendulum)


def test_generated_lorenz_matches_builtin_expression_order():
    """The generated Lorenz field emits the built-in field's operations in the same order
    (gradsolve tests/test_jax_field.py:82-94 requires 1e-15; here bitwise on the host)."""
    hf = HostField(lorenz, 3, 1)
    rng = np.random.default_rng(1)
    y = rng.uniform(-20.0, 20.0, (100, 3))
    rho = rng.uniform(0.0, 40.0, 100)
    for i in range(100):
        yy, r = y[i], rho[i]
        hand = np.array([10.0 * (yy[1] - yy[0]), r * yy[0] - yy[1] - yy[0] * yy[2],
                         yy[0] * yy[1] - (8.0 / 3.0) * yy[2]])
        assert np.array_equal(hf.f(0.0, yy, np.array([r])), hand)


def test_generated_stiff_jacobian_and_its_vjp():
    hf = HostField(robertson, 3, 3, stiff=True)
    rng = np.random.default_rng(2)
    for _ in range(30):
        y = rng.uniform(0.0, 1.0, 3)
        p = np.array([rng.uniform(0.02, 0.06), rng.uniform(2e7, 4e7), rng.uniform(8e3, 1.2e4)])
        J = hf.jac(y, p)
        np.testing.assert_allclose(J, cs_jac(robertson, 0.0, y, p, "y"), rtol=1e-14,
                                   atol=1e-14 * np.abs(J).max())
        Jb = rng.standard_normal((3, 3))
        yb, pb = hf.jac_vjp(y, p, Jb.reshape(-1))
        # J is linear in y and in p, so central differences are exact up to rounding
        fd_y = np.zeros(3)
        fd_p = np.zeros(3)
        for k in range(3):
            e = np.zeros(3)
            e[k] = 1e-3
            fd_y[k] = np.sum(Jb * (hf.jac(y + e, p) - hf.jac(y - e, p))) / 2e-3
            ep = np.zeros(3)
            ep[k] = 1e-3 * p[k]
            fd_p[k] = np.sum(Jb * (hf.jac(y, p + ep) - hf.jac(y, p - ep))) / (2e-3 * p[k])
        np.testing.assert_allclose(yb, fd_y, rtol=1e-8, atol=1e-8 * np.abs(fd_y).max())
        np.testing.assert_allclose(pb, fd_p, rtol=1e-7, atol=1e-7 * np.abs(fd_p).max())


@pytest.mark.parametrize("body,name", [
    (lambda t, y, p: y if y[0] > 0.0 else -y, "cond"),
    (lambda t, y, p: np.array([max(y[0], y[1]), y[1]]), "cond"),
    (lambda t, y, p: np.array([__import__("math").sin(y[0]), y[1]]), "float"),
])
def test_untraceable_right_hand_sides_raise(body, name):
    with pytest.raises(UnsupportedRHS) as ei:
        FG.trace(body, 2, 1)
    assert ei.value.primitive_name == name


def test_wrong_output_length_raises():
    with pytest.raises(UnsupportedRHS) as ei:
        FG.trace(lambda t, y, p: np.array([y[0]]), 2, 1)
    assert ei.value.primitive_name == "shape"


def forced_stiff(t, y, p):
    a, b, c = p[0], p[1], p[2]
    return np.stack([-a * (y[0] - y[1] * y[1]) + 0.4 * np.sin(3.0 * t),
                     y[0] - y[1] - b * y[1] ** 3 + 0.2 * np.cos(t) + c * np.sin(t) * y[0]])


def _cs_t(f, t, y, p):
    """Complex-step df/dt of the Python right-hand side."""
    h = 1e-30
    return np.imag(np.asarray(f(complex(t, h), y.astype(complex), p.astype(complex)),
                              dtype=complex)) / h


def test_time_dependent_stiff_field_jacobian_and_time_derivative():
    """A stiff field that reads t: J(t) against the complex-step Jacobian, df/dt against the
    complex step in t, and the VJPs of both (jac_vjp, dfdt_vjp) against central differences of
    the generated J and df/dt themselves."""
    hf = HostField(forced_stiff, 2, 3, stiff=True)
    assert hf.time_dependent
    rng = np.random.default_rng(4)
    for _ in range(20):
        t = rng.uniform(0.0, 5.0)
        y = rng.uniform(0.1, 1.0, 2)
        p = np.array([rng.uniform(10.0, 500.0), rng.uniform(0.3, 0.9), rng.uniform(-1.0, 1.0)])
        J = hf.jac(y, p, t)
        np.testing.assert_allclose(J, cs_jac(forced_stiff, t, y, p, "y"), rtol=1e-14,
                                   atol=1e-14 * np.abs(J).max())
        ft = hf.dfdt(t, y, p)
        np.testing.assert_allclose(ft, _cs_t(forced_stiff, t, y, p), rtol=1e-14, atol=1e-15)
        Jb = rng.standard_normal(4)
        ftb = rng.standard_normal(2)
        for fn, vjp, seed in ((lambda yy, pp: hf.jac(yy, pp, t).reshape(-1),
                               lambda: hf.jac_vjp(y, p, Jb, t), Jb),
                              (lambda yy, pp: hf.dfdt(t, yy, pp),
                               lambda: hf.dfdt_vjp(t, y, p, ftb), ftb)):
            yb, pb = vjp()
            fd_y = np.array([np.dot(seed, fn(y + e, p) - fn(y - e, p)) / 2e-6
                             for e in np.eye(2) * 1e-6])
            fd_p = np.array([np.dot(seed, fn(y, p + e) - fn(y, p - e)) / (2e-6 * max(1.0, p[k]))
                             for k, e in enumerate(np.eye(3) * 1e-6 * np.maximum(1.0, p))])
            np.testing.assert_allclose(yb, fd_y, rtol=1e-6, atol=1e-6 * max(1.0, np.abs(fd_y).max()))
            np.testing.assert_allclose(pb, fd_p, rtol=1e-6, atol=1e-6 * max(1.0, np.abs(fd_p).max()))


def test_library_registers_generated_field_and_reports_compile_errors():
    """gsx_field_register compiles through hiprtc on the CPU: ids are issued from
    GSX_FIELD_RUNTIME_BASE, layout queries answer for them (one extra checkpoint row for a
    time-dependent field), and a bad source returns GSX_E_INVALID with the compiler's log."""
    src = FG.emit_field(FG.trace(lorenz, 3, 1))
    fid = N.field_register(src, 3, 1, 0)
    assert fid >= N.FIELD_RUNTIME_BASE
    assert N.call("gsx_ckpt_rows", fid, 3) == 3
    src_t = FG.emit_field(FG.trace(forced_pendulum, 2, 2))
    fid_t = N.field_register(src_t, 2, 2, N.FIELD_TIME_DEPENDENT)
    assert fid_t > fid and N.call("gsx_ckpt_rows", fid_t, 2) == 3
    # a stiff field that reads t compiles its Rosenbrock23 and Rodas5P forwards (df/dt)
    src_s = FG.emit_field(FG.trace(forced_stiff, 2, 3), stiff=True)
    fid_s = N.field_register(src_s, 2, 3, N.FIELD_STIFF | N.FIELD_TIME_DEPENDENT)
    assert N.call("gsx_ckpt_rows", fid_s, 2) == 3
    assert N.call("gsx_field_release", fid_s) == 0
    with pytest.raises(ValueError, match="differs from the registered D=4"):
        N.field_register(src, 4, 1, 0)
    with pytest.raises(ValueError, match="error"):
        N.field_register("struct Field { this is not C++ };", 3, 1, 0)
    with pytest.raises(ValueError):
        N.call("gsx_ckpt_rows", fid, 2)  # wrong D for the registered field
    assert N.call("gsx_field_release", fid) == 0
    with pytest.raises(ValueError):
        N.call("gsx_ckpt_rows", fid, 3)


def piecewise(t, y, p):
    return np.array([y[1], -y[0] - p[0] * FG.maximum(y[1], 0.0)
                     + FG.where(y[0] > 0.5, 0.1, -0.1) * (y[0] > -1.0)
                     + 0.01 * FG.minimum(y[0] * y[1], p[1]) * FG.sign(y[0])])


def test_piecewise_subset_field_and_vjp():
    """Comparisons, where, maximum, minimum and sign (the reference translator's lt/gt,
    select, max, min, sign): f bitwise equal to the same Python function on floats, the VJP
    against central differences at points away from the kinks (rel 1e-7)."""
    hf = HostField(piecewise, 2, 2)
    rng = np.random.default_rng(5)
    checked = 0
    while checked < 40:
        y = rng.uniform(-2.0, 2.0, 2)
        p = rng.uniform(0.5, 2.0, 2)
        if min(abs(y[1]), abs(y[0] - 0.5), abs(y[0] + 1.0), abs(y[0] * y[1] - p[1]),
               abs(y[0])) < 1e-3:
            continue
        checked += 1
        want = np.asarray(piecewise(0.0, y, p), np.float64)
        assert np.array_equal(hf.f(0.0, y, p), want)
        v = rng.standard_normal(2)
        yb, pb = hf.vjp(0.0, y, p, v)
        fd_y, fd_p = np.zeros(2), np.zeros(2)
        for k in range(2):
            e = np.zeros(2)
            e[k] = 1e-7
            fd_y[k] = v @ (piecewise(0.0, y + e, p) - piecewise(0.0, y - e, p)) / 2e-7
            fd_p[k] = v @ (piecewise(0.0, y, p + e) - piecewise(0.0, y, p - e)) / 2e-7
        np.testing.assert_allclose(yb, fd_y, rtol=1e-7, atol=1e-7)
        np.testing.assert_allclose(pb, fd_p, rtol=1e-7, atol=1e-7)


def test_python_max_on_traced_values_names_the_alternative():
    with pytest.raises(UnsupportedRHS, match="maximum"):
        FG.trace(lambda t, y, p: np.array([max(y[0], y[1]), y[1]]), 2, 1)


def numpy_piecewise(t, y, p):
    return np.where(y > 0.25, np.clip(y, -1.5, 1.5) * p[0], -y) + \
        np.maximum(y[::-1], 0.1) - np.minimum(y * y[0], p[1]) + np.sign(y) * 0.01 + \
        np.fmax(np.sin(t), 0.0) * (y <= p[0])


def test_numpy_select_max_min_clip_sign_trace():
    """The same subset written with NumPy's own np.where / np.maximum / np.minimum / np.fmax /
    np.clip / np.sign and array comparisons (what a NumPy port of a jnp right-hand side uses):
    f bitwise equal to the Python function on floats, the VJP against central differences away
    from the kinks (rel 1e-7)."""
    hf = HostField(numpy_piecewise, 2, 2)
    assert hf.time_dependent
    rng = np.random.default_rng(6)
    checked = 0
    while checked < 40:
        t = rng.uniform(0.0, 6.0)
        y = rng.uniform(-2.0, 2.0, 2)
        p = rng.uniform(0.5, 2.0, 2)
        kinks = [*(y - 0.25), *(np.abs(y) - 1.5), *(y[::-1] - 0.1), *(y * y[0] - p[1]), *y,
                 np.sin(t), *(y - p[0])]
        if min(abs(k) for k in kinks) < 1e-3:
            continue
        checked += 1
        want = np.asarray(numpy_piecewise(t, y, p), np.float64)
        assert np.array_equal(hf.f(t, y, p), want)
        v = rng.standard_normal(2)
        yb, pb = hf.vjp(t, y, p, v)
        fd_y, fd_p = np.zeros(2), np.zeros(2)
        for k in range(2):
            e = np.zeros(2)
            e[k] = 1e-7
            fd_y[k] = v @ (numpy_piecewise(t, y + e, p) - numpy_piecewise(t, y - e, p)) / 2e-7
            fd_p[k] = v @ (numpy_piecewise(t, y, p + e) - numpy_piecewise(t, y, p - e)) / 2e-7
        np.testing.assert_allclose(yb, fd_y, rtol=1e-7, atol=1e-7)
        np.testing.assert_allclose(pb, fd_p, rtol=1e-7, atol=1e-7)


def allen_cahn(t, y, p):
    """A wide lattice field: a * (discrete Laplacian) + y - y^3 + b."""
    a, b = p[0], p[1]
    lap = np.roll(y, -1) - 2.0 * y + np.roll(y, 1)
    return a * lap + y - y * y * y + b


def test_register_limits_and_the_wide_lane():
    """Stiff fields stop at D = 12 (each lane of the stiff lanes factorises a D x D W in
    registers; gradsolve/warp/warp_rosenbrock.py:67,324-340 refuses its fused stiff kernel
    above the same 12). A non-stiff field may be up to 4096 wide: above 64 it compiles the wide
    Tsit5 lane at registration, and the register lanes refuse it before any device work."""
    src = FG.emit_field(FG.trace(lambda t, y, p: -p[0] * y, 13, 1), stiff=True)
    with pytest.raises(ValueError, match="register limit"):
        N.field_register(src, 13, 1, N.FIELD_STIFF)
    src12 = FG.emit_field(FG.trace(lambda t, y, p: -p[0] * y, 12, 1), stiff=True)
    fid12 = N.field_register(src12, 12, 1, N.FIELD_STIFF)
    assert N.call("gsx_field_release", fid12) == 0
    with pytest.raises(ValueError, match="outside 1..4096"):
        N.field_register(src, 4097, 1, 0)
    wide = FG.emit_field(FG.trace(allen_cahn, 100, 2))
    fid = N.field_register(wide, 100, 2, 0)
    assert N.call("gsx_ckpt_rows", fid, 100) == 100
    assert N.call("gsx_tsit5_wide_ws_rows", 0, 1) == 10
    assert N.call("gsx_tsit5_wide_ws_rows", 1, 8) == 6 * 8 + 14
    fake = 4096  # never dereferenced: the call stops before any device work
    with pytest.raises(N.NativeError, match="wide Tsit5 lane only"):
        N.call("gsx_tsit5_fwd", fid, 4, 100, 2, None, fake, fake, 0.0, 1.0, None, 1e-8, 1e-8,
               10, fake, fake, fake, fake, None, 0, None, 1, 0, None)
    with pytest.raises(ValueError, match="workspace"):
        N.call("gsx_tsit5_wide_fwd", fid, 4, 100, 2, None, fake, fake, 0.0, 1.0, None, 1e-8,
               1e-8, 10, fake, fake, fake, fake, None, 0, None, 1, None, None)
    with pytest.raises(N.NativeError, match="built in"):
        N.call("gsx_tsit5_wide_fwd", N.FIELD_LORENZ, 4, 3, 1, None, fake, fake, 0.0, 1.0, None,
               1e-8, 1e-8, 10, fake, fake, fake, fake, None, 0, None, 1, fake, None)
    assert N.call("gsx_field_release", fid) == 0
