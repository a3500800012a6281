"""Pin the pure-Python oracle for run-time fields (oracle/generic.py) against the C oracle.

On Lorenz (the Python right-hand side written in the C oracle's field order) the two must
record the same meshes bit for bit -- accepted and rejected counts, every accepted h, the
final states -- and the frozen-mesh replay must land on the recorded final state exactly.
The adjoints (analytic VJPs in C, complex-step VJPs in Python) agree to 1e-12 relative.
"""
import numpy as np

from gradsolve_amd import workloads as W
from oracle import generic as G
from oracle import oracle as O


def lorenz(t, y, p):
    return np.array([10.0 * (y[1] - y[0]), p[0] * y[0] - y[1] - y[0] * y[2],
                     y[0] * y[1] - (8.0 / 3.0) * y[2]])


def test_generic_record_matches_c_oracle_bitwise():
    e = W.lorenz(16, seed=5, noise=0.1)
    yf, tape, na, nr, st = G.tsit5_record(lorenz, e.y0, e.params, 0.0, 1.0, rtol=1e-8,
                                          atol=1e-8)
    oyf, otape, ona, onr, ost = O.tsit5_record(O.LORENZ, e.y0, e.params, 0.0, 1.0, rtol=1e-8,
                                               atol=1e-8)
    assert np.array_equal(na, ona) and np.array_equal(nr, onr)
    assert np.array_equal(st, ost)
    assert np.array_equal(tape, otape)
    assert np.array_equal(yf, oyf)
    assert np.array_equal(G.tsit5_replay(lorenz, e.y0, e.params, 0.0, tape), yf)


def test_generic_adjoint_matches_c_oracle():
    e = W.lorenz(6, seed=6, noise=0.1)
    yf, tape, na, nr, st = G.tsit5_record(lorenz, e.y0, e.params, 0.0, 1.0, rtol=1e-6,
                                          atol=1e-6)
    ybar = 2.0 * yf
    g0, gp = G.tsit5_adjoint(lorenz, e.y0, e.params, 0.0, tape, ybar)
    og0, ogp = O.tsit5_adjoint(O.LORENZ, e.y0, e.params, 0.0, tape, ybar)
    np.testing.assert_allclose(g0, og0, rtol=1e-12, atol=1e-12 * np.abs(og0).max())
    np.testing.assert_allclose(gp, ogp, rtol=1e-12, atol=1e-12 * np.abs(ogp).max())
