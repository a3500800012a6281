"""Pins the oracle's Rosenbrock23 record against vectors written by the reference's own numpy
oracle (``tests/_oracles/gradsolve_g2_spike_oracle.py:138-184``, fixtures made by
``tests/golden/make_rb23_golden.py``).

The fixtures fix the accept/reject sequence (and so the error estimate (h/6)(k1 - 2k2 + k3),
the -1/3 controller exponent, the clamp, the WRMS norm and the controller constants the Tsit5
record shares), the accepted step sizes and the final states.

Tolerances: accepted and rejected counts identical on every trajectory; tape entries within
1e-5 relative (the reference oracle solves W with LAPACK's ``np.linalg.solve`` three times per
step, the oracle factorises once; the controller magnifies that rounding difference); final
states within 1e-10 of the state's magnitude.

The mutation checks plant plausible mistakes in a short Python restatement (not in the oracle
under test) and show that each one changes the counts of the committed fixture, so a passing
count comparison rules them out.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "rb23_reference_oracle.npz")

CASES = [("c3", O.ROBERTSON), ("c3n", O.ROBERTSON), ("rish", O.ROBERTSON),
         ("lin", O.LINEAR_DECAY)]


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def _case(g, pre):
    t0, t1 = g[pre + "_t"]
    rtol, atol = g[pre + "_tol"]
    return g[pre + "_y0"], g[pre + "_params"], float(t0), float(t1), float(rtol), float(atol)


@pytest.mark.parametrize("pre,fid", CASES)
def test_oracle_record_matches_reference_oracle(gold, pre, fid):
    y0, p, t0, t1, rtol, atol = _case(gold, pre)
    yf, tape, na, nr, st = O.rb23_record(fid, y0, p, t0, t1, rtol=rtol, atol=atol)
    assert np.array_equal(st, gold[pre + "_status"])
    assert np.array_equal(na, gold[pre + "_n_acc"]), "accepted counts differ"
    assert np.array_equal(nr, gold[pre + "_n_rej"]), "rejected counts differ"
    ref = gold[pre + "_y_final"]
    scale = np.abs(ref).max(axis=1, keepdims=True)
    assert np.max(np.abs(yf - ref) / scale) < 1e-10
    if pre + "_tape" in gold:
        rt = gold[pre + "_tape"]
        assert tape.shape == rt.shape
        nz = rt != 0
        assert np.array_equal(tape == 0, ~nz), "zero tails differ"
        rel = np.abs(tape[nz] - rt[nz]) / rt[nz]
        assert rel.max() < 1e-5, rel.max()


def test_fixture_invariants(gold):
    """The fixture itself: every trajectory reaches t1, the tape sums to t1 - t0, Robertson
    conserves y0 + y1 + y2 = 1 (a linear invariant every Rosenbrock step keeps to rounding)."""
    for pre in ("c3", "rish", "lin"):
        _, _, t0, t1, _, _ = _case(gold, pre)
        assert np.all(gold[pre + "_status"] == 0)
        assert np.allclose(gold[pre + "_tape"].sum(axis=1), t1 - t0, rtol=0, atol=1e-11)
    for pre in ("c3", "c3n", "rish"):
        s = gold[pre + "_y_final"].sum(axis=1)
        s0 = gold[pre + "_y0"].sum(axis=1)
        assert np.max(np.abs(s - s0)) < 1e-12
    # y' = -k y: the final state is close to y0 exp(-k) at rtol 1e-6
    y0, k = gold["lin_y0"], gold["lin_params"][:, 0]
    exact = y0 * np.exp(-k)[:, None]
    err = np.abs(gold["lin_y_final"] - exact)
    assert np.all(err <= 1e-4 * np.abs(y0) + 1e-8)


_D = 1.0 / (2.0 + np.sqrt(2.0))
_E32 = 6.0 + np.sqrt(2.0)


def _rob_f(y, k):
    return np.array([-k[0] * y[0] + k[2] * y[1] * y[2],
                     k[0] * y[0] - k[2] * y[1] * y[2] - k[1] * y[1] * y[1],
                     k[1] * y[1] * y[1]])


def _rob_j(y, k):
    return np.array([[-k[0], k[2] * y[2], k[2] * y[1]],
                     [k[0], -k[2] * y[2] - 2.0 * k[1] * y[1], -k[2] * y[1]],
                     [0.0, 2.0 * k[1] * y[1], 0.0]])


def _counts(y0, k, t1, rtol, atol, expo=-1.0 / 3.0, err_w=(1.0, -2.0, 1.0), fac_max=5.0,
            norm="rms"):
    """A Rosenbrock23 record (the method of Shampine and Reichelt 1997 with an I-controller)
    with switches that plant a mistake: the controller exponent, the error combination's
    weights, the growth clamp, or the norm."""
    y = np.array(y0, float)
    t, dt = 0.0, t1 / 100.0
    acc = rej = 0
    while t < t1 and acc + rej < 2000:  # the fixture's trajectories take < 500 attempts
        h = min(dt, t1 - t)
        W = np.eye(3) - h * _D * _rob_j(y, k)
        F0 = _rob_f(y, k)
        k1 = np.linalg.solve(W, F0)
        F1 = _rob_f(y + 0.5 * h * k1, k)
        k2 = np.linalg.solve(W, F1 - k1) + k1
        yn = y + h * k2
        F2 = _rob_f(yn, k)
        k3 = np.linalg.solve(W, F2 - _E32 * (k2 - F1) - 2.0 * (k1 - F0))
        e = (h / 6.0) * (err_w[0] * k1 + err_w[1] * k2 + err_w[2] * k3)
        sc = atol + rtol * np.maximum(np.abs(y), np.abs(yn))
        q = e / sc
        err = float(np.sqrt(np.mean(q * q))) if norm == "rms" else float(np.max(np.abs(q)))
        if err <= 1.0:
            t, y = t + h, yn
            acc += 1
        else:
            rej += 1
        dt = h * min(max(0.9 * max(err, 1e-10) ** expo, 0.2), fac_max)
    return acc, rej


@pytest.mark.parametrize("mutant", [
    {},
    {"expo": -0.2},
    {"err_w": (1.0, -2.0, 0.0)},
    {"err_w": (1.0, 2.0, 1.0)},
    {"fac_max": 10.0},
    {"norm": "max"},
])
def test_fixture_detects_planted_mistakes(gold, mutant):
    """Counts of the first 3 stiff-workload trajectories under each planted mistake: the
    correct method reproduces the fixture exactly, every mutant misses at least one."""
    y0, k, t0, t1, rtol, atol = _case(gold, "c3")
    got = [_counts(y0[j], k[j], t1, rtol, atol, **mutant) for j in range(3)]
    ref = [(int(gold["c3_n_acc"][j]), int(gold["c3_n_rej"][j])) for j in range(3)]
    if mutant:
        assert got != ref, f"planted mistake {mutant} is not caught by the fixture"
    else:
        assert got == ref
