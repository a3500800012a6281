"""Writes ``rb23_reference_oracle.npz``: Rosenbrock23 records produced by the reference's own
numpy oracle, ``tests/_oracles/gradsolve_g2_spike_oracle.py`` (``rosenbrock23_adaptive``,
``:138-184``; ``robertson_jac``, ``:53-69``; ``constant_jac``, ``:129-135``).

Run once, in a checkout that has the reference tree at ``GRADSOLVE_REFERENCE`` (default
``/root/reference``): ``python tests/golden/make_rb23_golden.py``. The oracle file imports only
numpy; it is loaded by file path, so no other part of the reference is imported. Only the
numbers it produces are committed (inputs, accepted-step tapes, counts, statuses, final
states); the tests read the ``.npz`` and never the reference.

Cases (every input drawn from ``numpy.random.default_rng(seed)``, parameters first):

* ``c3``: the stiff Robertson workload -- k = [0.04, 3e7, 1e4] * exp(U(log 0.8, log 1.25)),
  y0 = [1, 0, 0], t in [0, 100], rtol 1e-6, atol 1e-10, dt0 = 1 -- 128 trajectories with
  full tapes (seed 11) and 1024 more with counts and final states only (seed 12).
* ``rish``: the reference's own three-trajectory Robertson batch
  (``tests/test_adjoint_replay.py:224-229``, t1 = 0.5, rtol 1e-6, atol 1e-9).
* ``lin``: f = -k y with the constant Jacobian -k I (``constant_jac``), D = 2, k spanning
  stiff and non-stiff values, t in [0, 1], rtol 1e-6, atol 1e-9.
"""
from __future__ import annotations

import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "rb23_reference_oracle.npz")


def load_reference_oracle():
    root = os.environ.get("GRADSOLVE_REFERENCE", "/root/reference")
    path = os.path.join(root, "tests", "_oracles", "gradsolve_g2_spike_oracle.py")
    spec = importlib.util.spec_from_file_location("gradsolve_rb23_numpy_oracle", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def robertson_f(p):
    """The reference Robertson field, in its textual association
    (``tests/test_adjoint_replay.py:100-106``)."""
    k1, k2, k3 = float(p[0]), float(p[1]), float(p[2])

    def f(t, y):
        y0, y1, y2 = float(y[0]), float(y[1]), float(y[2])
        d0 = -k1 * y0 + k3 * y1 * y2
        d1 = k1 * y0 - k3 * y1 * y2 - k2 * y1 * y1
        d2 = k2 * y1 * y1
        return np.array([d0, d1, d2], dtype=np.float64)

    return f


def run(mod, fs_jacs, y0, t0, t1, rtol, atol, keep_tape=True):
    n, D = y0.shape
    yf = np.zeros((n, D))
    n_acc = np.zeros(n, np.int64)
    n_rej = np.zeros(n, np.int64)
    status = np.zeros(n, np.int32)
    tapes = []
    dt0 = (t1 - t0) / 100.0
    for j in range(n):
        f, jac = fs_jacs[j]
        y, dts, nr, st = mod.rosenbrock23_adaptive(f, jac, y0[j], t0, t1, rtol, atol, dt0,
                                                    max_steps=50000)
        yf[j] = y
        n_acc[j] = len(dts)
        n_rej[j] = nr
        status[j] = st
        tapes.append(np.asarray(dts, dtype=np.float64))
    out = dict(y0=y0, y_final=yf, n_acc=n_acc, n_rej=n_rej, status=status,
               t=np.array([t0, t1]), tol=np.array([rtol, atol]))
    if keep_tape:
        S = int(n_acc.max())
        tape = np.zeros((n, S))
        for j, d in enumerate(tapes):
            tape[j, :len(d)] = d
        out["tape"] = tape
    return out


def robertson_k(n, seed):
    rng = np.random.default_rng(seed)
    return np.array([0.04, 3.0e7, 1.0e4]) * np.exp(
        rng.uniform(np.log(0.8), np.log(1.25), (n, 3)))


def main():
    mod = load_reference_oracle()
    arrays = {}

    def add(prefix, d, **extra):
        for k, v in {**d, **extra}.items():
            arrays[f"{prefix}_{k}"] = v

    for prefix, n, seed, keep in (("c3", 128, 11, True), ("c3n", 1024, 12, False)):
        k = robertson_k(n, seed)
        y0 = np.tile(np.array([1.0, 0.0, 0.0]), (n, 1))
        fj = [(robertson_f(k[j]), (lambda y, _p, p=k[j]: mod.robertson_jac(y, p)))
              for j in range(n)]
        add(prefix, run(mod, fj, y0, 0.0, 100.0, 1e-6, 1e-10, keep_tape=keep), params=k,
            seed=np.array(seed))
        print(f"{prefix}: n={n} mean accepted {arrays[prefix + '_n_acc'].mean():.1f}",
              flush=True)

    y0 = np.array([[1.0, 0.0, 0.0], [0.9, 0.05, 0.05], [0.8, 0.1, 0.1]])
    k = np.array([[0.4, 20.0, 5.0], [0.5, 25.0, 6.0], [0.35, 15.0, 4.0]])
    fj = [(robertson_f(k[j]), (lambda y, _p, p=k[j]: mod.robertson_jac(y, p)))
          for j in range(3)]
    add("rish", run(mod, fj, y0, 0.0, 0.5, 1e-6, 1e-9), params=k)

    kk = np.array([0.5, 3.0, 40.0, 700.0, 5000.0, 1.0e5])
    rng = np.random.default_rng(13)
    y0 = rng.uniform(0.5, 2.0, (kk.size, 2))
    fj = []
    for kj in kk:
        A = -kj * np.eye(2)
        fj.append(((lambda t, y, kj=kj: -kj * y), mod.constant_jac(A)))
    add("lin", run(mod, fj, y0, 0.0, 1.0, 1e-6, 1e-9), params=kk[:, None].copy(),
        seed=np.array(13))

    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)", flush=True)


if __name__ == "__main__":
    sys.exit(main())
