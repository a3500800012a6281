"""Pins the oracle's Tsit5 record, replay and adjoint (CPU only).

Each check is one the reference itself uses (gradsolve tests/test_steps.py,
tests/test_tsit5_replay.py, tests/test_grad_closure_y0.py) or one the mathematics fixes:
order conditions of the published tableau, local order against a high-accuracy integrator,
closed forms, invariants of the recorded mesh, central finite differences on the frozen
mesh, and an independent reverse-mode differentiation (torch.autograd, float64) of the same
replay.
"""
import numpy as np
import pytest
import scipy.integrate as si
import torch

from gradsolve_amd import workloads as W
from oracle import oracle as O

C, A, B, E = O.tableau()


def _A_matrix():
    a = np.zeros((7, 7))
    rows = [(1, [0]), (2, [0, 1]), (3, [0, 1, 2]), (4, [0, 1, 2, 3]), (5, [0, 1, 2, 3, 4])]
    k = 0
    for r, cols in rows:
        for c in cols:
            a[r, c] = A[k]
            k += 1
    a[6, :6] = B  # FSAL row: a7j = b_j
    return a


def test_tableau_order_conditions():
    """Fifth-order conditions of b (all 17 trees up to order 5) and the row-sum condition."""
    a = _A_matrix()
    c = np.concatenate([[0.0], C, [1.0]])
    b = np.concatenate([B, [0.0]])
    assert np.allclose(a.sum(axis=1), c, atol=1e-15)
    conds = [
        (b.sum(), 1.0),
        (b @ c, 1 / 2),
        (b @ c**2, 1 / 3), (b @ a @ c, 1 / 6),
        (b @ c**3, 1 / 4), (b @ (c * (a @ c)), 1 / 8), (b @ a @ c**2, 1 / 12),
        (b @ a @ a @ c, 1 / 24),
        (b @ c**4, 1 / 5), (b @ (c**2 * (a @ c)), 1 / 10), (b @ ((a @ c) ** 2), 1 / 20),
        (b @ (c * (a @ c**2)), 1 / 15), (b @ a @ c**3, 1 / 20),
        (b @ (c * (a @ a @ c)), 1 / 30), (b @ a @ (c * (a @ c)), 1 / 40),
        (b @ a @ a @ c**2, 1 / 60), (b @ a @ a @ a @ c, 1 / 120),
    ]
    for got, want in conds:
        assert abs(got - want) < 5e-15, (got, want)


def test_tableau_embedded_weights():
    """e = b - bhat with e7 = -1/66, the embedded pair of order 4 (sum(e) = 0 and the
    order-1..4 quadrature conditions of bhat)."""
    c = np.concatenate([[0.0], C, [1.0]])
    b = np.concatenate([B, [0.0]])
    bhat = b - E
    assert E[6] == -1.0 / 66.0
    assert abs(E.sum()) < 1e-15
    for q in range(4):
        assert abs(bhat @ c**q - 1.0 / (q + 1)) < 1e-14
    # and the embedded method is genuinely lower order: its c^4 condition fails
    assert abs(bhat @ c**4 - 1 / 5) > 1e-4


def test_tableau_matches_native_kernels():
    """The constants compiled into the HIP kernels equal the oracle's (host-side query)."""
    from gradsolve_amd import tableau as T
    t = T.tsit5()
    assert np.array_equal(np.array(t["c"]), C)
    assert np.array_equal(np.array(t["a"]), A)
    assert np.array_equal(np.array(t["b"]), B)
    assert np.array_equal(np.array(t["e"]), E)


def _lorenz_rhs(rho):
    def f(t, y):
        return [10.0 * (y[1] - y[0]), rho * y[0] - y[1] - y[0] * y[2],
                y[0] * y[1] - 8.0 / 3.0 * y[2]]
    return f


def test_local_order_against_dop853():
    """One replay step of size h on Lorenz: local error ~ h^6 (slope 5.5-6.8, as
    gradsolve tests/test_steps.py:99-140 requires against expm)."""
    y0 = np.array([[1.0, 2.0, 20.0]])
    p = np.array([[28.0]])
    hs = np.array([0.02, 0.01, 0.005, 0.0025])
    errs = []
    for h in hs:
        y1 = O.tsit5_replay(O.LORENZ, y0, p, 0.0, np.array([[h]]))[0]
        ref = si.solve_ivp(_lorenz_rhs(28.0), (0, h), y0[0], method="DOP853", rtol=1e-14,
                           atol=1e-14).y[:, -1]
        errs.append(np.linalg.norm(y1 - ref))
    slope = np.polyfit(np.log(hs), np.log(errs), 1)[0]
    assert 5.5 < slope < 6.8, slope


def test_dt_zero_is_identity_bitwise():
    e = W.lorenz(16, seed=3)
    tape = np.zeros((16, 5))
    yf = O.tsit5_replay(O.LORENZ, e.y0, e.params, 0.0, tape)
    assert np.array_equal(yf, e.y0)
    g0, gp = O.tsit5_adjoint(O.LORENZ, e.y0, e.params, 0.0, tape, np.ones((16, 3)))
    assert np.array_equal(g0, np.ones((16, 3)))
    assert np.array_equal(gp, np.zeros((16, 1)))


def test_record_invariants_and_replay():
    """Mesh invariants (gradsolve tests/test_tsit5_replay.py:56-105): tape [n, S = max n_acc],
    zero tail, sum(h) = t1 - t0, h > 0, status 0; the replay of the tape reproduces the
    recorded final state bitwise (tsit5_step's B7*k7 term is an exact +0.0)."""
    e = W.lorenz_small(seed=0)
    yf, tape, na, nr, st = O.tsit5_record(O.LORENZ, e.y0, e.params, 0.0, 1.0, rtol=1e-8,
                                          atol=1e-8)
    assert tape.shape == (1024, na.max())
    assert np.all(st == 0)
    for j in range(1024):
        assert np.all(tape[j, :na[j]] > 0) and np.all(tape[j, na[j]:] == 0)
    assert np.max(np.abs(tape.sum(axis=1) - 1.0)) < 1e-12
    assert np.array_equal(O.tsit5_replay(O.LORENZ, e.y0, e.params, 0.0, tape), yf)


def test_record_against_dop853_final_state():
    e = W.lorenz(32, seed=1, noise=0.1)
    yf, *_ = O.tsit5_record(O.LORENZ, e.y0, e.params, 0.0, 1.0, rtol=1e-8, atol=1e-8)
    for j in range(32):
        ref = si.solve_ivp(_lorenz_rhs(e.params[j, 0]), (0, 1), e.y0[j], method="DOP853",
                           rtol=1e-13, atol=1e-13).y[:, -1]
        assert np.max(np.abs(yf[j] - ref) / np.abs(ref)) < 5e-6


def test_exponential_decay_closed_form():
    """y' = -k y: y(t1) = y0 e^{-k t1} (gradsolve tests/test_tsit5_replay.py:48-50, rtol 2e-6)."""
    rng = np.random.default_rng(0)
    k = rng.uniform(0.1, 3.0, (50, 1))
    y0 = rng.uniform(0.5, 2.0, (50, 2))
    yf, tape, na, nr, st = O.tsit5_record(O.LINEAR_DECAY, y0, k, 0.0, 1.0, rtol=1e-6,
                                          atol=1e-9)
    assert np.max(np.abs(yf - y0 * np.exp(-k)) / (y0 * np.exp(-k))) < 2e-6


def test_exhausted_max_steps_status():
    e = W.lorenz(4, seed=2)
    yf, tape, na, nr, st = O.tsit5_record(O.LORENZ, e.y0, e.params, 0.0, 1.0, rtol=1e-8,
                                          atol=1e-8, max_steps=10)
    assert np.all(st == 1)
    assert np.all(na + nr == 10)


def test_teacher_forced_self_consistency():
    """Given its own tape, the teacher-forced oracle reproduces every entry bitwise; a 1-ulp
    perturbation of one entry changes only later entries."""
    e = W.lorenz(64, seed=5)
    yf, tape, na, nr, st = O.tsit5_record(O.LORENZ, e.y0, e.params, 0.0, 1.0, rtol=1e-8,
                                          atol=1e-8)
    ho, nro, rej_at = O.tsit5_teacher_forced(O.LORENZ, e.y0, e.params, 0.0, 1.0, tape, na,
                                             rtol=1e-8, atol=1e-8)
    assert np.array_equal(ho, tape)
    assert np.array_equal(nro, nr)
    assert np.array_equal(rej_at.sum(axis=1), nr)
    t2 = tape.copy()
    t2[:, 5] = np.nextafter(t2[:, 5], 1.0)
    ho2, _, _ = O.tsit5_teacher_forced(O.LORENZ, e.y0, e.params, 0.0, 1.0, t2, na, rtol=1e-8,
                                       atol=1e-8)
    assert np.array_equal(ho2[:, :6], tape[:, :6])


# --- adjoint ----------------------------------------------------------------------------

def _torch_tsit5_step(f, y, h, p):
    """tsit5_step (all 7 stages, B7*k7 kept) on torch tensors, using the oracle's tableau."""
    a = _A_matrix()
    k = [f(y, p)]
    for s in range(1, 7):
        z = y + h * sum(a[s, j] * k[j] for j in range(s))
        k.append(f(z, p))
    return y + h * sum(float(bb) * k[j] for j, bb in enumerate(np.concatenate([B, [0.0]])))


def _torch_lorenz(y, p):
    return torch.stack([10.0 * (y[1] - y[0]), p[0] * y[0] - y[1] - y[0] * y[2],
                        y[0] * y[1] - (8.0 / 3.0) * y[2]])


def test_adjoint_matches_torch_autograd():
    """Independent reverse mode: torch.autograd (float64, CPU) through a torch restatement
    of the replay, on the oracle's recorded tape."""
    e = W.lorenz(8, seed=6)
    yf, tape, na, nr, st = O.tsit5_record(O.LORENZ, e.y0, e.params, 0.0, 1.0, rtol=1e-8,
                                          atol=1e-8)
    g0, gp = O.tsit5_adjoint(O.LORENZ, e.y0, e.params, 0.0, tape, 2.0 * yf)
    for j in range(8):
        y = torch.tensor(e.y0[j], dtype=torch.float64, requires_grad=True)
        p = torch.tensor(e.params[j], dtype=torch.float64, requires_grad=True)
        z = y
        for s in range(na[j]):
            z = _torch_tsit5_step(_torch_lorenz, z, float(tape[j, s]), p)
        (z * z).sum().backward()
        assert np.max(np.abs(y.grad.numpy() - g0[j]) / np.abs(g0[j]).max()) < 1e-12
        assert abs(p.grad.item() - gp[j, 0]) / abs(gp[j, 0]) < 1e-12


def test_adjoint_frozen_tape_finite_differences():
    """Central differences on the frozen tape, step 1e-6*max(1,|x|), rel < 1e-6
    (gradsolve tests/test_grad_closure_y0.py:144-191)."""
    e = W.lorenz(32, seed=7, noise=0.1)
    yf, tape, na, nr, st = O.tsit5_record(O.LORENZ, e.y0, e.params, 0.0, 1.0, rtol=1e-8,
                                          atol=1e-8)
    g0, gp = O.tsit5_adjoint(O.LORENZ, e.y0, e.params, 0.0, tape, 2.0 * yf)

    def loss(y0, p):
        return np.sum(O.tsit5_replay(O.LORENZ, y0, p, 0.0, tape) ** 2, axis=1)

    eps = 1e-6 * np.maximum(1.0, np.abs(e.params[:, 0]))
    fd = (loss(e.y0, e.params + eps[:, None]) - loss(e.y0, e.params - eps[:, None])) / (2 * eps)
    assert np.max(np.abs(fd - gp[:, 0]) / np.abs(gp[:, 0])) < 1e-6


def test_linear_decay_adjoint_closed_form():
    """For y' = -k y on a fine tape, d sum(y(1)) / dk -> -sum(y0) e^{-k}."""
    k = np.array([[0.7]])
    y0 = np.array([[1.5, 0.5]])
    yf, tape, na, *_ = O.tsit5_record(O.LINEAR_DECAY, y0, k, 0.0, 1.0, rtol=1e-12, atol=1e-14)
    g0, gp = O.tsit5_adjoint(O.LINEAR_DECAY, y0, k, 0.0, tape, np.ones((1, 2)))
    assert abs(gp[0, 0] - (-2.0 * np.exp(-0.7))) < 1e-10
    assert np.allclose(g0, np.exp(-0.7), rtol=1e-10)


@pytest.mark.parametrize("fid,D,P", [(O.LORENZ, 3, 1), (O.VDP, 2, 1), (O.ROBERTSON, 3, 3),
                                     (O.LINEAR_DECAY, 4, 1)])
def test_field_vjp_matches_autograd(fid, D, P):
    rng = np.random.default_rng(fid)
    y = rng.uniform(0.1, 2.0, D)
    p = rng.uniform(0.5, 3.0, P)
    v = rng.standard_normal(D)
    yb, pb = O.field_vjp(fid, 0.0, y, p, v)
    yt = torch.tensor(y, requires_grad=True)
    pt = torch.tensor(p, requires_grad=True)
    dy = torch.tensor(O.field(fid, 0.0, y, p))
    f = {O.LORENZ: lambda y, p: _torch_lorenz(y, p),
         O.VDP: lambda y, p: torch.stack([y[1], p[0] * ((1 - y[0] * y[0]) * y[1] - y[0])]),
         O.ROBERTSON: lambda y, p: torch.stack([
             -p[0] * y[0] + p[2] * y[1] * y[2],
             p[0] * y[0] - p[2] * y[1] * y[2] - p[1] * y[1] * y[1],
             p[1] * y[1] * y[1]]),
         O.LINEAR_DECAY: lambda y, p: -p[0] * y}[fid]
    out = f(yt, pt)
    assert np.allclose(out.detach().numpy(), dy.numpy(), rtol=1e-15, atol=0)
    (out * torch.tensor(v)).sum().backward()
    assert np.allclose(yt.grad.numpy(), yb, rtol=1e-13, atol=1e-13)
    assert np.allclose(pt.grad.numpy(), pb, rtol=1e-13, atol=1e-13)
