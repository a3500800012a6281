"""Host-side API behaviour that needs no GPU: routing, validation, registration."""
import numpy as np
import pytest

import gradsolve_amd as gs
from gradsolve_amd import dispatch as D
from gradsolve_amd import fields as F
from gradsolve_amd import problems as P


def test_choose_engine_cells():
    assert D.choose_engine(3, False, False) == "hip_tsit5"
    assert D.choose_engine(3, False, True) == "hip_replay"
    assert D.choose_engine(3, True, False) == "hip_rosenbrock23"
    assert D.choose_engine(3, True, True) == "hip_rosenbrock23"
    assert D.choose_engine(1000, False, True) == "tsit5_replay"
    assert D.choose_engine(1000, True, True) == "rodas5p_replay"
    assert D.choose_engine(3, True, True, stiff_fused_enabled=False) == "rodas5p_replay"


def test_decision_map_agrees_with_choose_engine():
    for row in D.DECISION_MAP:
        assert D.choose_engine(row["dim"], row["stiff"], row["need_grad"]) == row["engine"]


def test_remat_policy():
    assert D.choose_remat(3, False) is True
    assert D.choose_remat(3, True) is False
    assert D.choose_remat(32, True) is True


def test_bindings():
    assert F.binding_for(P.DiffEqGPULorenz()).field_id == 1
    assert F.binding_for(P.Lorenz()).n_params == 3
    assert F.binding_for(P.Robertson()).stiff
    assert F.binding_for(P.LinearDecay(2)).dim == 2
    assert F.binding_for(P.LinearDecay(7)) is None


def test_register_jax_field_identifies_native_field():
    class MyLorenz:
        name = "my_lorenz_for_test"
        dim = 3
        t0, t1 = 0.0, 1.0
        is_stiff = False

        def f_jax(self, t, y, p):
            return np.array([10.0 * (y[1] - y[0]), y[0] * (p[0] - y[2]) - y[1],
                             y[0] * y[1] - 8.0 / 3.0 * y[2]])

    b = gs.register_jax_field(MyLorenz.name, MyLorenz().f_jax, 3, 1)
    assert b.field_id == 1
    assert F.binding_for(MyLorenz()) is b
    assert gs.register_jax_field(MyLorenz.name, MyLorenz().f_jax, 3, 1) is b  # idempotent


def test_register_jax_field_compiles_unknown_rhs_and_rejects_untraceable():
    """A right-hand side no hand-written field computes is traced and compiled (hiprtc, no GPU
    needed) into a field of its own; one that branches on the state raises UnsupportedRHS."""
    b = gs.register_jax_field("cubic_for_test", lambda t, y, p: -p[0] * y ** 3, 2, 1)
    assert b.field_id >= 1000 and b.source is not None and "struct Field" in b.source
    with pytest.raises(gs.UnsupportedRHS):
        gs.register_jax_field("branchy_for_test",
                              lambda t, y, p: y if y[0] > 0.0 else -y, 2, 1)


def test_problem_f_jax_shapes():
    y = np.array([1.0, 2.0, 3.0])
    assert P.DiffEqGPULorenz().f_jax(0.0, y, np.array([28.0])).shape == (3,)
    assert P.Robertson().f_jax(0.0, y, np.array([0.04, 3e7, 1e4])).shape == (3,)
    K = np.ones((4, 4)) - np.eye(4)
    assert P.Kuramoto(K).f_jax(0.0, np.zeros(4), np.ones(4)).shape == (4,)


def test_validation_errors():
    e_y0 = np.zeros((4, 3))
    with pytest.raises(ValueError, match="invalid wrt"):
        gs.grad_closure(P.DiffEqGPULorenz(), e_y0, np.ones((4, 1)), wrt="rho")
    with pytest.raises(ValueError, match="shape"):
        gs.solve(P.DiffEqGPULorenz(), np.zeros((4, 2)), np.ones((4, 1)))
    with pytest.raises(ValueError, match="unknown engine"):
        gs.solve(P.DiffEqGPULorenz(), e_y0, np.ones((4, 1)), engine="nope")
    with pytest.raises(ValueError, match="does not include"):
        gs.solve(P.DiffEqGPULorenz(), e_y0, np.ones((4, 1)), engine="diffrax")
    # fused=False skips compiling the problem's f_jax, so nothing serves dim 7
    with pytest.raises(ValueError, match="no native field"):
        gs.solve(P.LinearDecay(7), np.zeros((4, 7)), np.ones((4, 1)), fused=False)
    with pytest.raises(RuntimeError, match="no CPU engine"):
        gs.solve(P.DiffEqGPULorenz(), e_y0, np.ones((4, 1)), device="cpu")


class _AllenCahn100:
    """A 100-wide lattice (above NVAR_CEILING): the reference's general-engine cells."""
    name = "allen_cahn_100"
    dim = 100
    t0 = 0.0
    t1 = 1.0
    is_stiff = False

    def f_jax(self, t, y, p):
        lap = np.roll(y, -1) - 2.0 * y + np.roll(y, 1)
        return p[0] * lap + y - y * y * y + p[1]


def test_high_nvar_cells_route_to_the_wide_lane():
    """gradsolve/dispatch.py:256-258 names tsit5_replay (gradient) and diffrax (forward) above
    NVAR_CEILING; both run on the wide lane once the field is compiled (fused='auto'), with the
    substitution on route.reason. A built-in field asked for the general lane keeps the
    register lane."""
    import gradsolve_amd.api as A
    prob = _AllenCahn100()
    params = np.ones((4, 2))
    pre = []
    A._translate(prob, params, "auto", pre)
    assert F.binding_for(prob).source is not None
    name, reasons = A._resolve("auto", prob, True, None, None, pre)
    assert name == D.HIP_TSIT5_WIDE and "alias:tsit5_replay" in reasons
    name, reasons = A._resolve("auto", prob, False, None, None, pre)
    assert name == D.HIP_TSIT5_WIDE and "diffrax-not-installed" in reasons
    name, reasons = A._resolve("tsit5_replay", prob, True, None, None)
    assert name == D.HIP_TSIT5_WIDE
    name, _ = A._resolve("tsit5_replay", P.DiffEqGPULorenz(), True, None, None)
    assert name == D.HIP_REPLAY
    with pytest.raises(ValueError, match="diffrax"):
        A._resolve("diffrax", prob, False, None, None)
