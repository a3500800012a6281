"""The public API on the GPU: solve / grad_closure with the reference's signatures, torch
autograd through the native adjoint, the wrt forms, replay at new inputs, and routes."""
import numpy as np
import pytest
import torch

import gradsolve_amd as gs
from gradsolve_amd import problems as P
from gradsolve_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_solve_lorenz_matches_oracle(gpu):
    e = W.lorenz(512, seed=0)
    res = gs.solve(P.DiffEqGPULorenz(), e.y0, e.params, rtol=1e-8, atol=1e-8, device="cuda")
    assert res.route.actual == "hip_tsit5" and res.solver == "hip_tsit5"
    oyf, _, ona, onr, _ = O.tsit5_record(O.LORENZ, e.y0, e.params, 0.0, 1.0, rtol=1e-8,
                                         atol=1e-8)
    same = res.accepted_steps == ona
    assert same.mean() > 0.999
    assert np.max(np.abs(res.y_final[same] - oyf[same]) / np.abs(oyf[same])) < 1e-10
    assert res.accepted_steps.dtype == np.int64 and res.rejected_steps.shape == (512,)


def test_solve_failed_rows_are_nan(gpu):
    e = W.lorenz(8, seed=1)
    p = P.DiffEqGPULorenz()
    import gradsolve_amd.api as A
    res = A._solve_tsit5(p, e.y0, e.params, rtol=1e-8, atol=1e-8, device="cuda", max_steps=5)
    assert np.all(np.isnan(res.y_final))


@pytest.mark.parametrize("wrt", ["params", "y0", ("y0", "params")])
def test_grad_closure_wrt_forms(gpu, wrt):
    e = W.lorenz(256, seed=2)
    prob = P.DiffEqGPULorenz()
    cl = gs.grad_closure(prob, e.y0, e.params, wrt=wrt, rtol=1e-8, atol=1e-8, device="cuda")
    assert cl.route.actual == "hip_replay"
    y0 = torch.tensor(e.y0, device="cuda", requires_grad=True)
    p = torch.tensor(e.params, device="cuda", requires_grad=True)
    if wrt == "params":
        out = cl(p)
    elif wrt == "y0":
        out = cl(y0)
    else:
        out = cl(y0, p)
    (out * out).sum().backward()
    tape = cl.dts_padded()
    oyf = O.tsit5_replay(O.LORENZ, e.y0, e.params, 0.0, tape)
    og0, ogp = O.tsit5_adjoint(O.LORENZ, e.y0, e.params, 0.0, tape, 2.0 * oyf)
    if wrt in ("params", ("y0", "params")):
        assert np.max(np.abs(p.grad.cpu().numpy() - ogp) / np.abs(ogp)) < 1e-10
    if wrt in ("y0", ("y0", "params")):
        g = y0.grad.cpu().numpy()
        assert np.max(np.abs(g - og0) / np.abs(og0).max(axis=1, keepdims=True)) < 1e-10


def test_closure_replays_frozen_mesh_at_new_inputs(gpu):
    e = W.lorenz(128, seed=3)
    prob = P.DiffEqGPULorenz()
    cl = gs.grad_closure(prob, e.y0, e.params, wrt=("y0", "params"), rtol=1e-8, atol=1e-8,
                         device="cuda")
    tape = cl.dts_padded()
    p2 = e.params * 1.001
    y2 = e.y0 + 1e-3
    yt = torch.tensor(y2, device="cuda", requires_grad=True)
    pt = torch.tensor(p2, device="cuda", requires_grad=True)
    out = cl(yt, pt)
    oyf = O.tsit5_replay(O.LORENZ, y2, p2, 0.0, tape)
    assert np.max(np.abs(out.detach().cpu().numpy() - oyf) / np.abs(oyf)) < 1e-12
    (out * out).sum().backward()
    og0, ogp = O.tsit5_adjoint(O.LORENZ, y2, p2, 0.0, tape, 2.0 * oyf)
    assert np.max(np.abs(pt.grad.cpu().numpy() - ogp) / np.abs(ogp)) < 1e-10


def test_lorenz_three_params_gradients(gpu):
    """The 'lorenz' problem differentiates all three columns [sigma, beta, rho]."""
    e = W.lorenz(64, seed=4)
    params = np.concatenate([np.full((64, 1), 10.0), np.full((64, 1), 8.0 / 3.0), e.params],
                            axis=1)
    cl = gs.grad_closure(P.Lorenz(), e.y0, params, rtol=1e-8, atol=1e-8, device="cuda")
    p = torch.tensor(params, device="cuda", requires_grad=True)
    out = cl(p)
    (out * out).sum().backward()
    tape = cl.dts_padded()
    oyf = O.tsit5_replay(O.LORENZ3, e.y0, params, 0.0, tape)
    og0, ogp = O.tsit5_adjoint(O.LORENZ3, e.y0, params, 0.0, tape, 2.0 * oyf)
    assert np.max(np.abs(p.grad.cpu().numpy() - ogp) / np.abs(ogp)) < 1e-10


def test_robertson_grad_closure(gpu):
    e = W.robertson(256, seed=5)
    cl = gs.grad_closure(P.Robertson(), e.y0, e.params, rtol=1e-6, atol=1e-10, device="cuda")
    assert cl.route.actual == "hip_rosenbrock23"
    p = torch.tensor(e.params, device="cuda", requires_grad=True)
    out = cl(p)
    (out * torch.tensor([1.0, 10.0, 1.0], device="cuda", dtype=torch.float64)).sum().backward()
    tape = cl.dts_padded()
    og0, ogp = O.rb23_adjoint(O.ROBERTSON, e.y0, e.params, 0.0, tape,
                              np.tile([1.0, 10.0, 1.0], (256, 1)))
    assert np.max(np.abs(p.grad.cpu().numpy() - ogp) / np.abs(ogp)) < 1e-9
    res = gs.solve(P.Robertson(), e.y0, e.params, rtol=1e-6, atol=1e-10, device="cuda")
    assert res.route.actual == "hip_rosenbrock23"
    assert np.array_equal(res.accepted_steps, cl.n_acc)


def test_exhausted_max_steps_raises(gpu):
    e = W.lorenz(4, seed=6)
    from gradsolve_amd import fields as F
    from gradsolve_amd.closure import make_tsit5_closure
    with pytest.raises(RuntimeError, match="exhausted max_steps"):
        make_tsit5_closure(F.binding_for(P.DiffEqGPULorenz()), P.DiffEqGPULorenz(), e.y0,
                           e.params, wrt=("params",), rtol=1e-8, atol=1e-8,
                           device=torch.device("cuda:0"), max_steps=5)


def test_engine_aliases_route_to_native(gpu):
    e = W.lorenz(16, seed=7)
    res = gs.solve(P.DiffEqGPULorenz(), e.y0, e.params, engine="cuda_tsit5", rtol=1e-8,
                   atol=1e-8)
    assert res.route.actual == "hip_tsit5" and res.route.reason == "alias:cuda_tsit5"
    cl = gs.grad_closure(P.DiffEqGPULorenz(), e.y0, e.params, engine="warp_replay",
                         rtol=1e-8, atol=1e-8)
    assert cl.route.actual == "hip_replay"


def test_float32_closure_against_float64_replay_of_its_tape(gpu):
    """precision='float32' (tests/test_f32_replay.py:88-94 of the reference): the FP32 lane's
    gradient against the FP64 replay adjoint of the same FP32-recorded tape, rel-L2 <= 1e-3
    and cosine >= 1 - 1e-6; the primal within 1e-4. Lorenz rho = 28, n = 2048, rtol 1e-6,
    atol 1e-9, y0 = [1,0,0] + 0.01 N(0,1) (seed 0)."""
    n = 2048
    rng = np.random.default_rng(0)
    y0 = np.tile(np.array([1.0, 0.0, 0.0]), (n, 1)) + 0.01 * rng.standard_normal((n, 3))
    params = np.full((n, 1), 28.0)
    prob = P.DiffEqGPULorenz()
    cl = gs.grad_closure(prob, y0, params, wrt="params", rtol=1e-6, atol=1e-9,
                         precision="float32", device="cuda")
    p = torch.tensor(params, device="cuda", requires_grad=True)
    yf = cl(p)
    (yf * yf).sum().backward()
    g32 = p.grad.cpu().numpy()[:, 0]
    tape = cl.dts_padded()
    oyf = O.tsit5_replay(O.LORENZ, y0, params, 0.0, tape)
    _, ogp = O.tsit5_adjoint(O.LORENZ, y0, params, 0.0, tape, 2.0 * oyf)
    g64 = ogp[:, 0]
    rel_l2 = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    cos = g32 @ g64 / (np.linalg.norm(g32) * np.linalg.norm(g64))
    prim = np.max(np.abs(yf.detach().cpu().numpy() - oyf) / np.abs(oyf).max(axis=1,
                                                                          keepdims=True))
    print(f"float32 lane vs float64 replay of its tape: grad rel-L2 {rel_l2:.2e}, cosine "
          f"1-{1 - cos:.1e}, primal {prim:.2e}", flush=True)
    assert rel_l2 <= 1e-3 and cos >= 1 - 1e-6 and prim <= 1e-4
    res = gs.solve(prob, y0, params, rtol=1e-6, atol=1e-9, precision="float32", device="cuda")
    assert np.max(np.abs(res.y_final - oyf) / np.abs(oyf).max(axis=1, keepdims=True)) < 1e-4


def test_float32_rejected_on_the_rodas5p_lane(gpu):
    """Rodas5P is float64 throughout; float32 raises naming the precision, as the reference
    does for its f64-only lanes (gradsolve/api.py:262-274)."""
    e = W.robertson(4, seed=0)
    with pytest.raises(ValueError, match="float32"):
        gs.grad_closure(P.Robertson(), e.y0, e.params, engine="rodas5p_replay",
                        precision="float32", device="cuda")
