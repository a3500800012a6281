"""The Kuramoto engine (D = 32, coupling on FP64 MFMA, 16 trajectories per wavefront) against
the oracle: forward (counts, teacher-forced tape entries, final state replayed on the same
tape), saveat (bracket-and-re-step on the device tape), the adjoint with save-time and
final-state cotangents, central finite differences through the device replay, the public API
route, and ragged ensemble sizes.

Workload: K ~ U(0, 2) 32 x 32 with zero diagonal (shared), omega ~ N(0, 1), theta0 ~ U(0, 2 pi),
t in [0, 10], rtol = atol = 1e-8, 64 save times linspace(0, 10, 64).

Tolerances (the coupling sums run in MFMA order with fused multiply-adds and the device sin/cos
is within 1 ulp of the C library, so the device and the oracle round differently):
teacher-forced step sizes 1e-6 relative (the controller amplifies rounding through the error
estimate's cancellation); states and saves replayed on the same tape 1e-10 relative to the
row's largest entry; gradients 1e-8.
"""
import numpy as np
import pytest
import torch

import gradsolve_amd as gs
from gradsolve_amd import kuramoto as KU
from gradsolve_amd import problems as P
from gradsolve_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TS = np.linspace(0.0, 10.0, 64)
TF_REL = 1e-6
STATE_REL = 1e-10
GRAD_REL = 1e-8


def _setup(n, seed, dev):
    e = W.kuramoto32(n, seed=seed)
    K = torch.as_tensor(e.aux, device=dev)
    y0 = torch.as_tensor(np.ascontiguousarray(e.y0.T), device=dev)
    om = torch.as_tensor(np.ascontiguousarray(e.params.T), device=dev)
    return e, K, y0, om


def _rowrel(a, b):
    return np.max(np.abs(a - b).reshape(a.shape[0], -1), axis=1) / \
        np.max(np.abs(b).reshape(b.shape[0], -1), axis=1)


def _record(e, K, y0, om, dev, ts=TS):
    tsd = None if ts is None else torch.as_tensor(ts, device=dev)
    rec = KU.record(K, y0, om, t0=0.0, t1=10.0, rtol=1e-8, atol=1e-8, ts=tsd)
    torch.cuda.synchronize()
    return rec


def test_forward_against_oracle(gpu):
    n = 1000
    print(f"case kuramoto forward n={n}", flush=True)
    e, K, y0, om = _setup(n, 0, gpu)
    rec = _record(e, K, y0, om, gpu)
    assert int((rec.status != 0).sum().item()) == 0
    na = rec.n_acc.cpu().numpy().astype(np.int64)
    nr = rec.n_rej.cpu().numpy().astype(np.int64)
    oyf, _, ona, onr, ost = O.tsit5_record(O.KURAMOTO, e.y0, e.params, 0.0, 10.0, rtol=1e-8,
                                           atol=1e-8, aux=e.aux)
    same = (na == ona) & (nr == onr)
    print(f"accepted steps mean {na.mean():.1f} max {na.max()}; counts equal to the oracle "
          f"{same.mean():.4f}", flush=True)
    assert same.mean() > 0.99
    tape = rec.padded_tape().cpu().numpy()
    ho, nro, rej_at = O.tsit5_teacher_forced(O.KURAMOTO, e.y0, e.params, 0.0, 10.0, tape, na,
                                             rtol=1e-8, atol=1e-8, aux=e.aux)
    nz = tape != 0
    rel = np.abs(ho[nz] - tape[nz]) / tape[nz]
    print(f"teacher-forced step sizes: max rel {rel.max():.2e}, median {np.median(rel):.2e}",
          flush=True)
    assert rel.max() < TF_REL
    ryf = O.tsit5_replay(O.KURAMOTO, e.y0, e.params, 0.0, tape, aux=e.aux)
    yf = rec.y_final.t().cpu().numpy()
    d = _rowrel(yf, ryf)
    print(f"final state vs oracle replay of the device tape: max {d.max():.2e}", flush=True)
    assert d.max() < STATE_REL
    # recorded start-of-step states: step s of trajectory j is the replay of its first s steps
    states = rec.states.cpu().numpy()
    for s in (0, 7, int(na.min()) - 1):
        ry = O.tsit5_replay(O.KURAMOTO, e.y0, e.params, 0.0, tape[:, :s], aux=e.aux)
        assert np.max(_rowrel(states[s], ry)) < STATE_REL


def test_saveat_and_adjoint_against_oracle(gpu):
    n = 500
    print(f"case kuramoto saveat+adjoint n={n}", flush=True)
    e, K, y0, om = _setup(n, 1, gpu)
    rec = _record(e, K, y0, om, gpu)
    ys = KU.saveat(rec)
    torch.cuda.synchronize()
    tape = rec.padded_tape().cpu().numpy()
    ryf, rys = O.tsit5_replay_saveat(O.KURAMOTO, e.y0, e.params, 0.0, tape, TS, aux=e.aux)
    ysn = ys.cpu().numpy()
    yf = rec.y_final.t().cpu().numpy()
    assert np.array_equal(ysn[:, 0], e.y0)       # ts[0] = t0: the save is y0
    assert np.array_equal(ysn[:, -1], yf)        # ts[-1] = t1: the save is y_final
    d = _rowrel(ysn, rys)
    print(f"saves vs oracle on the device tape: max {d.max():.2e}", flush=True)
    assert d.max() < STATE_REL
    rng = np.random.default_rng(5)
    target = rng.standard_normal(ysn.shape)
    ysbar = 2.0 * (ysn - target) / ysn.size       # d mean((ys - target)^2) / d ys
    yfbar = rng.standard_normal(yf.shape)
    g0, gom = KU.adjoint(rec, ysbar=torch.as_tensor(ysbar, device=gpu),
                         yfbar=torch.as_tensor(np.ascontiguousarray(yfbar.T), device=gpu))
    torch.cuda.synchronize()
    og0, ogom = O.tsit5_adjoint_saveat(O.KURAMOTO, e.y0, e.params, 0.0, tape, TS, ysbar,
                                       yfbar, aux=e.aux)
    r0 = _rowrel(g0.t().cpu().numpy(), og0)
    rp = _rowrel(gom.t().cpu().numpy(), ogom)
    print(f"adjoint vs oracle: y0bar max {r0.max():.2e}, omega_bar max {rp.max():.2e}",
          flush=True)
    assert r0.max() < GRAD_REL and rp.max() < GRAD_REL
    # final-state cotangent alone
    g0f, gomf = KU.adjoint(rec, yfbar=torch.as_tensor(np.ascontiguousarray(yfbar.T), device=gpu))
    og0f, ogomf = O.tsit5_adjoint(O.KURAMOTO, e.y0, e.params, 0.0, tape, yfbar, aux=e.aux)
    assert _rowrel(gomf.t().cpu().numpy(), ogomf).max() < GRAD_REL
    assert _rowrel(g0f.t().cpu().numpy(), og0f).max() < GRAD_REL


def test_grad_closure_route_and_finite_differences(gpu):
    n = 4
    e = W.kuramoto32(n, seed=2)
    prob = P.Kuramoto(e.aux)
    print("case kuramoto closure n=4", flush=True)
    cl = gs.grad_closure(prob, e.y0, e.params, wrt="params", saveat=TS, rtol=1e-8, atol=1e-8,
                         device="cuda")
    assert cl.route.actual == "hip_kuramoto"
    rng = np.random.default_rng(9)
    w = torch.as_tensor(rng.standard_normal((n, TS.size, 32)), device="cuda")
    p = torch.tensor(e.params, device="cuda", requires_grad=True)
    ys = cl(p)
    assert tuple(ys.shape) == (n, TS.size, 32)
    (ys * w).sum().backward()
    g = p.grad.cpu().numpy()
    # central differences through the device replay of the frozen mesh
    j = 1
    fd = np.zeros(32)
    for i in range(32):
        eps = 1e-6 * max(1.0, abs(e.params[j, i]))
        up, dn = e.params.copy(), e.params.copy()
        up[j, i] += eps
        dn[j, i] -= eps
        lu = (cl(torch.as_tensor(up, device="cuda")) * w).sum().item()
        ld = (cl(torch.as_tensor(dn, device="cuda")) * w).sum().item()
        fd[i] = (lu - ld) / (2 * eps)
    rel = np.linalg.norm(fd - g[j]) / np.linalg.norm(g[j])
    print(f"omega gradient vs central differences on the frozen mesh: rel-L2 {rel:.2e}",
          flush=True)
    assert rel < 1e-6
    res = gs.solve(prob, e.y0, e.params, saveat=TS, rtol=1e-8, atol=1e-8, device="cuda")
    assert res.route.actual == "hip_kuramoto" and res.y_saved.shape == (n, TS.size, 32)
    assert np.array_equal(res.ts_saved, TS)
    assert np.array_equal(res.y_saved[:, -1], res.y_final)
    with pytest.raises(ValueError, match="sorted"):
        gs.solve(prob, e.y0, e.params, saveat=TS[::-1], device="cuda")


@pytest.mark.parametrize("n", [1, 17, 33])
def test_ragged_sizes_and_save_edges(gpu, n):
    """Ensembles that do not fill the last wavefront; duplicate save times, saves at t0 and
    at t1, and saves on either side of a recorded step boundary."""
    print(f"case kuramoto ragged n={n}", flush=True)
    e, K, y0, om = _setup(n, 3, gpu)
    ts = np.array([0.0, 0.0, 1e-3, 2.5, 2.5, 7.0, 10.0, 10.0])
    rec = _record(e, K, y0, om, gpu, ts=ts)
    ys = KU.saveat(rec).cpu().numpy()
    tape = rec.padded_tape().cpu().numpy()
    _, rys = O.tsit5_replay_saveat(O.KURAMOTO, e.y0, e.params, 0.0, tape, ts, aux=e.aux)
    assert np.max(_rowrel(ys, rys)) < STATE_REL
    assert np.array_equal(ys[:, 0], e.y0) and np.array_equal(ys[:, 1], e.y0)
    yf = rec.y_final.t().cpu().numpy()
    assert np.array_equal(ys[:, 6], yf) and np.array_equal(ys[:, 7], yf)
    assert np.array_equal(ys[:, 3], ys[:, 4])
    ysbar = np.random.default_rng(n).standard_normal(ys.shape)
    g0, gom = KU.adjoint(rec, ysbar=torch.as_tensor(ysbar, device=gpu))
    og0, ogom = O.tsit5_adjoint_saveat(O.KURAMOTO, e.y0, e.params, 0.0, tape, ts, ysbar,
                                       aux=e.aux)
    assert _rowrel(gom.t().cpu().numpy(), ogom).max() < GRAD_REL
    assert _rowrel(g0.t().cpu().numpy(), og0).max() < GRAD_REL


def test_fp32_mfma_variant_error_against_fp64(gpu):
    """FP32 variant (FP32 stages and coupling on FP32 MFMA, compensated state, FP64 t, h, norm
    and controller): final states and saves against the FP64 oracle's own solution, and the
    gradient against the FP64 oracle adjoint on the FP32-recorded mesh (the reference's f32
    acceptance, tests/test_f32_replay.py:88-94: rel-L2 <= 1e-3, cosine >= 1 - 1e-6). The errors
    are printed; the FP64 oracle's discretization error (rtol 1e-8 against rtol 1e-11) is
    printed beside them as the scale."""
    n = 256
    print(f"case kuramoto fp32 n={n}", flush=True)
    e, K, y0, om = _setup(n, 4, gpu)
    tsd = torch.as_tensor(TS, device=gpu)
    rec = KU.record(K, y0, om, t0=0.0, t1=10.0, rtol=1e-8, atol=1e-8, ts=tsd, fp32=True)
    ys = KU.saveat(rec)
    torch.cuda.synchronize()
    assert int((rec.status != 0).sum().item()) == 0
    na = rec.n_acc.cpu().numpy()
    oyf, otape, ona, _, _ = O.tsit5_record(O.KURAMOTO, e.y0, e.params, 0.0, 10.0, rtol=1e-8,
                                           atol=1e-8, aux=e.aux)
    _, oys = O.tsit5_replay_saveat(O.KURAMOTO, e.y0, e.params, 0.0, otape, TS, aux=e.aux)
    fyf, ftape, _, _, _ = O.tsit5_record(O.KURAMOTO, e.y0, e.params, 0.0, 10.0, rtol=1e-11,
                                         atol=1e-11, aux=e.aux)
    yf32 = rec.y_final.t().cpu().numpy()
    r_state = _rowrel(yf32, oyf)
    r_saves = _rowrel(ys.cpu().numpy(), oys)
    r_disc = _rowrel(oyf, fyf)
    print(f"FP32 vs FP64 oracle: final state max {r_state.max():.2e} (median "
          f"{np.median(r_state):.2e}), saves max {r_saves.max():.2e}; FP64 rtol 1e-8 vs 1e-11: "
          f"max {r_disc.max():.2e}; accepted steps FP32 {na.mean():.1f} vs FP64 {ona.mean():.1f}",
          flush=True)
    assert r_state.max() < 1e-4 and r_saves.max() < 1e-4
    rng = np.random.default_rng(6)
    ysn = ys.cpu().numpy()
    ysbar = 2.0 * (ysn - rng.standard_normal(ysn.shape)) / ysn.size
    g0, gom = KU.adjoint(rec, ysbar=torch.as_tensor(ysbar, device=gpu))
    tape = rec.padded_tape().cpu().numpy()
    _, ogom = O.tsit5_adjoint_saveat(O.KURAMOTO, e.y0, e.params, 0.0, tape, TS, ysbar,
                                     aux=e.aux)
    g = gom.t().cpu().numpy().ravel()
    og = ogom.ravel()
    rel_l2 = np.linalg.norm(g - og) / np.linalg.norm(og)
    cos = g @ og / (np.linalg.norm(g) * np.linalg.norm(og))
    print(f"FP32 omega gradient vs FP64 adjoint on the FP32 mesh: rel-L2 {rel_l2:.2e}, cosine "
          f"1-{1 - cos:.1e}", flush=True)
    assert rel_l2 <= 1e-3 and cos >= 1 - 1e-6
    # the public route
    cl = gs.grad_closure(P.Kuramoto(e.aux), e.y0[:8], e.params[:8], wrt="params", saveat=TS,
                         rtol=1e-8, atol=1e-8, precision="float32", device="cuda")
    assert cl.route.actual == "hip_kuramoto" and cl.rec.flags != 0


def test_phases_outside_the_fast_sincos_range(gpu):
    """Trajectories whose phases exceed 2^20 (one wavefront of three, all 32 phases shifted by
    2^21) take the range-checked sin/cos (a wavefront-uniform redo of the field's products) in
    the forward, the saves and both VJP forms of the adjoint; the other wavefronts keep the
    fast path. Everything is compared with the oracle (C library sin/cos) as above; the large
    phases loosen the step-size control (scale atol + rtol |theta| ~ 2e-2), so counts are few."""
    n = 48
    print(f"case kuramoto large phases n={n}", flush=True)
    e, _, _, om = _setup(n, 6, gpu)
    y0n = e.y0.copy()
    y0n[16:32] += 2.0 ** 21
    K = torch.as_tensor(e.aux, device=gpu)
    y0 = torch.as_tensor(np.ascontiguousarray(y0n.T), device=gpu)
    ts = np.linspace(0.0, 10.0, 9)
    rec = _record(e, K, y0, om, gpu, ts=ts)
    assert int((rec.status != 0).sum().item()) == 0
    na = rec.n_acc.cpu().numpy().astype(np.int64)
    nr = rec.n_rej.cpu().numpy().astype(np.int64)
    _, _, ona, onr, _ = O.tsit5_record(O.KURAMOTO, y0n, e.params, 0.0, 10.0, rtol=1e-8,
                                       atol=1e-8, aux=e.aux)
    same = (na == ona) & (nr == onr)
    print(f"accepted steps: shifted {na[16:32].mean():.1f}, others {na[:16].mean():.1f}; "
          f"counts equal to the oracle {same.mean():.4f}", flush=True)
    assert same.mean() > 0.99
    tape = rec.padded_tape().cpu().numpy()
    ryf, rys = O.tsit5_replay_saveat(O.KURAMOTO, y0n, e.params, 0.0, tape, ts, aux=e.aux)
    yf = rec.y_final.t().cpu().numpy()
    ys = KU.saveat(rec).cpu().numpy()
    # relative to the row's largest entry: 1e-10 of 2^21 is 2e-4 rad, so the shifted rows are
    # compared on the phase differences from their first phase instead
    dfin = _rowrel(yf - yf[:, :1], ryf - ryf[:, :1])
    dsav = _rowrel(ys - ys[:, :, :1], rys - rys[:, :, :1])
    print(f"final phase differences vs oracle replay: max {dfin.max():.2e}; saves "
          f"{dsav.max():.2e}", flush=True)
    assert dfin.max() < STATE_REL and dsav.max() < STATE_REL
    ysbar = np.random.default_rng(7).standard_normal(ys.shape) / ys.size
    g0, gom = KU.adjoint(rec, ysbar=torch.as_tensor(ysbar, device=gpu))
    og0, ogom = O.tsit5_adjoint_saveat(O.KURAMOTO, y0n, e.params, 0.0, tape, ts, ysbar,
                                       aux=e.aux)
    r0 = _rowrel(g0.t().cpu().numpy(), og0)
    rp = _rowrel(gom.t().cpu().numpy(), ogom)
    print(f"adjoint vs oracle: y0bar max {r0.max():.2e}, omega_bar max {rp.max():.2e}",
          flush=True)
    assert r0.max() < GRAD_REL and rp.max() < GRAD_REL
