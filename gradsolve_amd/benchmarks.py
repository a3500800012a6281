"""Throughput of the engines on each workload (one JSON line per workload).

    python -m gradsolve_amd.benchmarks --workload uneven|robertson|lorenz1m|kuramoto|wide|all

Every measurement: inputs resident in HBM, buffers preallocated, one warm-up, then the best
of ``--repeat`` timed calls with device synchronisation on both sides (the reference's
benchmark protocol, gradsolve benchmarks/reverse_mode_vs_diffrax.py:88-98). Reported:
trajectories/s and accepted steps/s for the forward and for forward + gradient, and the
algorithmic FP64 TFLOP/s of the forward + adjoint against the 78.6 TFLOP/s peak.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import _native as N
from . import autotune as AT
from . import hip_tsit5 as H
from . import persist as PS
from . import workcounts as WC
from . import workloads as W


def _best(fn, repeat):
    fn()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(repeat):
        torch.cuda.synchronize()
        t = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t)
    return best


def _soa(a, dev):
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a, np.float64).T), device=dev)


def _log(msg):
    print(f"[benchmarks] {msg}", file=sys.stderr, flush=True)


def uneven(args, dev):
    """Uneven lengths: Lorenz, N = 4,194,304, rho ~ U(0, 60), t1 ~ U(0.5, 10), rtol = atol =
    1e-6: persistent work-queue launch (autotuned, FP64 and FP32, graph-captured) against
    the one-trajectory-per-thread launch."""
    e = W.lorenz_uneven(args.n or (1 << 22), seed=0)
    n = e.y0.shape[0]
    y0, p = _soa(e.y0, dev), _soa(e.params, dev)
    t1 = torch.as_tensor(e.t1_per, device=dev)
    kw = dict(t0=0.0, t1=float(e.t1_per.max()), rtol=1e-6, atol=1e-6, max_steps=100000,
              t1_per=t1)
    out = {"workload": "lorenz_uneven", "n": n, "rtol": 1e-6}
    for fp32 in (False, True):
        tag = "fp32" if fp32 else "fp64"
        ch = AT.tune(N.FIELD_LORENZ, y0, p, fp32=fp32, sample=args.tune_sample, **kw)
        _log(f"{tag} autotune: {ch.table} -> waves {ch.waves_per_simd}, C {ch.ckpt_interval}")
        est = PS.estimate_steps(N.FIELD_LORENZ, y0, p, t0=0.0, t1=kw["t1"], t1_per=t1,
                                rtol=1e-6, atol=1e-6)
        eng = PS.PersistentTsit5(N.FIELD_LORENZ, 3, 1, n, est=est,
                                 waves_per_simd=ch.waves_per_simd,
                                 ckpt_interval=ch.ckpt_interval, fp32=fp32, device=dev)
        eng.plan_longest_first(t1)
        eng.run_forward(y0, p, **kw)
        torch.cuda.synchronize()
        acc = int(eng.n_acc.to(torch.int64).sum().item())
        trials = acc + int(eng.n_rej.to(torch.int64).sum().item())

        def fwd():
            eng.forward(y0, p, **kw)

        def both():
            eng.forward(y0, p, **kw)
            torch.mul(eng.y_final, 2.0, out=eng.ybar)
            eng.adjoint(p, eng.ybar, t0=0.0)

        tf = _best(fwd, args.repeat)
        tb = _best(both, args.repeat)
        cap = PS.CapturedStep(eng, y0, p, lambda yf, o: torch.mul(yf, 2.0, out=o), **kw)
        tg = _best(cap.replay, args.repeat)
        bad = int((eng.status != 0).sum().item())
        flops = WC.LORENZ_TSIT5_TRIAL_FLOPS * trials + WC.LORENZ_TSIT5_ADJ_STEP_FLOPS * acc
        out[f"persistent_{tag}"] = {
            "waves_per_simd": ch.waves_per_simd, "ckpt_interval": ch.ckpt_interval,
            "autotune_ms": ch.table, "lanes": eng.lanes, "workspace_GB": round(eng.bytes() / 1e9, 2),
            "fwd_ms": round(tf * 1e3, 3), "fwd_adj_ms": round(tb * 1e3, 3),
            "graph_fwd_adj_ms": round(tg * 1e3, 3),
            "fwd_traj_per_s": round(n / tf, 1), "fwd_adj_traj_per_s": round(n / tb, 1),
            "fwd_acc_steps_per_s": round(acc / tf, 1),
            "fwd_adj_acc_steps_per_s": round(acc / tb, 1),
            "fwd_adj_tflops": round(flops / tb / 1e12, 3),
            "roofline_frac": round(flops / tb / 1e12 / WC.PEAK_FP64_TFLOPS, 4),
            "mean_accepted": round(acc / n, 2),
            "min_accepted": int(eng.n_acc.min().item()), "max_accepted": int(eng.n_acc.max().item()),
            "status_nonzero": bad}
        _log(f"{tag}: {out[f'persistent_{tag}']}")
        del eng, cap
        torch.cuda.empty_cache()
    # one trajectory per thread, same arithmetic
    est = PS.estimate_steps(N.FIELD_LORENZ, y0, p, t0=0.0, t1=kw["t1"], t1_per=t1, rtol=1e-6,
                            atol=1e-6)
    cap = est.max_steps * 2 + 64
    # its own best checkpoint interval (the pair adjoint at C = 2 included)
    per_c = {}
    bad_pt = 0
    for C in (2, 4, 8, 16):
        ws = H.Workspace(N.FIELD_LORENZ, 3, 1, n, cap, C, dev)

        def both_plain():
            ws.forward(y0, p, t0=0.0, t1=kw["t1"], rtol=1e-6, atol=1e-6, max_steps=100000,
                       t1_per=t1)
            torch.mul(ws.y_final, 2.0, out=ws.ybar)
            ws.adjoint(p, ws.ybar, t0=0.0)

        per_c[C] = _best(both_plain, args.repeat)
        bad_pt = max(bad_pt, int((ws.status != 0).sum().item()))
        del ws
    C_pt = min(per_c, key=per_c.get)
    tb = per_c[C_pt]
    out["per_trajectory_fp64"] = {"fwd_adj_ms": round(tb * 1e3, 3),
                                  "fwd_adj_traj_per_s": round(n / tb, 1),
                                  "ckpt_interval": C_pt,
                                  "fwd_adj_ms_by_ckpt": {c: round(v * 1e3, 3)
                                                         for c, v in per_c.items()},
                                  "status_nonzero": bad_pt}
    out["persistent_speedup"] = round(tb * 1e3 / out["persistent_fp64"]["fwd_adj_ms"], 3)
    return out


def robertson(args, dev):
    """Stiff: Robertson, N = 262,144, Rosenbrock23 rtol 1e-6 atol 1e-10, t in [0, 100];
    gradient of sum(y_final * [1, 10, 1]) w.r.t. k1..k3. The same workload on the Rodas5P
    lane is reported beside it (``rodas5p``)."""
    out = _stiff_lane(args, dev, "rb23", "robertson_rb23")
    out["rb23_fp32"] = _stiff_lane(args, dev, "rb23", "robertson_rb23_fp32", flags=N.FP32)
    out["rodas5p"] = _stiff_lane(args, dev, "rodas5p", "robertson_rodas5p")
    return out


def _stiff_lane(args, dev, method, name, flags=0):
    fwd_entry, adj_entry = f"gsx_{method}_fwd", f"gsx_{method}_adjoint"
    e = W.robertson(args.n or (1 << 18), seed=0)
    n = e.y0.shape[0]
    y0, p = _soa(e.y0, dev), _soa(e.params, dev)
    rec = H.record(N.FIELD_ROBERTSON, y0, p, t0=0.0, t1=100.0, rtol=1e-6, atol=1e-10,
                   max_steps=50000, tape_cap=1024, ckpt_interval=8, entry=fwd_entry,
                   fp32=bool(flags & N.FP32))
    torch.cuda.synchronize()
    S = rec.max_steps_taken()
    acc = int(rec.n_acc.to(torch.int64).sum().item())
    rej = int(rec.n_rej.to(torch.int64).sum().item())
    out = {"workload": name, "n": n, "mean_accepted": round(acc / n, 2),
           "mean_rejected": round(rej / n, 2), "max_accepted": S}
    wvec = torch.tensor([1.0, 10.0, 1.0], dtype=torch.float64, device=dev)[:, None]
    best = None
    # C = 2 is the Rosenbrock23 pair adjoint (FP64); the Rodas5P lane and the FP32
    # Rosenbrock23 lane have segment adjoints only
    for C in ((2, 4, 8, 16) if method == "rb23" and not flags else (4, 8, 16)):
        ws = H.Workspace(N.FIELD_ROBERTSON, 3, 3, n, S + 16, C, dev)
        ybar = (wvec * torch.ones((3, n), dtype=torch.float64, device=dev)).contiguous()

        def fwd():
            N.call(fwd_entry, N.FIELD_ROBERTSON, n, 3, 3, None, y0.data_ptr(),
                   p.data_ptr(), 0.0, 100.0, None, 1e-6, 1e-10, 50000, ws.y_final.data_ptr(),
                   ws.n_acc.data_ptr(), ws.n_rej.data_ptr(), ws.status.data_ptr(),
                   ws.tape.data_ptr(), ws.cap, ws.ckpt.data_ptr(), C, flags,
                   H._stream_handle(None))

        def both():
            fwd()
            N.call(adj_entry, N.FIELD_ROBERTSON, n, 3, 3, None, None, p.data_ptr(),
                   0.0, ws.tape.data_ptr(), ws.cap, ws.n_acc.data_ptr(), ws.ckpt.data_ptr(), C,
                   ybar.data_ptr(), ws.y0bar.data_ptr(), ws.pbar.data_ptr(), flags,
                   H._stream_handle(None))

        tf = _best(fwd, args.repeat)
        tb = _best(both, args.repeat)
        out[f"C{C}"] = {"fwd_ms": round(tf * 1e3, 3), "fwd_adj_ms": round(tb * 1e3, 3)}
        if best is None or tb < best[0]:
            best = (tb, tf, C)
        del ws
    tb, tf, C = best
    trial_fl, adj_fl = ((WC.ROBERTSON_RB23_TRIAL_FLOPS, WC.ROBERTSON_RB23_ADJ_STEP_FLOPS)
                        if method == "rb23" else
                        (WC.ROBERTSON_RODAS5P_TRIAL_FLOPS, WC.ROBERTSON_RODAS5P_ADJ_STEP_FLOPS))
    fl_fwd = trial_fl * (acc + rej)
    fl_all = fl_fwd + adj_fl * acc
    out.update({"ckpt_interval": C, "fwd_traj_per_s": round(n / tf, 1),
                "fwd_adj_traj_per_s": round(n / tb, 1),
                "fwd_acc_steps_per_s": round(acc / tf, 1),
                "fwd_adj_acc_steps_per_s": round(acc / tb, 1),
                "fwd_tflops": round(fl_fwd / tf / 1e12, 3),
                "fwd_adj_tflops": round(fl_all / tb / 1e12, 3),
                "roofline_frac": round(fl_all / tb / 1e12 / WC.PEAK_FP64_TFLOPS, 4)})
    return out


def lorenz1m(args, dev):
    """The headline workload's forward alone and forward + gradient (FP64 and FP32)."""
    e = W.lorenz_1m(seed=0, n=args.n or (1 << 20))
    n = e.y0.shape[0]
    y0, p = _soa(e.y0, dev), _soa(e.params, dev)
    kw = dict(t0=0.0, t1=1.0, rtol=1e-8, atol=1e-8, max_steps=100000)
    rec = H.record(N.FIELD_LORENZ, y0, p, tape_cap=256, ckpt_interval=8, **kw)
    torch.cuda.synchronize()
    cap = rec.max_steps_taken() + 16
    acc = int(rec.n_acc.to(torch.int64).sum().item())
    out = {"workload": "lorenz_1m", "n": n, "mean_accepted": round(acc / n, 3)}
    for tag, flags in (("fp64", 0), ("fp32", N.FP32), ("fp64_strict", N.STRICT)):
        ws = H.Workspace(N.FIELD_LORENZ, 3, 1, n, cap, 4, dev)
        ws.flags = flags

        def fwd_only():
            ws.forward(y0, p, record=False, **kw)

        def both():
            ws.forward(y0, p, **kw)
            torch.mul(ws.y_final, 2.0, out=ws.ybar)
            ws.adjoint(p, ws.ybar, t0=0.0)

        tf = _best(fwd_only, args.repeat)
        tb = _best(both, args.repeat)
        out[tag] = {"fwd_ms": round(tf * 1e3, 3), "fwd_adj_ms": round(tb * 1e3, 3),
                    "fwd_traj_per_s": round(n / tf, 1), "fwd_adj_traj_per_s": round(n / tb, 1),
                    "fwd_adj_acc_steps_per_s": round(acc / tb, 1)}
        del ws
    return out


# FP64 MFMA flops of the Kuramoto kernels per trajectory: one field evaluation forms K sin and
# K cos (2 x 2 x 32 x 32); a field VJP adds K^T (v cos), K^T (v sin).
KURA_RHS_FLOPS = 4 * 32 * 32
KURA_TRIAL_FLOPS = 6 * KURA_RHS_FLOPS              # k2..k7 (k1 carried over)
KURA_RESTEP_FLOPS = 6 * KURA_RHS_FLOPS             # k1..k6 of one save re-step
# the adjoint of one step: the couplings of its 6 stage inputs (each once; the stage VJPs reuse
# their diagonal terms) and the 6 transposed products
KURA_VJP_STEP_FLOPS = 6 * KURA_RHS_FLOPS + 6 * KURA_RHS_FLOPS


def kuramoto(args, dev):
    """Kuramoto-32 with saveat: N = 131,072, K ~ U(0, 2) shared, omega ~ N(0, 1),
    theta0 ~ U(0, 2 pi), t in [0, 10], rtol = atol = 1e-8, 64 save times linspace(0, 10, 64);
    loss mean((ys - target)^2) with target ~ N(0, 1) (torch generator seed 1 on the device);
    gradient with respect to omega (and theta0). FP64 (coupling on FP64 MFMA) and the FP32
    variant (coupling on FP32 MFMA)."""
    from . import kuramoto as KU
    e = W.kuramoto32(args.n or (1 << 17), seed=0)
    n = e.y0.shape[0]
    K = torch.as_tensor(e.aux, device=dev)
    y0, om = _soa(e.y0, dev), _soa(e.params, dev)
    ts = torch.as_tensor(np.linspace(0.0, 10.0, 64), device=dev)
    k = int(ts.numel())
    g = torch.Generator(device=dev).manual_seed(1)
    target = torch.randn((n, k, 32), generator=g, dtype=torch.float64, device=dev)
    ys = torch.empty((n, k, 32), dtype=torch.float64, device=dev)
    ysbar = torch.empty_like(ys)
    mu = torch.empty_like(ys)
    y0bar = torch.empty((32, n), dtype=torch.float64, device=dev)
    ombar = torch.empty_like(y0bar)
    scale = 2.0 / ys.numel()
    neg_target = target.mul_(-scale)  # the loss data, scaled once (in place)
    del target
    st = H._stream_handle(None)
    out = {"workload": "kuramoto32_saveat", "n": n, "nsave": k}
    for tag, fp32 in (("fp64", False), ("fp32", True)):
        rec = KU.record(K, y0, om, t0=0.0, t1=10.0, rtol=1e-8, atol=1e-8, ts=ts, fp32=fp32)
        torch.cuda.synchronize()
        assert int((rec.status != 0).sum().item()) == 0
        acc = int(rec.n_acc.to(torch.int64).sum().item())
        trials = acc + int(rec.n_rej.to(torch.int64).sum().item())
        resteps = int((rec.save_step >= 0).sum().item())

        def fwd_record():
            N.call("gsx_kuramoto_fwd", n, 32, K.data_ptr(), y0.data_ptr(), om.data_ptr(), 0.0,
                   10.0, 1e-8, 1e-8, 100000, ts.data_ptr(), k, rec.y_final.data_ptr(),
                   rec.n_acc.data_ptr(), rec.n_rej.data_ptr(), rec.status.data_ptr(),
                   rec.tape.data_ptr(), rec.cap, rec.states.data_ptr(),
                   rec.save_step.data_ptr(), rec.save_t.data_ptr(), rec.flags, st)

        def fwd():
            fwd_record()
            KU.saveat(rec, out=ys)

        def cot():
            # scale (ys - target) in one pass (bit for bit (ys - target) * scale whenever scale
            # is a power of two, as at power-of-two ensemble sizes)
            torch.add(neg_target, ys, alpha=scale, out=ysbar)

        def both():
            fwd()
            cot()
            KU.adjoint(rec, ysbar=ysbar, mu=mu, out=(y0bar, ombar))

        def fwd_only():
            KU.forward_only(K, y0, om, t0=0.0, t1=10.0, rtol=1e-8, atol=1e-8, fp32=fp32)

        tf0 = _best(fwd_only, args.repeat)
        tf = _best(fwd, args.repeat)
        tb = _best(both, args.repeat)
        # per-kernel times inside one forward + gradient, HIP events on the launch stream
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
        torch.cuda.synchronize()
        ev[0].record()
        fwd_record()
        ev[1].record()
        KU.saveat(rec, out=ys)
        ev[2].record()
        cot()
        ev[3].record()
        KU.adjoint(rec, ysbar=ysbar, mu=mu, out=(y0bar, ombar))
        ev[4].record()
        torch.cuda.synchronize()
        ms = [ev[i].elapsed_time(ev[i + 1]) for i in range(4)]
        fl_fwd = KURA_TRIAL_FLOPS * trials + KURA_RHS_FLOPS * n
        fl_sav = KURA_RESTEP_FLOPS * resteps
        fl_adj = KURA_VJP_STEP_FLOPS * (acc + resteps)
        peak = WC.PEAK_FP64_TFLOPS if not fp32 else 2 * WC.PEAK_FP64_TFLOPS
        out[tag] = {
            "mean_accepted": round(acc / n, 2), "max_accepted": rec.max_steps_taken(),
            "mean_trials": round(trials / n, 2),
            "workspace_GB": round((rec.states.numel() + 4 * ys.numel()) * 8 / 1e9, 2),
            "fwd_only_ms": round(tf0 * 1e3, 3), "fwd_saveat_ms": round(tf * 1e3, 3),
            "fwd_adj_ms": round(tb * 1e3, 3),
            "fwd_traj_per_s": round(n / tf, 1), "fwd_adj_traj_per_s": round(n / tb, 1),
            "fwd_adj_acc_steps_per_s": round(acc / tb, 1),
            "kernel_ms": {"kura_fwd_kernel": round(ms[0], 3),
                          "kura_saveat_kernel": round(ms[1], 3), "cotangent": round(ms[2], 3),
                          "kura_adj_kernel": round(ms[3], 3)},
            "mfma_tflops": {"kura_fwd_kernel": round(fl_fwd / ms[0] / 1e9, 2),
                            "kura_saveat_kernel": round(fl_sav / ms[1] / 1e9, 2),
                            "kura_adj_kernel": round(fl_adj / ms[3] / 1e9, 2)},
            "mfma_peak_tflops": peak,
            "mfma_frac_of_peak": round((fl_fwd + fl_sav + fl_adj) / tb / 1e12 / peak, 4)}
        _log(f"kuramoto {tag}: {out[tag]}")
        del rec
        torch.cuda.empty_cache()
    return out


def copy(args, dev):
    """A device copy of 2^27 float64 (1 GiB read, 1 GiB written), repeated: the calibration
    kernel for the HBM byte counters (FETCH_SIZE / WRITE_SIZE read for it against its known
    traffic)."""
    n = args.n or (1 << 27)
    src = torch.ones(n, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    t = _best(lambda: dst.copy_(src), args.repeat)
    return {"workload": "copy", "elements": n, "bytes_read": n * 8, "bytes_written": n * 8,
            "ms": round(t * 1e3, 3), "GB_per_s": round(2 * n * 8 / t / 1e9, 1)}


def wide(args, dev):
    """High-NVAR non-stiff: the periodic Allen-Cahn lattice (workloads.allen_cahn, D = 256 by
    default, N = 16,384, t in [0, 2], rtol = atol = 1e-8) compiled at run time and integrated on
    the wide Tsit5 lane (state and stages in an HBM workspace): the record, and record +
    adjoint of sum(y_final^2), at checkpoint intervals 4, 8 and 16."""
    from . import hip_wide as HW
    from .fields import register_field
    D = args.dim or 256
    e = W.allen_cahn(args.n or (1 << 14), D=D, seed=0)
    n = e.y0.shape[0]
    b = register_field(f"bench_allen_cahn_{D}", W.allen_cahn_rhs, D, 2)
    fid = b.field_id
    y0, p = _soa(e.y0, dev), _soa(e.params, dev)
    kw = dict(t0=e.t0, t1=e.t1, rtol=e.rtol, atol=e.atol)
    rec = HW.record(fid, y0, p, max_steps=100000, ckpt_interval=8, **kw)
    torch.cuda.synchronize()
    assert int((rec.status != 0).sum().item()) == 0
    S = rec.max_steps_taken()
    acc = int(rec.n_acc.to(torch.int64).sum().item())
    rej = int(rec.n_rej.to(torch.int64).sum().item())
    cap = S + 16
    out = {"workload": f"allen_cahn{D}_wide", "n": n, "D": D, "mean_accepted": round(acc / n, 2),
           "mean_rejected": round(rej / n, 2), "max_accepted": S}
    st = H._stream_handle(None)
    ws0 = HW.workspace(0, D, n, dev)
    y_final = torch.empty((D, n), dtype=torch.float64, device=dev)
    na = torch.empty(n, dtype=torch.int32, device=dev)
    nr = torch.empty(n, dtype=torch.int32, device=dev)
    status = torch.empty(n, dtype=torch.int32, device=dev)
    tape = torch.empty((cap, n), dtype=torch.float64, device=dev)
    y0bar = torch.empty((D, n), dtype=torch.float64, device=dev)
    pbar = torch.empty((2, n), dtype=torch.float64, device=dev)
    ybar = torch.empty((D, n), dtype=torch.float64, device=dev)
    best = None
    for C in (4, 8, 16):
        ckpt = H._alloc_ckpt(fid, D, n, cap, C, dev)
        ws1 = HW.workspace(1, D, n, dev, C)

        def fwd(record=True):
            N.call("gsx_tsit5_wide_fwd", fid, n, D, 2, None, y0.data_ptr(), p.data_ptr(),
                   kw["t0"], kw["t1"], None, kw["rtol"], kw["atol"], 100000,
                   y_final.data_ptr(), na.data_ptr(), nr.data_ptr(), status.data_ptr(),
                   tape.data_ptr() if record else None, cap if record else 0,
                   ckpt.data_ptr() if record else None, C, ws0.data_ptr(), st)

        def both():
            fwd()
            torch.mul(y_final, 2.0, out=ybar)
            N.call("gsx_tsit5_wide_adjoint", fid, n, D, 2, None, None, p.data_ptr(), kw["t0"],
                   tape.data_ptr(), cap, na.data_ptr(), ckpt.data_ptr(), C, ybar.data_ptr(),
                   y0bar.data_ptr(), pbar.data_ptr(), ws1.data_ptr(), st)

        tf0 = _best(lambda: fwd(False), args.repeat)
        tf = _best(fwd, args.repeat)
        tb = _best(both, args.repeat)
        assert int((status != 0).sum().item()) == 0
        out[f"C{C}"] = {"fwd_only_ms": round(tf0 * 1e3, 3), "record_ms": round(tf * 1e3, 3),
                        "record_adj_ms": round(tb * 1e3, 3),
                        "adj_workspace_GB": round(ws1.numel() * 8 / 1e9, 2)}
        _log(f"wide D={D} n={n} C={C}: {out[f'C{C}']}")
        if best is None or tb < best[0]:
            best = (tb, tf, tf0, C)
        del ckpt, ws1
    tb, tf, tf0, C = best
    out.update({"ckpt_interval": C, "fwd_traj_per_s": round(n / tf0, 1),
                "fwd_adj_traj_per_s": round(n / tb, 1),
                "fwd_acc_steps_per_s": round(acc / tf0, 1),
                "fwd_adj_acc_steps_per_s": round(acc / tb, 1),
                "fwd_workspace_GB": round(ws0.numel() * 8 / 1e9, 3)})
    return out


WORKLOADS = {"uneven": uneven, "robertson": robertson, "lorenz1m": lorenz1m,
             "kuramoto": kuramoto, "wide": wide, "copy": copy}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="all", choices=["all", *WORKLOADS])
    ap.add_argument("--n", type=int, default=0)
    ap.add_argument("--dim", type=int, default=0, help="state dimension of the wide workload")
    ap.add_argument("--repeat", type=int, default=5)
    # the persistent autotune runs on the whole ensemble: its lanes' tails depend on how many
    # trajectories each lane gets, which a subsample would change
    ap.add_argument("--tune-sample", type=int, default=1 << 22)
    args = ap.parse_args(argv)
    dev = torch.device("cuda:0")
    names = [w for w in WORKLOADS if w != "copy"] if args.workload == "all" else [args.workload]
    for name in names:
        _log(f"workload {name}")
        print(json.dumps(WORKLOADS[name](args, dev)), flush=True)


if __name__ == "__main__":
    main()
