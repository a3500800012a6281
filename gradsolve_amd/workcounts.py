"""Algorithmic work per unit, used for roofline accounting.

Counted on the oracle with an op-counting scalar (every add/sub/mul/div/sqrt = 1 flop, so a
fused multiply-add counts 2; pow counted apart); tests pin these constants against that
count. Checkpoint re-integration inside the adjoint is implementation overhead and is not
counted, and neither are the stage times t + c_i h of a field that does not read t (Lorenz,
Robertson): the kernels never form them.
"""

# Lorenz (D=3, P=1) Tsit5.
LORENZ_TSIT5_TRIAL_FLOPS = 260      # one adaptive trial step (FSAL: 6 field calls) + controller
LORENZ_TSIT5_TRIAL_POWS = 1
LORENZ_TSIT5_REPLAY_STEP_FLOPS = 252  # one tsit5_step (7 field calls)
LORENZ_TSIT5_ADJ_STEP_FLOPS = 495   # stage recompute (no k7) + vector-Jacobian products

# Robertson (D=3, P=3) on the stiff lanes: one adaptive trial step (Jacobian, one LU of W, the
# stage solves, the error norm and the controller) and one adjoint step (stage recompute and
# the transposed products through the solves, the field and the Jacobian).
ROBERTSON_RB23_TRIAL_FLOPS = 198
ROBERTSON_RB23_ADJ_STEP_FLOPS = 302
ROBERTSON_RODAS5P_TRIAL_FLOPS = 932
ROBERTSON_RODAS5P_ADJ_STEP_FLOPS = 2528

# HBM bytes (algorithmic) for Lorenz, float64.
TAPE_BYTES_PER_ACCEPTED_STEP = 8
LORENZ_FWD_IO_BYTES = 24 + 8 + 24 + 12   # y0, rho in; y_final, n_acc/n_rej/status out
LORENZ_ADJ_IO_BYTES = 8 + 24 + 24 + 24 + 8   # rho, ybar in (+ checkpoint 0 = y0); y0bar, pbar out

# MI355X published peaks.
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


# HBM bytes per launch of the headline workload's kernels (N = 2^20 Lorenz, rtol 1e-8), by
# checkpoint interval: rocprofv3 FETCH_SIZE and WRITE_SIZE per dispatch, scaled by what the same
# counters read for a 1 GiB device copy (FETCH_SIZE 0.5 GiB -> x2.0; WRITE_SIZE 1 GiB -> x1.0).
# Algorithmic bytes: the forward reads y0 and p and writes the tape (8 B per accepted step), a
# checkpoint (24 B) at the start and every C steps, y_final and the counts; the adjoint reads
# them back with the cotangent and writes the two gradients.
HEADLINE_TRAFFIC_BYTES = {
    # C: {kernel: (measured bytes, algorithmic bytes)}; at C = 2 the adjoint launch is
    # tsit5_adj_pair_kernel (kept under the adjoint's key).
    2: {"tsit5_fwd_kernel": (2.544e9, 2.442e9), "tsit5_adj_kernel": (2.533e9, 2.442e9)},
    4: {"tsit5_fwd_kernel": (1.801e9, 1.738e9), "tsit5_adj_kernel": (1.807e9, 1.738e9)},
    8: {"tsit5_fwd_kernel": (1.419e9, 1.386e9), "tsit5_adj_kernel": (1.447e9, 1.386e9)},
    16: {"tsit5_fwd_kernel": (1.239e9, 1.211e9), "tsit5_adj_kernel": (1.254e9, 1.211e9)},
}
