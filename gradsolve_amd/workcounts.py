"""Algorithmic work per unit, used for roofline accounting.

Counted on the oracle with an op-counting scalar (every add/sub/mul/div/sqrt = 1 flop, so a
fused multiply-add counts 2; pow counted apart); tests pin these constants against that
count. Checkpoint re-integration inside the adjoint is implementation overhead and is not
counted.
"""

# Lorenz (D=3, P=1) Tsit5.
LORENZ_TSIT5_TRIAL_FLOPS = 271      # one adaptive trial step (FSAL: 6 field calls) + controller
LORENZ_TSIT5_TRIAL_POWS = 1
LORENZ_TSIT5_REPLAY_STEP_FLOPS = 264  # one tsit5_step (7 field calls)
LORENZ_TSIT5_ADJ_STEP_FLOPS = 517   # stage recompute (no k7) + vector-Jacobian products

# HBM bytes (algorithmic) for Lorenz, float64.
TAPE_BYTES_PER_ACCEPTED_STEP = 8
LORENZ_FWD_IO_BYTES = 24 + 8 + 24 + 12   # y0, rho in; y_final, n_acc/n_rej/status out
LORENZ_ADJ_IO_BYTES = 8 + 24 + 24 + 24 + 8   # rho, ybar in (+ checkpoint 0 = y0); y0bar, pbar out

# MI355X published peaks.
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
