"""The roofline work constants equal the oracle's op-counted work."""
from gradsolve_amd import workcounts as WC
from oracle import oracle as O


def test_lorenz_tsit5_counts():
    assert O.count_ops(0, O.LORENZ, 3, 1) == (WC.LORENZ_TSIT5_TRIAL_FLOPS,
                                              WC.LORENZ_TSIT5_TRIAL_POWS)
    assert O.count_ops(1, O.LORENZ, 3, 1)[0] == WC.LORENZ_TSIT5_REPLAY_STEP_FLOPS
    assert O.count_ops(2, O.LORENZ, 3, 1)[0] == WC.LORENZ_TSIT5_ADJ_STEP_FLOPS


def test_robertson_stiff_lane_counts():
    assert O.count_ops(3, O.ROBERTSON, 3, 3) == (WC.ROBERTSON_RB23_TRIAL_FLOPS, 1)
    assert O.count_ops(4, O.ROBERTSON, 3, 3)[0] == WC.ROBERTSON_RB23_ADJ_STEP_FLOPS
    assert O.count_ops(5, O.ROBERTSON, 3, 3) == (WC.ROBERTSON_RODAS5P_TRIAL_FLOPS, 1)
    assert O.count_ops(6, O.ROBERTSON, 3, 3)[0] == WC.ROBERTSON_RODAS5P_ADJ_STEP_FLOPS
