"""FP32 path (bit-identical to the GPU) against the FP64 build of the same
physics on the same random streams: nearly every history takes the same
discrete path, and the continuous results agree to FP32 rounding."""


def test_fp32_error_against_fp64():
    from precision_report import report
    r = report(config=2, n=50000)
    assert r["same_path_fraction"] > 0.9999
    assert r["events_matched"] > 3000 and r["events_only_one_side"] <= 2
    assert r["deposited_energy_rel"]["p99"] < 1e-5
    assert r["interaction_position_cm"]["p99"] < 1e-4
    assert r["timestamp_us"]["p99"] < 1e-8
    assert r["exit_energy_rel"]["max"] < 1e-6
