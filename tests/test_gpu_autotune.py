"""Autotuner of the window decoders' geometry (turbo_cpp_amd.autotune)."""
import pytest

import turbo_cpp_amd as T
from turbo_cpp_amd import autotune

pytestmark = pytest.mark.gpu


def test_tuner_picks_a_ber_equivalent_candidate():
    """On a small calibration batch every candidate is timed; the chosen one
    has its bit errors inside the exact decoder's 95% interval and is the
    fastest such candidate; waves_per_cu is honoured by the plan."""
    print("[gpu-test] autotune K=1024", flush=True)
    r = autotune.tune(1024, n_blocks=256, ebn0_db=(0.25, 0.5, 0.75), min_segments=(16, 32, 64),
                      warmups=(16, 32), waves=(0, 2), extrinsic_bits=(16, 32), reps=2)
    table = r["table"]
    assert len(table) >= 8
    best = r["best"]
    assert best["qualifies"]
    assert best["ms"] == min(c["ms"] for c in table if c["qualifies"])
    for p, (be, ke) in zip(r["points"], best["per_point"]):  # inside both intervals at every point
        assert p["bit_interval"][0] <= be <= p["bit_interval"][1]
        assert p["block_interval"][0] <= ke <= p["block_interval"][1]
    # the default geometry (S=32, g=32, all waves) qualifies: it is the BER-validated design point
    default = [c for c in table if c["segment_length"] == 32 and c["warmup"] == 32 and c["waves_per_cu"] == 0
               and c["extrinsic_bits"] == 16]
    assert default and default[0]["qualifies"]
    plan = T.Plan([1024] * 8, **autotune.options_for(r))
    assert plan.post_floats == 8 * 1024
