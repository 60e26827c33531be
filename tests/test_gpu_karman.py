"""D2Q9 Karman street (512x128, disc D = 16, Re = 100): the FP32 run against
the FP64 GPU oracle run of the same inputs over a long horizon. Bar: Strouhal
number and mean drag within 1% (the parity bar for long runs), with vortex
shedding present (lift oscillates)."""
import numpy as np
import pytest

from openhomelbm_amd import karman

pytestmark = pytest.mark.gpu


def test_fp32_long_run_matches_fp64_strouhal_and_drag():
    r = karman.compare(steps=16000, window=9000)
    assert r["f64"]["lift_amplitude"] > 1e-3, "no shedding"
    assert 0.1 < r["f64"]["strouhal"] < 0.3
    assert r["f32_rel_err"]["strouhal"] < 0.01
    assert r["f32_rel_err"]["mean_drag"] < 0.01
