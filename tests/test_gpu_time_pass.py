"""wf_time_pass, the back-to-back timing aid the benchmark uses for the
dominant kernel: it repeats a pair pass that reads only its inputs, so the
state after it equals the state after one pass, bit for bit."""
import numpy as np
import pytest

from weldformgpu_amd import _abi

from harness import randomize, small_c1, to_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [_abi.WF_FP64, _abi.WF_MIXED])
def test_time_pass_repeats_without_changing_results(wflib, precision):
    sc = randomize(small_c1(0.4), seed=5)
    sc.cfg.precision = precision
    print(f"[case] time_pass precision {precision} n={sc.n}", flush=True)
    with to_gpu(sc, wflib) as d:
        d.nsearch()
        d.CalcRateTensors()
        d.CalcAccel()
        a1, d1 = d.get("a"), d.get("strrate")
        ms_a = d.time_pass("accel", 5)
        ms_r = d.time_pass("rates", 5)
        assert ms_a > 0.0 and ms_r > 0.0
        assert np.array_equal(d.get("a"), a1)
        assert np.array_equal(d.get("strrate"), d1)
        with pytest.raises(RuntimeError):
            d.time_pass("stress", 5)
