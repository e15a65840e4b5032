"""The library's peak microbenchmarks (r2d_debug_peaks), which the bench
reports its rooflines against: each rate is positive, below the published
peak it measures (HBM3E 8 TB/s; LDS 256 CUs x 256 B per clock at 2.4 GHz)
and above a floor that a working stream clears by far.  The LDS rate is
also above 128 B per clock per CU, which rules out the earlier CDNA
generation's LDS rate as gfx950's peak."""
import pytest

pytestmark = pytest.mark.gpu

HBM_PEAK, LDS_PEAK = 8000.0, 256 * 256 * 2.4


def test_measured_peaks_are_sane(gpu_ctx):
    p = gpu_ctx.measured_peaks()
    print("measured peaks (GB/s):", p, flush=True)
    assert 1000.0 < p["hbm_read"] <= 1.05 * HBM_PEAK
    assert 1000.0 < p["hbm_copy"] <= 1.05 * HBM_PEAK
    assert 256 * 128 * 2.4 < p["lds_read"] <= 1.05 * LDS_PEAK
