"""Peak microbenchmarks (LDS read stream, integer VALU stream) on the device:
the measured figures are positive and within a plausible band of the
published gfx950 ones (per CU 256 B/clk of LDS reads and 64 integer lanes
per clock, at up to 2.4 GHz)."""
import pytest

import edge_matching_puzzle_amd as emp

pytestmark = pytest.mark.gpu


def test_measured_peaks_are_plausible(gpu_context):
    pk = gpu_context.peaks()
    info = gpu_context.device_info()
    cus = info["compute_units"]
    lds_published = cus * 256 * 2.4          # GB/s
    valu_published = cus * 64 * 2.4          # G lane-ops/s
    assert 0.2 * lds_published < pk["lds_gbs"] < 1.05 * lds_published, pk
    assert 0.2 * valu_published < pk["valu_gops"] < 2.5 * valu_published, pk
