"""The committed counter summary feeds the bench line's roofline: the dominant kernel's HBM traffic and
VALU instructions per wave are found in it, and the issue fraction follows from them in closed form."""
import sys

from conftest import REPO

sys.path[:0] = [str(REPO), str(REPO / "dynamic-occupancy-grid-map_amd")]
import bench  # noqa: E402
import profiling  # noqa: E402

SUMMARY = REPO / "profiles" / "hbm_traffic_plain_full_batch32.md"


def test_roofline_kernel_rows_are_found():
    name = bench.PROFILED_NAME["rs_merge"]
    traffic = profiling.lookup_traffic(str(SUMMARY), name)
    grid, valu = profiling.lookup_valu(str(SUMMARY), name)
    # at least the algorithmic bytes of 32 vehicles (50 N + 4 B each)
    assert traffic and traffic > 32 * (50 * 2_000_000 + 4 * 200_000)
    assert grid == 32 * 500224 and 500 < valu < 5000


def test_valu_issue_fraction_closed_form():
    grid, valu = profiling.lookup_valu(str(SUMMARY), bench.PROFILED_NAME["rs_merge"])
    launch_s = 1e-3
    got = bench.valu_issue("rs_merge", "plain_full", 32, launch_s)
    assert got["waves"] == grid // 64
    assert abs(got["frac"] - (grid // 64) * valu * 4 / (1024 * 2.4e9 * launch_s)) < 1e-4
    assert bench.valu_issue("rs_merge", "plain_full", 7, launch_s) is None  # no summary for that batch
