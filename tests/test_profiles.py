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
    # at least the algorithmic bytes of 32 vehicles
    assert traffic and traffic > 32 * bench.kernel_bytes("rs_merge", 2_000_000, 200_000, 1_440_000, 3)
    assert grid == 32 * 500224 and 500 < valu < 5000


def test_valu_issue_fraction_closed_form():
    grid, valu = profiling.lookup_valu(str(SUMMARY), bench.PROFILED_NAME["rs_merge"])
    launch_s = 1e-3
    got = bench.valu_issue("rs_merge", "plain_full", 32, 32 * 2_000_000, launch_s)
    assert got["waves"] == grid // 64
    assert abs(got["frac"] - (grid // 64) * valu * 4 / (1024 * 2.4e9 * launch_s)) < 1e-4
    assert bench.valu_issue("rs_merge", "plain_full", 7, 7 * 2_000_000, launch_s) is None  # no summary for that batch


def test_sort_rows_follow_the_tuned_pass_form():
    """The sort stages are looked up under the kernel of the pass form the tuning gives the step, and the
    counter summary holds a row for the form the committed table selects."""
    from dogm_ctypes import TUNING_BUILTIN, tuning_override
    try:
        for form, kern in (("reduce_scan", "k_sort_downsweep<true>"), ("onesweep", "k_sort_onesweep<true>")):
            tuning_override(dict(TUNING_BUILTIN, sort=form))
            assert bench.profiled_name("sort_gather", 32, 32 * 2_000_000) == kern
    finally:
        tuning_override(None)
    name = bench.profiled_name("sort_gather", 32, 32 * 2_000_000)
    assert profiling.lookup_traffic(str(SUMMARY), name) > 32 * 2_000_000 * 16
