"""Runs the C++ re-expression of the reference's specs (tests/cpp/dogm_spec.cpp) built against the
drop-in header include/dogm/dogm.h and libdogm_hip.so."""
import subprocess

import pytest

pytestmark = pytest.mark.gpu


def test_cpp_dogm_spec(repo):
    exe = repo / "tests" / "cpp" / "dogm_spec"
    if not exe.exists():
        pytest.fail(f"{exe} not built: run dynamic-occupancy-grid-map_amd/build.py")
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "0 failure(s)" in res.stdout
