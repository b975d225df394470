"""Host sanitizer run: the checker's every stage (f32 and f64, all CDF, draw and ego modes), its
extraction and moment-sum helpers, and the library's host-side JSON reader (csrc/dogm_json.h), built
with AddressSanitizer and UndefinedBehaviorSanitizer (oracle/Makefile `sanitize`, g++ only: no device
code is built with a sanitizer); any report fails the run."""
import shutil
import subprocess

import pytest

from conftest import REPO


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_checker_and_json_reader_under_asan_ubsan():
    r = subprocess.run(["make", "-s", "-C", str(REPO / "oracle"), "sanitize"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "clean" in r.stdout
