"""The committed fixture tests/golden/small_3frames.npz against the checker (CPU) and the HIP path
(GPU), bit for bit."""
import numpy as np
import pytest

from golden.make_golden import FRAMES, SEED, scenario, run
from oracle_ctypes import OracleDOGM

FIXTURE = __import__("pathlib").Path(__file__).resolve().parent / "golden" / "small_3frames.npz"


def _compare(out, ref):
    for k, v in out.items():
        r = ref[k]
        if v.dtype.kind == "f":
            assert np.array_equal(v.view(np.uint32), r.view(np.uint32)), k
        else:
            assert np.array_equal(v, r), k


def _inputs(ref):
    p, _, _ = scenario()
    meas = [ref["meas"][f] for f in range(FRAMES)]
    poses = [tuple(float(v) for v in ref["poses"][f]) for f in range(FRAMES)]
    return p, meas, poses


def test_checker_reproduces_fixture():
    ref = np.load(FIXTURE)
    p, meas, poses = _inputs(ref)
    _compare(run(OracleDOGM(p, seed=SEED), meas, poses), ref)


@pytest.mark.gpu
def test_gpu_reproduces_fixture():
    from dogm_ctypes import DOGM
    ref = np.load(FIXTURE)
    p, meas, poses = _inputs(ref)
    _compare(run(DOGM(p, seed=SEED), meas, poses), ref)
