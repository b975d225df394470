"""The C-ABI library loads and exports every function include/dogm/dogm_c.h declares (no GPU calls)."""
import re

import dogm_ctypes


def _declared(repo):
    text = (repo / "include" / "dogm" / "dogm_c.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dogm_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(repo):
    names = _declared(repo)
    assert len(names) >= 40
    for n in names:
        assert hasattr(dogm_ctypes.lib, n), n
    assert sorted(dogm_ctypes.EXPORTED) == names


def test_status_strings_and_kernel_names():
    assert dogm_ctypes.lib.dogm_status_string(0) == b"ok"
    assert dogm_ctypes.lib.dogm_status_string(1) == b"invalid argument"
    assert dogm_ctypes.lib.dogm_status_string(99) == b"unknown status"
    names = dogm_ctypes.DOGM.kernel_names()
    assert len(names) == dogm_ctypes.lib.dogm_kernel_count()
    assert "predict" in names and "rs_merge" in names
    assert dogm_ctypes.lib.dogm_kernel_name(-1) is None


def test_struct_layouts():
    assert dogm_ctypes.GRID_CELL_DTYPE.itemsize == 64
    assert dogm_ctypes.MEAS_CELL_DTYPE.itemsize == 16
    import ctypes
    assert ctypes.sizeof(dogm_ctypes.dogm_params) == 44


def test_null_handle_calls_are_rejected():
    lib = dogm_ctypes.lib
    assert lib.dogm_particle_prediction(None, 0.1) == 1
    assert lib.dogm_update_grid(None, None, 0.0, 0.0, 0.0, 0.1, 0) == 1
    assert lib.dogm_get_grid_size(None) == 0
    assert lib.dogm_last_error(None) == b"null handle"
