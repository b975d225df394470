"""The tuning table reader (dogm_tuning.hip) on the host: first matching entry wins, ranges are
closed, unknown keys are ignored, malformed tables and invalid choices are rejected, and the
override takes precedence.  No GPU is needed: these calls touch no device."""
import json

import pytest

from dogm_ctypes import DogmError, lib, tuning_load, tuning_lookup, tuning_override

BUILTIN = {"sort": "reduce_scan", "merge_slice": 0, "pred_block": 512, "seg_items": 8, "lanes": 2}


@pytest.fixture
def table(tmp_path):
    def write(obj):
        p = tmp_path / "t.json"
        p.write_text(obj if isinstance(obj, str) else json.dumps(obj))
        return p
    yield write
    tuning_override(None)
    tuning_load(None)


def test_first_matching_entry_wins(table):
    tuning_load(table({"generated_by": "test", "entries": [
        {"class": "a", "vehicles": [1, 1], "particles": [0, 999], "sort": "onesweep", "merge_slice": 1,
         "pred_block": 256, "seg_items": 4, "lanes": 1, "step_ms": 1.0, "tried": [{"x": [1, 2]}]},
        {"vehicles": [1, 8], "particles": [0, 10**12], "seg_items": 4},
        {"vehicles": [1, 8], "particles": [0, 10**12], "sort": "onesweep"}]}))
    assert lib.dogm_tuning_entries() == 3
    assert tuning_lookup(1, 999) == {"sort": "onesweep", "merge_slice": 1, "pred_block": 256, "seg_items": 4,
                                     "lanes": 1}
    assert tuning_lookup(1, 1000) == dict(BUILTIN, seg_items=4)  # closed range [0, 999] ends at 999
    assert tuning_lookup(8, 0) == dict(BUILTIN, seg_items=4)
    assert tuning_lookup(9, 0) == BUILTIN  # no entry: built-in choices


def test_override_takes_precedence(table):
    tuning_load(table({"entries": [{"vehicles": [1, 100], "particles": [0, 10**12], "lanes": 1}]}))
    tuning_override({"sort": "onesweep"})
    assert tuning_lookup(4, 100) == dict(BUILTIN, sort="onesweep")
    tuning_override(None)
    assert tuning_lookup(4, 100) == dict(BUILTIN, lanes=1)


@pytest.mark.parametrize("bad", [
    "{",
    {"entries": {}},
    {"entries": [{"vehicles": [1], "particles": [0, 1]}]},
    {"entries": [{"vehicles": [2, 1], "particles": [0, 1]}]},
    {"entries": [{"vehicles": [1, 1], "particles": [0, 1], "sort": "bitonic"}]},
    {"entries": [{"vehicles": [1, 1], "particles": [0, 1], "pred_block": 128}]},
    {"entries": [{"vehicles": [1, 1], "particles": [0, 1], "lanes": 3}]},
])
def test_malformed_tables_are_rejected_and_keep_the_old_one(table, bad):
    tuning_load(table({"entries": [{"vehicles": [1, 1], "particles": [0, 10], "seg_items": 4}]}))
    with pytest.raises(DogmError):
        tuning_load(table(bad))
    assert lib.dogm_tuning_entries() == 1 and tuning_lookup(1, 5)["seg_items"] == 4


def test_invalid_override_is_rejected():
    with pytest.raises(DogmError):
        tuning_override({"seg_items": 6})
