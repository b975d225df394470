"""Every launch choice of the tuning gives the same results: each choice alone and all together,
on a plain grid (against the checker) and on a batch of uneven grids (against the built-in
choices), for fused steps and for step-graph replays; and a lookback wait that gives up reports
DOGM_ERR_DEVICE_CHECK instead of hanging.

The choices change launch shapes only (dynamic-occupancy-grid-map_amd/csrc/dogm_tuning.hip); the
stage sequence is the reference's updateGrid (dogm/src/dogm.cu:115-131).
"""
import pytest

from dogm_ctypes import DOGM, Batch, DogmError, tuning_override
from oracle_ctypes import OracleDOGM
from parity_util import assert_int_equal, assert_vehicle_equal, small_params, small_scene

pytestmark = pytest.mark.gpu

SEED = 321
CHOICES = [
    {"sort": "onesweep"},
    {"merge_slice": 1},
    {"merge_slice": 2},
    {"pred_block": 256},
    {"seg_items": 4},
    {"lanes": 1},
    {"sort": "onesweep", "merge_slice": 1, "pred_block": 256, "seg_items": 4, "lanes": 1},
]
FLEET = [(24, 1200, 120), (40, 6000, 600), (64, 9000, 900), (96, 30000, 3000)]


@pytest.fixture(autouse=True)
def _builtin_afterwards():
    yield
    tuning_override(None)


def _name(c):
    return ",".join(f"{k}={v}" for k, v in c.items())


@pytest.mark.parametrize("choice", CHOICES, ids=_name)
def test_plain_frames_match_checker(choice):
    tuning_override(choice)
    p = small_params(gs=64, res=0.5, n=20000, b=2000)
    sc = small_scene(gs=64, res=0.5, seed=11)
    g, o = DOGM(p, seed=SEED), OracleDOGM(p, seed=SEED)
    try:
        for f in range(3):
            m, pose = sc.measurement(f), sc.ego_pose(f)
            g.update_grid(m, *pose, 0.1)
            o.update_grid(m, *pose, 0.1)
            assert_vehicle_equal(f"{_name(choice)} frame {f}", g.get_particles(), o.get_particles(),
                                 g.get_grid_cells(), o.get_grid_cells())
            assert_int_equal(f"{_name(choice)} frame {f} ancestors", g.get_resampled_indices(),
                             o.get_resampled_indices())
    finally:
        g.close()
        o.close()


def _run_batch(choice, graph):
    tuning_override(choice)
    vs = []
    for k, (gs, n, b) in enumerate(FLEET):
        vs.append((small_params(gs=gs, res=0.5, n=n, b=b), small_scene(gs=gs, res=0.5, seed=20 + k)))
    out = []
    with Batch() as batch:
        batch.set_graph(graph)
        ids = [batch.add_vehicle(p, SEED, k) for k, (p, _) in enumerate(vs)]
        for f in range(4):
            batch.update(ids, [sc.measurement(f) for _, sc in vs], [sc.ego_pose(f) for _, sc in vs], 0.1)
            out.append([(batch.get_particles(i), batch.get_grid_cells(i), batch.get_resampled_indices(i))
                        for i in ids])
    return out


@pytest.mark.parametrize("choice", CHOICES, ids=_name)
def test_batch_frames_match_builtin_choices(choice):
    ref = _run_batch({}, graph=False)
    for graph in (False, True):  # graph: captured on the first frame, replayed on the next three
        got = _run_batch(choice, graph=graph)
        for f, (gf, rf) in enumerate(zip(got, ref)):
            for k, (g, r) in enumerate(zip(gf, rf)):
                tag = f"{_name(choice)} graph={graph} frame {f} vehicle {k}"
                assert_vehicle_equal(tag, g[0], r[0], g[1], r[1])
                assert_int_equal(f"{tag} ancestors", g[2], r[2])


def test_lookback_give_up_is_reported(monkeypatch):
    """A host-forced give-up report (negative poll limit: every tile that looks back reports a give-up
    but still waits as usual) reaches the host as DOGM_ERR_DEVICE_CHECK after the step, every kernel
    having finished normally; a handle created without the forced limit runs cleanly."""
    tuning_override({"sort": "onesweep"})
    monkeypatch.setenv("DOGM_LOOKBACK_SPIN", "-1")
    p = small_params(gs=64, res=0.5, n=20000, b=2000)  # ten sort tiles: every later tile looks back
    sc = small_scene(gs=64, res=0.5, seed=12)
    g = DOGM(p, seed=SEED)
    monkeypatch.delenv("DOGM_LOOKBACK_SPIN")
    try:
        with pytest.raises(DogmError) as ei:
            g.update_grid(sc.measurement(0), *sc.ego_pose(0), 0.1)
        assert ei.value.status == 4 and "lookback" in str(ei.value)
    finally:
        g.close()
    ok, o = DOGM(p, seed=SEED), OracleDOGM(p, seed=SEED)
    try:
        ok.update_grid(sc.measurement(0), *sc.ego_pose(0), 0.1)
        o.update_grid(sc.measurement(0), *sc.ego_pose(0), 0.1)
        assert_vehicle_equal("after a forced give-up report", ok.get_particles(), o.get_particles(),
                             ok.get_grid_cells(), o.get_grid_cells())
    finally:
        ok.close()
        o.close()
