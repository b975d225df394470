"""Pins the checker (oracle/) stage by stage against facts that do not come from the checker itself:
the reference's own test intents (dogm/test/dogm_spec.cpp), closed forms, the textbook
Dempster-Shafer combination evaluated in float64, brute-force float64 sums, exact-rational
recomputations, invariants and distribution tests.  CPU only."""
import bisect
import math
from fractions import Fraction

import numpy as np
import pytest
from scipy import stats

import oracle_ctypes as oc
from oracle_ctypes import GRID_FIELDS, MEAS_CELL_DTYPE, OracleDOGM
from parity_util import small_params, small_scene


def meas_grid(gs, occ=None, free=None, lik=1.0, pA=1.0):
    m = np.zeros(gs * gs, MEAS_CELL_DTYPE)
    m["likelihood"] = lik
    m["p_A"] = pA
    if occ is not None:
        m["occ_mass"] = occ
    if free is not None:
        m["free_mass"] = free
    return m


def spec_params(**kw):
    # dogm_spec.cpp:12-22 (freespace_discount is left unset there).
    base = dict(gs=10, res=1.0, n=2, b=1, persistence_prob=0.5, stddev_process_noise_position=0.0,
                stddev_process_noise_velocity=0.0, birth_prob=0.02, stddev_velocity=10.0, init_max_velocity=30.0)
    base.update(kw)
    return small_params(**base)


def grid_dict(o):
    return o.get_grid_cells()


# ---- ego motion (dogm.cu:161-203; dogm_spec.cpp:10-66) ------------------------------------------------
def test_ego_motion_reference_intent():
    o = OracleDOGM(spec_params())
    st = np.array([[2.5, 3.5, 1.25, -0.75], [7.25, 1.5, -2.0, 0.5]], np.float32)
    o.set_particles(st, [32, 17], [0.25, 0.5], [1, 0])
    o.update_pose(10.0, 10.0, 0.0)                      # first pose: stored, no shift
    assert o.position()[:2] == (10.0, 10.0)
    assert np.array_equal(o.get_particles()["state"], st)
    o.update_pose(10.5, 10.5, 0.0)                      # below one resolution: ignored
    assert o.position()[:2] == (10.0, 10.0)
    assert np.array_equal(o.get_particles()["state"], st)
    o.update_pose(13.0, 10.0, 0.0)                      # +3 cells: particles move by +3 in x
    assert o.position()[:2] == (13.0, 10.0)
    exp = st.copy()
    exp[:, 0] += 3.0
    assert np.array_equal(o.get_particles()["state"], exp)
    o.update_pose(13.0, 8.0, 0.7)                       # -2 cells in y; yaw stored, not applied
    exp[:, 1] -= 2.0
    assert np.array_equal(o.get_particles()["state"], exp)
    assert o.position() == (13.0, 8.0, pytest.approx(0.7))


def test_ego_motion_grid_shift():
    gs = 12
    o = OracleDOGM(small_params(gs=gs, res=1.0, n=4, b=1))
    cells = o.get_grid_cells()
    vals = np.arange(gs * gs, dtype=np.float64) + 1.0
    for f in GRID_FIELDS:
        cells[f] = vals * (1 + GRID_FIELDS.index(f))
    cells["start_idx"] = np.arange(gs * gs)
    cells["end_idx"] = np.arange(gs * gs) + 7
    o.set_grid_cells(cells)
    o.update_pose(0.0, 0.0, 0.0)
    o.update_pose(2.6, -1.2, 0.0)   # x_move = -int(2.6) = -2, y_move = -int(-1.2) = 1
    g = o.get_grid_cells()
    old = vals.reshape(gs, gs)      # [y, x]
    exp = np.zeros((gs, gs))
    # new[x, y] = old[x - 2, y + 1] where the source lies strictly inside (0, gs) on both axes.
    for y in range(gs):
        for x in range(gs):
            sx, sy = x - 2, y + 1
            if 0 < sx < gs and 0 < sy < gs:
                exp[y, x] = old[sy, sx]
    assert np.array_equal(g["free_mass"].reshape(gs, gs), exp * 3)
    assert np.array_equal(g["start_idx"].reshape(gs, gs), np.where(exp > 0, exp - 1, 0))
    # Source column 0 is never copied (strict lower bound, ego_motion_compensation.cu:38).
    assert np.all(g["occ_mass"].reshape(gs, gs)[:, 2] == 0)


# ---- prediction (predict.cu:16-52; dogm_spec.cpp:68-103) -----------------------------------------------
def test_predict_reference_intent_closed_form():
    o = OracleDOGM(spec_params())
    st = np.array([[2.5, 3.5, 1.25, -0.75], [9.1, 1.5, 20.0, 0.5]], np.float32)
    o.set_particles(st, [32, 19], [0.25, 0.5], [1, 0])
    o.particle_prediction(0.1)
    p = o.get_particles()
    dt = np.float32(0.1)
    exp = st + dt * np.stack([st[:, 2], st[:, 3], np.zeros(2, np.float32), np.zeros(2, np.float32)], 1)
    assert np.array_equal(p["state"], exp.astype(np.float32))
    # Second particle leaves the grid (x = 11.1 > 9): weight 0; index clamped to the border.
    assert p["weight"][0] == np.float32(0.25) * np.float32(0.5) and p["weight"][1] == 0.0
    assert list(p["grid_cell_idx"]) == [2 + 10 * 3, 9 + 10 * 1]


def test_predict_noise_distribution():
    n = 40000
    p = small_params(gs=100, res=1.0, n=n, b=1, stddev_process_noise_position=0.3,
                     stddev_process_noise_velocity=2.5)
    o = OracleDOGM(p, seed=99)
    rng = np.random.default_rng(0)
    st = np.stack([rng.uniform(20, 80, n), rng.uniform(20, 80, n), rng.normal(0, 5, n), rng.normal(0, 5, n)],
                  1).astype(np.float32)
    o.set_particles(st, np.zeros(n, np.int32), np.full(n, 1e-4, np.float32), np.zeros(n, np.uint8))
    o.particle_prediction(0.1)
    q = o.get_particles()["state"].astype(np.float64)
    s = st.astype(np.float64)
    res = [(q[:, 0] - s[:, 0] - 0.1 * s[:, 2]) / 0.3, (q[:, 1] - s[:, 1] - 0.1 * s[:, 3]) / 0.3,
           (q[:, 2] - s[:, 2]) / 2.5, (q[:, 3] - s[:, 3]) / 2.5]
    for r in res:
        assert abs(r.mean()) < 0.03 and abs(r.std() - 1) < 0.03
        assert stats.kstest(r, "norm").pvalue > 1e-3
    c = np.corrcoef(np.stack(res))
    assert np.max(np.abs(c - np.eye(4))) < 0.03


# ---- assignment (dogm.cu:262-281) ---------------------------------------------------------------------
def test_assignment_is_stable_sort_with_bounds():
    gs, n = 20, 5000
    o = OracleDOGM(small_params(gs=gs, res=1.0, n=n, b=1))
    rng = np.random.default_rng(1)
    idx = rng.integers(0, gs * gs, n).astype(np.int32)
    idx[:50] = 7  # a long run
    st = np.zeros((n, 4), np.float32)
    st[:, 2] = np.arange(n)
    w = rng.random(n).astype(np.float32)
    o.set_particles(st, idx, w, np.zeros(n, np.uint8))
    o.particle_assignment()
    p = o.get_particles()
    order = np.argsort(idx, kind="stable")
    assert np.array_equal(p["grid_cell_idx"], idx[order])
    assert np.array_equal(p["state"][:, 2].astype(np.int64), order)
    assert np.array_equal(o.get_weight_array(), w[order])
    g = o.get_grid_cells()
    u, first, counts = np.unique(idx[order], return_index=True, return_counts=True)
    assert np.array_equal(g["start_idx"][u], first) and np.array_equal(g["end_idx"][u], first + counts - 1)
    empty = np.setdiff1d(np.arange(gs * gs), u)
    assert np.all(g["start_idx"][empty] == -1) and np.all(g["end_idx"][empty] == -1)


# ---- occupancy / mass update (mass_update.cu:16-93) -----------------------------------------------------
def _prepared(gs=16, n=3000, seed=4, pA=1.0, heavy=False):
    p = small_params(gs=gs, res=1.0, n=n, b=300)
    o = OracleDOGM(p, seed=seed)
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, gs * gs // 2, n).astype(np.int32)
    w = rng.uniform(0, 2.0 / n, n).astype(np.float32)
    if heavy:
        w[idx == 3] = 0.2  # cell 3 sums to > 1: normalised
    st = np.stack([rng.uniform(0, gs, n), rng.uniform(0, gs, n), rng.normal(0, 3, n), rng.normal(0, 3, n)],
                  1).astype(np.float32)
    o.set_particles(st, idx, w, np.zeros(n, np.uint8))
    m = meas_grid(gs, occ=rng.choice([0.0, 0.7], gs * gs, p=[0.7, 0.3]),
                  free=rng.choice([0.0, 0.4], gs * gs, p=[0.6, 0.4]), lik=rng.uniform(0.5, 1.0, gs * gs), pA=pA)
    m["free_mass"][m["occ_mass"] > 0] = 0.0
    o.update_measurement_grid(m)  # first measurement also initialises; particles are set again below
    o.set_particles(st, idx, w, np.zeros(n, np.uint8))
    cells = o.get_grid_cells()
    cells["free_mass"] = rng.uniform(0, 0.6, gs * gs)
    o.set_grid_cells(cells)
    o.particle_assignment()
    return o, m


def test_occupancy_against_textbook_dempster_shafer():
    o, m = _prepared(heavy=True)
    gs = o.grid_size
    before = o.get_grid_cells()
    wa = o.get_weight_array().astype(np.float64)
    idx = o.get_particles()["grid_cell_idx"]
    o.grid_cell_occupancy_update(0.1)
    g = o.get_grid_cells()
    alpha = float(np.float32(0.01) ** np.float32(0.1))
    pB = 0.02
    msum = np.bincount(idx, weights=wa, minlength=gs * gs)
    for c in range(gs * gs):
        mo = min(msum[c], 1.0) if before["start_idx"][c] != -1 else 0.0
        mf = min(alpha * float(before["free_mass"][c]), 1.0 - mo)
        th1 = 1.0 - mo - mf
        o2, f2 = float(m["occ_mass"][c]), float(m["free_mass"][c])
        th2 = 1.0 - o2 - f2
        K = mo * f2 + mf * o2
        occ = (mo * o2 + mo * th2 + th1 * o2) / (1.0 - K)
        free = (mf * f2 + mf * th2 + th1 * f2) / (1.0 - K)
        rb = occ * pB * (1.0 - mo) / (mo + pB * (1.0 - mo)) if (mo + pB * (1.0 - mo)) > 0 else 0.0
        for name, ref in (("occ_mass", occ), ("free_mass", free), ("pred_occ_mass", mo),
                          ("new_born_occ_mass", rb), ("pers_occ_mass", occ - rb)):
            assert abs(g[name][c] - ref) <= 1e-5 * abs(ref) + 2e-7, (c, name, g[name][c], ref)
    assert np.all(g["occ_mass"] + g["free_mass"] <= 1.0 + 1e-6)
    assert np.all(g["occ_mass"] >= 0) and np.all(g["free_mass"] >= 0)
    assert np.array_equal(o.get_born_masses(), g["new_born_occ_mass"])
    # The overfull cell: weights divided by its sum, prediction clamped to 1 (mass_update.cu:51-59, 79-83).
    assert g["pred_occ_mass"][3] == 1.0
    w_after = o.get_particles()["weight"][idx == 3].astype(np.float64)
    assert abs(w_after.sum() - 1.0) < 1e-5


# ---- persistent update (update_persistent_particles.cu:16-87) -----------------------------------------
@pytest.mark.parametrize("pA", [1.0, 0.6])
def test_persistent_update_preserves_cell_mass(pA):
    o, m = _prepared(pA=pA, seed=5)
    o.grid_cell_occupancy_update(0.1)
    o.update_persistent_particles()
    g = o.get_grid_cells()
    idx = o.get_particles()["grid_cell_idx"]
    wp = o.get_weight_array().astype(np.float64)
    w = o.get_particles()["weight"].astype(np.float64)
    lik = m["likelihood"].astype(np.float64)[idx]
    C = o.grid_cell_count
    post_sum = np.bincount(idx, weights=wp, minlength=C)
    lw_sum = np.bincount(idx, weights=lik * w, minlength=C)
    occ_cells = np.nonzero(g["start_idx"] != -1)[0]
    for c in occ_cells:
        rho_p = float(g["pers_occ_mass"][c])
        if lw_sum[c] > 0 and g["pred_occ_mass"][c] > 0:
            assert abs(g["mu_A"][c] - rho_p / lw_sum[c]) <= 1e-5 * abs(rho_p / lw_sum[c]) + 1e-7
            assert abs(post_sum[c] - rho_p) <= 1e-5 * abs(rho_p) + 1e-7, (c, post_sum[c], rho_p)
    empty = g["start_idx"] == -1
    assert np.all(g["mu_A"][empty] == 0) and np.all(g["mu_UA"][empty] == 0)


# ---- births (init_new_particles.cu:18-195) ----------------------------------------------------------------
@pytest.mark.parametrize("pA", [1.0, 0.75])
def test_births_counts_positions_weights(pA):
    o, m = _prepared(pA=pA, seed=6)
    o.grid_cell_occupancy_update(0.1)
    o.update_persistent_particles()
    o.initialize_new_particles()
    g = o.get_grid_cells()
    b = o.get_birth_particles()
    B, gs = o.new_born_particle_count, o.grid_size
    born = o.get_born_masses().astype(np.float32)
    # Exact proportional split: n_j = floor(B S_j / S) - floor(B S_{j-1} / S) within one particle.
    S = [Fraction(0)]
    for x in born:
        S.append(S[-1] + Fraction(float(x)))
    counts = np.bincount(b["grid_cell_idx"], minlength=gs * gs)
    assert counts.sum() == B
    ends = np.cumsum(counts)
    exact = [math.floor(B * S[j + 1] / S[-1]) for j in range(gs * gs)]
    assert max(abs(int(e) - x) for e, x in zip(ends, exact)) <= 1
    assert np.all(np.diff(b["grid_cell_idx"]) >= 0)  # cell j owns a contiguous run of slots
    cell = b["grid_cell_idx"]
    assert np.array_equal(b["state"][:, 0], (cell % gs).astype(np.float32) + np.float32(0.5))
    assert np.array_equal(b["state"][:, 1], cell.astype(np.float32) / np.float32(gs) + np.float32(0.5))
    v = b["state"][:, 2:].astype(np.float64).ravel() / o.params.stddev_velocity
    assert stats.kstest(v, "norm").pvalue > 1e-4
    for j in np.nonzero(counts)[0]:
        sl = cell == j
        n = counts[j]
        nu_A = int(np.float32(np.float32(n) * np.float32(pA)).round()) if pA != 1.0 else n
        n_assoc = min(nu_A + 1, n)
        assert int(b["associated"][sl].sum()) == n_assoc
        assert np.all(b["associated"][sl][:n_assoc] == 1)
        wa = pA * float(born[j]) / nu_A if nu_A > 0 else 0.0
        assert np.allclose(b["weight"][sl][b["associated"][sl] == 1], wa, rtol=1e-6, atol=1e-12)
        assert g["w_A"][j] == pytest.approx(wa, rel=1e-6, abs=1e-12)
    if pA == 1.0:
        assert np.all(b["associated"] == 1)


# ---- moments (statistical_moments.cu:16-106) -------------------------------------------------------------
def test_moments_brute_force():
    o, m = _prepared(seed=7)
    for st in ("grid_cell_occupancy_update", "update_persistent_particles", "initialize_new_particles",
               "statistical_moments"):
        getattr(o, st)(*((0.1,) if st == "grid_cell_occupancy_update" else ()))
    g = o.get_grid_cells()
    p = o.get_particles()
    w = o.get_weight_array().astype(np.float64)
    vx, vy = p["state"][:, 2].astype(np.float64), p["state"][:, 3].astype(np.float64)
    idx = p["grid_cell_idx"]
    C = o.grid_cell_count
    s = [np.bincount(idx, weights=x, minlength=C) for x in (w * vx, w * vy, w * vx * vx, w * vy * vy, w * vx * vy)]
    rho = g["pers_occ_mass"].astype(np.float64)
    for c in range(C):
        if g["start_idx"][c] == -1 or rho[c] <= 0:
            assert g["mean_x_vel"][c] == 0 and g["var_x_vel"][c] == 0 and g["covar_xy_vel"][c] == 0
            continue
        mx, my = s[0][c] / rho[c], s[1][c] / rho[c]
        ref = dict(mean_x_vel=mx, mean_y_vel=my, var_x_vel=s[2][c] / rho[c] - mx * mx,
                   var_y_vel=s[3][c] / rho[c] - my * my, covar_xy_vel=s[4][c] / rho[c] - mx * my)
        scale = s[2][c] / rho[c] + s[3][c] / rho[c] + 1e-6
        for k, r in ref.items():
            assert abs(g[k][c] - r) <= 1e-5 * scale, (c, k, g[k][c], r)


# ---- resampling (resampling.cu:17-68) -----------------------------------------------------------------
def _spacing_draws(seed, instance, epoch, n, total):
    """Independent recomputation of the sorted draws: exponentials from the pinned Philox/log,
    exact integer prefix, double-precision ratio."""
    S, s = [], 0
    for k in range(n + 1):
        if k % 4 == 0:
            words = oc.philox((k // 4, epoch, 4, instance), (seed & 0xffffffff, seed >> 32))
        u = ((words[k % 4] >> 8) + 1) * 2.0 ** -24
        e = -np.float32(oc.log_f32(u))
        s += oc.fx_from_f32(float(e))
        S.append(s)
    sn = S[n]
    sh = max(0, sn.bit_length() - 53)
    inv = 1.0 / float(sn >> sh)
    tot_d = oc.fx_to_f64(total)
    out = []
    for k in range(n):
        u = float(S[k] >> sh) * inv
        if not u < 1.0:
            u = float.fromhex("0x1.fffffffffffffp-1")
        out.append(oc.fx_from_f64(tot_d * u))
    return out


def test_resampling_ancestors_exact():
    n, b = 600, 60
    p = small_params(gs=8, res=1.0, n=n, b=b)
    o = OracleDOGM(p, seed=77, instance=3)
    rng = np.random.default_rng(8)
    st = rng.normal(0, 5, (n, 4)).astype(np.float32)
    w = (rng.random(n) ** 4 * 1e-2).astype(np.float32)
    w[::17] = 0.0
    w[5] = -1e-9  # a negative posterior weight counts as zero
    o.set_particles(st, rng.integers(0, 64, n), rng.random(n).astype(np.float32), np.zeros(n, np.uint8))
    o.set_weight_array(w)
    bw = (rng.random(b) * 1e-3).astype(np.float32)
    o.set_birth_particles(rng.normal(0, 1, (b, 4)).astype(np.float32), rng.integers(0, 64, b), bw,
                          rng.integers(0, 2, b))
    o.epoch = 5
    before = o.get_particles()
    births = o.get_birth_particles()
    o.resampling()
    anc = o.get_resampled_indices()
    joint = np.concatenate([w, bw])
    cdf, s = [], 0
    for x in joint:
        s += oc.fx_from_f32(float(max(x, 0.0)))
        cdf.append(s)
    draws = _spacing_draws(77, 3, 5, n, s)
    assert draws == sorted(draws)
    exp = [min(bisect.bisect_left(cdf, r), n + b - 1) for r in draws]
    assert list(anc) == exp
    after = o.get_particles()
    total = np.float32(oc.fx_to_f32(s))
    assert np.all(after["weight"] == total / np.float32(n))
    src = np.concatenate([before["state"], births["state"]])
    assert np.array_equal(after["state"], src[anc])
    assert not np.any(np.isin(anc, np.nonzero(joint <= 0)[0]))


def test_resampling_is_multinomial():
    n = 20000
    p = small_params(gs=8, res=1.0, n=n, b=4)
    o = OracleDOGM(p, seed=1)
    rng = np.random.default_rng(9)
    groups = rng.integers(0, 20, n)
    w = (groups + 1).astype(np.float32) * 1e-5
    o.set_weight_array(w)
    o.set_birth_particles(np.zeros((4, 4), np.float32), np.zeros(4), np.zeros(4, np.float32), np.zeros(4))
    o.resampling()
    anc = o.get_resampled_indices()
    obs = np.bincount(groups[anc], minlength=20)
    exp_p = np.bincount(groups, weights=w.astype(np.float64), minlength=20)
    exp_p /= exp_p.sum()
    assert stats.chisquare(obs, exp_p * n).pvalue > 1e-3
    # Draws are the order statistics of uniforms: the normalized CDF at the ancestors is ~U(0,1).
    cdf = np.cumsum(w.astype(np.float64))
    u = cdf[anc] / cdf[-1]
    assert stats.kstest(u, "uniform").pvalue > 1e-4


# ---- first-frame initialisation (init_new_particles.cu:76-124) -----------------------------------------
def test_initialize_particles_split_and_state():
    gs, n = 16, 5000
    p = small_params(gs=gs, res=1.0, n=n, b=10, init_max_velocity=12.0)
    o = OracleDOGM(p, seed=2)
    rng = np.random.default_rng(10)
    occ = np.where(rng.random(gs * gs) < 0.2, 0.7, 0.0).astype(np.float32)
    o.update_measurement_grid(meas_grid(gs, occ=occ))
    q = o.get_particles()
    counts = np.bincount(q["grid_cell_idx"], minlength=gs * gs)
    assert counts.sum() == n and np.all(counts[occ == 0] == 0)
    nz = counts[occ > 0]
    assert nz.max() - nz.min() <= 1  # equal masses -> equal split
    c = q["grid_cell_idx"]
    assert np.array_equal(q["state"][:, 0], (c % gs).astype(np.float32) + 0.5)
    assert np.array_equal(q["state"][:, 1], (c // gs).astype(np.float32) + 0.5)  # integer division
    v = q["state"][:, 2:].ravel()
    assert v.min() >= -12.0 and v.max() < 12.0
    assert stats.kstest((v + 12.0) / 24.0, "uniform").pvalue > 1e-3
    assert np.all(q["weight"] == np.float32(1.0) / np.float32(n))


def test_initialize_particles_zero_evidence_spreads_uniformly():
    gs, n = 10, 1000
    o = OracleDOGM(small_params(gs=gs, res=1.0, n=n, b=1))
    o.update_measurement_grid(meas_grid(gs))
    counts = np.bincount(o.get_particles()["grid_cell_idx"], minlength=gs * gs)
    assert np.all(counts == n // (gs * gs))


# ---- whole frames -----------------------------------------------------------------------------------------
def test_deterministic_and_instance_dependent():
    p = small_params()
    sc = small_scene()
    runs = []
    for inst in (0, 0, 1):
        o = OracleDOGM(p, seed=3, instance=inst)
        for f in range(2):
            x, y, yaw = sc.ego_pose(f)
            o.update_grid(sc.measurement(f), x, y, yaw, 0.1)
        runs.append(o.get_particles()["state"])
    assert np.array_equal(runs[0], runs[1])
    assert not np.array_equal(runs[0], runs[2])


def test_fp32_against_fp64_one_frame():
    """The f32 checker against its double-precision twin from identical inputs: masses within 1e-5."""
    p = small_params(gs=40, res=0.5, n=20000, b=2000)
    sc = small_scene(gs=40, res=0.5)
    f32, f64 = OracleDOGM(p, seed=4), OracleDOGM(p, seed=4, fp64=True)
    for o in (f32, f64):
        o.update_grid(sc.measurement(0), 0.0, 0.0, 0.0, 0.1)
    a, b = f32.get_grid_cells(), f64.get_grid_cells()
    for k in ("occ_mass", "free_mass", "pers_occ_mass", "new_born_occ_mass", "pred_occ_mass"):
        err = np.abs(a[k].astype(np.float64) - b[k])
        assert np.all(err <= 1e-5 * np.abs(b[k]) + 1e-7), (k, err.max())


def _load_twin(src, p, meas, seed):
    """An FP64 checker holding the f32 checker's state after the assignment stage."""
    twin = OracleDOGM(p, seed=seed, fp64=True)
    twin.set_flags(True, True)
    twin.update_measurement_grid(meas)
    parts = src.get_particles()
    twin.set_particles(state=parts["state"], grid_cell_idx=parts["grid_cell_idx"], weight=parts["weight"],
                       associated=parts["associated"])
    twin.set_grid_cells(src.get_grid_cells())
    twin.set_weight_array(src.get_weight_array())
    twin.epoch = src.epoch
    return twin


def test_fp32_against_fp64_moments_and_birth_weights():
    """200 x 200 cells: from identical state after the assignment, the f32 path's occupancy, persistent,
    birth and moment stages against the same stages in double precision.  Masses within 1e-5 relative;
    velocity mean / variance / covariance within 1e-5 of the cell's second moment (E[v^2] = var +
    mean^2, its square root for the mean, sqrt(E[vx^2] E[vy^2]) for the covariance); w_A and w_UA
    within 1e-5 relative wherever both assign the cell the same number of birth particles (the counts
    come from prefixes of f32 vs f64 born masses and may differ by one at a rounding boundary)."""
    gs = 200
    p = small_params(gs=gs, res=0.25, n=120_000, b=12_000)
    sc = small_scene(gs=gs, res=0.25, n_boxes=12, n_beams=720, ego_speed=4.0)
    f32 = OracleDOGM(p, seed=11)
    for f in range(2):
        f32.update_grid(sc.measurement(f), *sc.ego_pose(f), 0.1)
    m = sc.measurement(2)
    f32.update_measurement_grid(m)
    f32.update_pose(*sc.ego_pose(2))
    f32.particle_prediction(0.1)
    f32.particle_assignment()
    f64 = _load_twin(f32, p, m, seed=11)
    for o in (f32, f64):
        o.grid_cell_occupancy_update(0.1)
        o.update_persistent_particles()
        o.initialize_new_particles()
        o.statistical_moments()
    a, b = f32.get_grid_cells(), f64.get_grid_cells()
    occupied = (b["start_idx"] != -1) & (b["pers_occ_mass"] > 0)
    assert occupied.sum() > 1000
    for k in ("occ_mass", "free_mass", "pers_occ_mass", "new_born_occ_mass", "pred_occ_mass", "mu_A", "mu_UA"):
        err = np.abs(a[k].astype(np.float64) - b[k])
        assert np.all(err <= 1e-5 * np.abs(b[k]) + 1e-12), (k, float((err / np.maximum(np.abs(b[k]), 1e-30)).max()))
    m2x = b["var_x_vel"].astype(np.float64) + b["mean_x_vel"].astype(np.float64) ** 2
    m2y = b["var_y_vel"].astype(np.float64) + b["mean_y_vel"].astype(np.float64) ** 2
    scale = {"mean_x_vel": np.sqrt(m2x), "mean_y_vel": np.sqrt(m2y), "var_x_vel": m2x, "var_y_vel": m2y,
             "covar_xy_vel": np.sqrt(m2x * m2y)}
    for k, s in scale.items():
        err = np.abs(a[k].astype(np.float64) - b[k])[occupied]
        assert np.all(err <= 1e-5 * s[occupied]), (k, float((err / s[occupied]).max()))
        assert np.all(a[k][~occupied] == 0) and np.all(b[k][~occupied] == 0), k
    na = np.bincount(f32.get_birth_particles()["grid_cell_idx"], minlength=gs * gs)
    nb = np.bincount(f64.get_birth_particles()["grid_cell_idx"], minlength=gs * gs)
    same = na == nb
    assert same.mean() > 0.99
    for k in ("w_A", "w_UA"):
        err = np.abs(a[k].astype(np.float64) - b[k])[same]
        assert np.all(err <= 1e-5 * np.abs(b[k][same]) + 1e-30), k
