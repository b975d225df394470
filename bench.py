#!/usr/bin/env python3
"""Benchmark of the gfx950 DOGM update on one MI355X.

Measures two forms of the full per-frame filter (measurement copy, pose, predict, assignment,
occupancy, persistent update, births, moments, resampling) on the full-scale workload:
1200 x 1200 cells at 0.1 m, 2e6 persistent + 2e5 birth particles, 40 moving boxes seen by a
1080-beam 60 m lidar, ego at 10 m/s with 0.1 rad/s yaw rate (a grid shift every frame), dt 0.1 s.
The measurement grids of all frames are generated before timing and are resident in HBM; each step
consumes the next one.
 - plain: one dogm_update_grid call per frame (the reference's updateGrid, blocking);
 - batched (the headline `value`): K vehicles of the same workload per dogm_batch_update step, each
   with its own Philox instance, all fed the same frame sequence; frames/s counts K frames per step.
Prints one JSON line (stdout); per-kernel breakdowns go to stderr.

    python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--workload plain_full] [--batch 32] [--cpu-frames 24]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np
from ctypes import c_void_p as C_void

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "dynamic-occupancy-grid-map_amd"))

METRIC = "frames/s @2M particles,1.44M cells; p99 frame latency, batched vs plain"
HBM_PEAK_GBS = 8000.0
# Published figure for this workload size (Nuss et al. as quoted by the reference README:
# 2e6 particles, 2e5 births, 1.44e6 cells at 50 ms per frame on a GTX 980).
PUBLISHED_FPS = 20.0

WORKLOAD_TEXT = {
    "plain_full": "plain per-frame updateGrid, 1200x1200 cells at 0.1 m, 2e6 particles + 2e5 birth, 40 moving "
                  "boxes, 1080-beam 60 m lidar, ego 10 m/s at 0.1 rad/s (grid shift every frame), dt 0.1 s",
    "plain_small": "plain per-frame updateGrid, 250x250 cells at 0.2 m, 3e5 particles + 3e4 birth, 8 moving boxes, "
                   "360-beam 50 m lidar, stationary ego, dt 0.1 s",
}


def kernel_bytes(name: str, N: int, B: int, C: int, passes: int) -> float | None:
    """Algorithmic bytes one launch must move per vehicle (read once what it consumes, write once what
    it hands on); gathers through cell indices count their payload once."""
    table = {
        # ancestor index, its 16 B state and association in; state, idx and association out (the
        # resampling leaves the new set as ancestors, read here; the prior weights are equal after every
        # resampling, so predict leaves them in place and the gathering pass forms the predicted weights)
        "predict": 42 * N,
        "sort_downsweep": 16 * N,                # a pass before the last: keys + values in and out
        # the last pass: keys + values in, the 16 B state gathered, sorted particles (25 B) out (the
        # weight copy is skipped: the persistent update of the same step rewrites it)
        "sort_gather": (8 + 16 + 25) * N,
        "seg_occ": 8 * N,                        # keys + weights
        # start, free, meas occ/free in; 6 fields out; the step's measurement grid copied from the
        # caller's device buffer on the way (16 B in, 16 B out per cell)
        "mass_update": 72 * C,
        "seg_pers": 12 * N,                      # keys + weights in, likelihood-weighted weights out
        "seg_moments": 24 * N,                   # keys, lik-weighted and prior weights, vx, vy in; posterior out
        "rs_merge": 8 * N + 4 * B,               # joint weights in; ancestor indices out
        "birth_particles": 25 * B,
        "birth_cells": 20 * C,
        "reset_bounds": 8 * C,
        "scan_tiles": 4 * C,
    }
    return table.get(name)


# Profiled kernel names (rocprofv3) of the per-kernel timers.
PROFILED_NAME = {
    "predict": "k_predict", "seg_occ": "k_segreduce<OccOp>", "seg_pers": "k_segreduce<PersOp>",
    "seg_moments": "k_moments<4>", "mass_update": "k_mass_update", "rs_merge": "k_rs_merge_exact",
    "birth_cells": "k_scan_apply<SrcBorn,SinkBirth>",
}
# The scatter and gathering passes of the radix sort, by the pass form the tuning gives the step.
SORT_PROFILED_NAME = {
    "reduce_scan": {"sort_downsweep": "k_sort_downsweep<false>", "sort_gather": "k_sort_downsweep<true>"},
    "onesweep": {"sort_downsweep": "k_sort_onesweep<false>", "sort_gather": "k_sort_onesweep<true>"},
}


def profiled_name(kernel: str, vehicles: int, particles: int) -> str | None:
    """Kernel-name prefix of a bench stage in the rocprofv3 summaries; the sort stages resolve to the
    pass form the tuning table selects for a step of `vehicles` vehicles and `particles` particles."""
    if kernel in ("sort_downsweep", "sort_gather"):
        from dogm_ctypes import tuning_lookup
        return SORT_PROFILED_NAME[tuning_lookup(vehicles, particles)["sort"]][kernel]
    return PROFILED_NAME.get(kernel)


def measured_traffic(kernel: str, workload: str, vehicles: int, particles: int) -> tuple[int | None, str | None]:
    """HBM bytes per launch of `kernel` from the committed counter summary of this workload and
    batch size (FETCH_SIZE + WRITE_SIZE, scaled by a copy of known size), if one exists."""
    import profiling
    path = REPO / "profiles" / f"hbm_traffic_{workload}_batch{vehicles}.md"
    name = profiled_name(kernel, vehicles, particles)
    if name is None or not path.exists():
        return None, None
    return profiling.lookup_traffic(str(path), name), str(path.relative_to(REPO))


# Vector-ALU issue: a wave64 VALU instruction occupies a 16-lane SIMD for 4 cycles; 256 CUs x 4 SIMDs
# at the 2.4 GHz peak engine clock.
SIMDS = 1024
ENGINE_HZ = 2.4e9


def valu_issue(kernel: str, workload: str, vehicles: int, particles: int, avg_launch_s: float) -> dict | None:
    """Fraction of the VALU issue rate a launch of `kernel` uses: (waves x VALU instructions per wave
    x 4 cycles) / (SIMDs x clock x launch time), with the instruction count from the committed counter
    summary (SQ_INSTS_VALU / SQ_WAVES at the largest grid recorded, the batched launch)."""
    import profiling
    path = REPO / "profiles" / f"hbm_traffic_{workload}_batch{vehicles}.md"
    name = profiled_name(kernel, vehicles, particles)
    if name is None or not path.exists() or not avg_launch_s:
        return None
    row = profiling.lookup_valu(str(path), name)
    if row is None:
        return None
    grid, valu = row
    waves = grid // 64
    cycles = waves * valu * 4.0
    return {"valu_per_wave": valu, "waves": waves,
            "frac": round(cycles / (SIMDS * ENGINE_HZ * avg_launch_s), 4),
            "source": str(path.relative_to(REPO))}


def frame_bytes(N: int, B: int, C: int, shifted: bool) -> int:
    """Stage-boundary algorithmic bytes of one frame: 172 N + 29 B + 128 C (+16 N + 8 C with a
    grid shift)."""
    return 172 * N + 29 * B + 128 * C + ((16 * N + 8 * C) if shifted else 0)


def cpu_baseline(workload: str, frames: int, meas_list, poses, dt):
    """The checker's f32 restatement timed on this host's cores (its OpenMP loops over particles and
    cells, OMP_NUM_THREADS threads) on the first `frames` frames of the same workload."""
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle_ctypes  # noqa: E402  (checker: host-core baseline only)
    import scene
    from dogm_ctypes import Params
    p = Params(**scene.WORKLOADS[workload]["params"])
    o = oracle_ctypes.OracleDOGM(p)
    times = []
    for f in range(frames):
        x, y, yaw = poses[f]
        t0 = time.perf_counter()
        o.update_grid(meas_list[f], x, y, yaw, dt)
        times.append(time.perf_counter() - t0)
    o.close()
    warm = min(3, len(times) - 1)
    steady = times[warm:]
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), model)
    except OSError:
        pass
    return {
        "value": round(1.0 / float(np.mean(steady)), 4),
        "unit": "frames/s",
        "cores": oracle_ctypes.threads(),
        "kind": "port",
        "sample": f"{frames} frames of the same workload, the first {warm} (warm-up, particle initialisation) "
                  f"excluded; median {1e3 * float(np.median(steady)):.1f} ms/frame, "
                  f"p99 {1e3 * float(np.percentile(steady, 99)):.1f} ms/frame",
        "cpu_model": model,
        "nproc": os.cpu_count(),
    }


def _dominant(per_frame: dict) -> str:
    return max(per_frame, key=lambda k: per_frame[k][0])


def run_plain(args, p, meas_bufs, poses, dt):
    """One dogm_update_grid per frame; returns timing, the warm-up breakdown and the dominant kernel."""
    from dogm_ctypes import DOGM, lib
    g = DOGM(p, seed=123456)
    names = DOGM.kernel_names()
    total = args.warmup + args.steps
    for f in range(args.warmup):
        if f == 1:
            for k in names:
                g.set_kernel_timing(k, True)
        x, y, yaw = poses[f]
        g.update_grid(meas_bufs[f], x, y, yaw, dt)
    breakdown = {k: g.kernel_timing(k) for k in names}
    g.reset_kernel_timing()
    for k in names:
        g.set_kernel_timing(k, False)
    nwf = args.warmup - 1
    per_frame = {k: (ms / nwf, n) for k, (ms, n) in breakdown.items() if n}
    dominant = _dominant(per_frame)
    g.set_kernel_timing(dominant, True)
    lib.dogm_device_synchronize()
    lat = []
    t_start = time.perf_counter()
    for f in range(args.warmup, total):
        x, y, yaw = poses[f]
        t1 = time.perf_counter()
        g.update_grid(meas_bufs[f], x, y, yaw, dt)
        lat.append(time.perf_counter() - t1)
    lib.dogm_device_synchronize()
    elapsed = time.perf_counter() - t_start
    dom_ms, dom_n = g.kernel_timing(dominant)
    g.close()
    return {"elapsed": elapsed, "frames": args.steps, "lat": lat, "per_frame": per_frame, "dominant": dominant,
            "dom_avg_s": dom_ms / max(1, dom_n) / 1e3, "launches_per_step": dom_n / args.steps}


def run_batched(args, k, p, meas_bufs, poses, dt, sc=None):
    """k vehicles of the same workload (Philox instances 0..k-1) per dogm_batch_update step; in step f
    vehicle i consumes resident grid (f + i) mod G at that frame's pose, so the vehicles of a step
    read k distinct grids (each vehicle follows the scene's frame sequence from its own offset)."""
    from dogm_ctypes import Batch, DOGM, lib
    b = Batch()
    ids = [b.add_vehicle(p, 123456, i) for i in range(k)]
    names = DOGM.kernel_names()
    total = args.warmup + args.steps
    ng = len(meas_bufs)

    def step(f):
        b.update(ids, [meas_bufs[(f + i) % ng] for i in range(k)], [poses[(f + i) % ng] for i in range(k)], dt)

    for f in range(args.warmup):
        if f == 1:
            for n in names:
                b.set_kernel_timing(n, True)
        step(f)
    breakdown = {n: b.kernel_timing(n) for n in names}
    b.reset_kernel_timing()
    for n in names:
        b.set_kernel_timing(n, False)
    nwf = args.warmup - 1
    per_step = {n: (ms / nwf, c) for n, (ms, c) in breakdown.items() if c}
    dominant = _dominant(per_step)
    b.set_kernel_timing(dominant, True)
    lib.dogm_device_synchronize()
    lat = []
    t_start = time.perf_counter()
    for f in range(args.warmup, total):
        t1 = time.perf_counter()
        step(f)
        lat.append(time.perf_counter() - t1)
    lib.dogm_device_synchronize()
    elapsed = time.perf_counter() - t_start
    dom_ms, dom_n = b.kernel_timing(dominant)
    b.set_kernel_timing(dominant, False)
    # Roofline leg: the same vehicles with one lane, so that launches of the dominant kernel do not
    # overlap with the other lane's (two lanes run the kernel twice per step, concurrently, and each
    # launch then shares the device).
    b.set_lanes(1)
    step(args.warmup - 1)
    b.reset_kernel_timing()
    b.set_kernel_timing(dominant, True)
    nroof = min(args.steps, 10)
    for f in range(nroof):
        step(args.warmup + f)
    lib.dogm_device_synchronize()
    roof_ms, roof_n = b.kernel_timing(dominant)
    b.set_kernel_timing(dominant, False)
    b.set_lanes(0)  # back to the tuning's lane count
    extraction = run_extraction(args, b, ids, p, step, total) if args.extract_steps > 0 else None
    generation = run_generation(args, b, ids, sc, total) if (args.generation_steps > 0 and sc is not None) else None
    b.close()
    return {"elapsed": elapsed, "frames": args.steps * k, "lat": lat, "per_frame": per_step, "dominant": dominant,
            "dom_avg_s": dom_ms / max(1, dom_n) / 1e3, "vehicles": k, "extraction": extraction,
            "generation": generation,
            "launches_per_step": dom_n / args.steps,
            "roofline_leg": {"dom_avg_s": roof_ms / max(1, roof_n) / 1e3, "launches_per_step": roof_n / nroof,
                             "steps": nroof, "lanes": 1},
            "lanes": "tuning"}


EXTRACT_KERNELS = ("extract_mark", "extract_scan", "extract_compact")


def _percentiles(lat) -> dict:
    a = 1e3 * np.asarray(lat)
    return {"samples": int(a.size), "p50_ms": round(float(np.percentile(a, 50)), 4),
            "p99_ms": round(float(np.percentile(a, 99)), 4), "max_ms": round(float(a.max()), 4)}


def run_latency(args, p, sc, dt, meas_bufs) -> dict:
    """Latency percentiles over `args.latency_frames` frames of the same workload: the plain path (one
    blocking updateGrid per frame) and the batched path (one blocking step of `args.batch` vehicles
    per frame, vehicle i at frame f + i).  The ego pose advances every frame (a grid shift every frame);
    the measurement grids cycle through the timed region's resident grids (the scene's boxes leave the sensor range after
    about a minute of driving, so later frames of the scene would be empty).  The first
    `args.warmup` frames of each path are not counted."""
    from dogm_ctypes import Batch, DOGM
    nf = args.latency_frames
    bufs = [meas_bufs[f % len(meas_bufs)] for f in range(nf)]
    poses = [sc.ego_pose(f) for f in range(nf)]
    out = {"frames": nf, "warmup_frames_excluded": args.warmup,
           "note": (f"one latency sample per blocking call; grids cycle through {len(meas_bufs)} resident "
                    "grids while the ego pose advances every frame")}
    g = DOGM(p, seed=123456)
    lat = []
    for f in range(nf):
        x, y, yaw = poses[f]
        t1 = time.perf_counter()
        g.update_grid(bufs[f], x, y, yaw, dt)
        lat.append(time.perf_counter() - t1)
    g.close()
    out["plain"] = _percentiles(lat[args.warmup:])
    if args.batch > 0:
        k = args.batch
        b = Batch()
        ids = [b.add_vehicle(p, 123456, i) for i in range(k)]
        lat = []
        for f in range(nf):
            t1 = time.perf_counter()
            b.update(ids, [bufs[(f + i) % nf] for i in range(k)], [poses[(f + i) % nf] for i in range(k)], dt)
            lat.append(time.perf_counter() - t1)
        b.close()
        out["batched"] = dict(_percentiles(lat[args.warmup:]), vehicles=k)
    return out


def run_uneven(args) -> dict:
    """Config 4: `args.uneven` grids of uneven size (sides from {128, 256, 512, 768, 1024} at 0.1 m,
    N = 1.5 x cells, B = N / 10, each with its own scene, 720-beam lidar and ego motion) per
    dogm_batch_update step.  Every step generates the grids on the device from the scans
    (dogm_lidar_generate_many per side, into the vehicles' measurement buffers) and then updates all
    vehicles; both parts are timed, the update also alone."""
    import fleet as fl
    from dogm_ctypes import Batch, lib
    vehicles = fl.make_fleet(args.uneven, seed=args.seed + 100)
    sides = [v.side for v in vehicles]
    bank = fl.ScanBank(sides, frames=4)
    b = Batch()
    ids = [b.add_vehicle(v.params, 123456, i) for i, v in enumerate(vehicles)]
    by_side = {}
    for j, v in enumerate(vehicles):
        by_side.setdefault(v.side, []).append(j)
    outs = {sd: [b.measurement_buffer(ids[j]) for j in js] for sd, js in by_side.items()}
    byte_sum = 0
    for v in vehicles:
        c = v.side * v.side
        n, nb = v.params.particle_count, v.params.new_born_particle_count
        byte_sum += frame_bytes(n, nb, c, True)

    def generate(k):
        for sd, js in by_side.items():
            r = bank.ranges[sd][k % len(bank.ranges[sd])]
            bank.lidar[sd].generate_many(np.stack([r] * len(js)), outs[sd])

    def update(k):
        b.update(ids, [None] * len(ids), [v.pose(k) for v in vehicles], 0.1)

    for k in range(2):
        generate(k)
        update(k)
    lib.dogm_device_synchronize()
    steps = args.uneven_steps
    t_gen = t_upd = 0.0
    for k in range(2, 2 + steps):
        t0 = time.perf_counter()
        generate(k)
        t1 = time.perf_counter()
        update(k)
        t2 = time.perf_counter()
        t_gen += t1 - t0
        t_upd += t2 - t1
    frames = steps * len(ids)
    # The same steps with the update's kernel sequence replayed from a HIP graph (opt-in).
    b.set_graph(True)
    for k in range(2 + steps, 4 + steps):
        generate(k)
        update(k)
    lib.dogm_device_synchronize()
    t_upd_graph = 0.0
    for k in range(4 + steps, 4 + 2 * steps):
        generate(k)
        lib.dogm_device_synchronize()
        t1 = time.perf_counter()
        update(k)
        t_upd_graph += time.perf_counter() - t1
    b.close()
    bank.close()
    return {"vehicles": len(ids), "particles": sum(v.params.particle_count for v in vehicles),
            "sides": {str(sd): len(js) for sd, js in sorted(by_side.items())}, "steps": steps,
            "update_frames_per_s_graph": round(frames / t_upd_graph, 1),
            "frames_per_s": round(frames / (t_gen + t_upd), 1), "update_frames_per_s": round(frames / t_upd, 1),
            "ms_per_step": round((t_gen + t_upd) / steps * 1e3, 3), "update_ms_per_step": round(t_upd / steps * 1e3, 3),
            "generation_ms_per_step": round(t_gen / steps * 1e3, 3),
            "algorithmic_bytes_per_step": byte_sum,
            "update_hbm_fraction": round(byte_sum / (t_upd / steps) / (HBM_PEAK_GBS * 1e9), 4),
            "note": "grids generated on the device from lidar scans every step; ego shift every frame"}


def run_generation(args, b, ids, sc, total):
    """Measurement grids generated on the device from the scene's lidar ranges, straight into the
    vehicles' measurement buffers (dogm_lidar_generate_many, one call per step for all vehicles):
    grids per second and the algorithmic bytes of the three launches -- a 16-byte clear per cell,
    a 4-byte free-mass store per free sample, an 8-byte store per return, and the ranges."""
    from dogm_ctypes import Lidar, lib
    cs, sn = sc.beam_directions()
    c = sc.cfg
    ranges = [np.stack([sc.ranges(f)] * len(ids)) for f in range(min(total, 8))]
    outs = [b.measurement_buffer(v) for v in ids]
    cells = c.grid_size ** 2
    step = 0.5 * c.resolution
    n_steps = int(np.ceil(c.max_range / step)) + 1
    dist = np.arange(n_steps) * step
    free_samples = sum(int(np.count_nonzero(dist[None, :] < np.minimum(r[0], c.max_range)[:, None]))
                       for r in ranges) / len(ranges)
    returns = sum(int(np.count_nonzero(np.isfinite(r[0]))) for r in ranges) / len(ranges)
    per_grid = 16 * cells + 4 * free_samples + 8 * returns + 8 * c.n_beams
    with Lidar(**sc.lidar_params(), beam_cos=cs, beam_sin=sn) as lidar:
        for f in range(2):
            lidar.generate_many(ranges[f % len(ranges)], outs)
        lib.dogm_device_synchronize()
        reps = args.generation_steps
        t0 = time.perf_counter()
        for f in range(reps):
            lidar.generate_many(ranges[f % len(ranges)], outs)
        el = time.perf_counter() - t0
    grids = reps * len(ids)
    return {"grids_per_s": round(grids / el, 1), "us_per_call": round(el / reps * 1e6, 2), "grids_per_call": len(ids),
            "algorithmic_bytes_per_grid": int(per_grid),
            "achieved_GBs_wall": round(per_grid * grids / el / 1e9, 1),
            "note": "wall time of blocking calls (ranges upload, three launches, stream wait)"}


def run_extraction(args, b, ids, p, step, first_frame):
    """Result extraction inside the batched steps (after the timed frames): the demo's per-frame
    post-processing -- 8-bit occupancy image and the cells passing the velocity test at the demo's
    thresholds 0.7 / 4.0 (image_creation.cpp:19-128, main.cpp:51-52) -- for every vehicle of the step.
    Reports each extraction kernel's average launch time against its algorithmic bytes, and the
    read-back size of the compact results against a full getGridCells."""
    from dogm_ctypes import EXTRACT_IMAGE, EXTRACT_VELOCITY, lib
    min_occ, min_vel = 0.7, 4.0
    b.set_extraction(EXTRACT_IMAGE | EXTRACT_VELOCITY, min_occ, min_vel)
    step(first_frame - 1)  # allocates the extraction buffers
    for n in EXTRACT_KERNELS:
        b.set_kernel_timing(n, True)
    lib.dogm_device_synchronize()
    t0 = time.perf_counter()
    for f in range(args.extract_steps):
        step(first_frame - 1 - (f % 2))
    lib.dogm_device_synchronize()
    ms_step = (time.perf_counter() - t0) / args.extract_steps * 1e3
    timing = {n: b.kernel_timing(n) for n in EXTRACT_KERNELS}
    for n in EXTRACT_KERNELS:
        b.set_kernel_timing(n, False)
    gs = int(np.float32(p.size) / np.float32(p.resolution))
    C = gs * gs
    units = -(-C // 256)
    units += (-units) % 8
    counts = []
    occupied = 0
    for vid in ids:
        _, cnt = b.get_cells_with_velocity(vid, capacity=0)
        counts.append(cnt)
        occupied += int(np.count_nonzero(b.get_occupancy(vid) >= np.float32(min_occ)))
    kv = len(ids)
    passing = sum(counts)
    # mark: occ + free of every cell, the five moment fields of cells at or above the occupancy
    # threshold; writes the image byte, one 64-bit ballot per 64 cells and one count per unit.
    # scan: counts in, offsets out.  compact: counts, offsets and ballots of every unit, 4 bytes per
    # passing cell.
    algo = {
        "extract_mark": kv * (8 * C + 1 * C + 8 * (units * 4) + 4 * units) + 20 * occupied,
        "extract_scan": kv * (4 * units + 4 * (units + 1)),
        "extract_compact": kv * (4 * units + 4 * units + 8 * units * 4) + 4 * passing,
    }
    kernels = {}
    for n in EXTRACT_KERNELS:
        ms, launches = timing[n]
        avg_s = ms / max(1, launches) / 1e3
        per_launch = algo[n] * args.extract_steps / max(1, launches)  # the step's vehicles share its launches
        gbs = per_launch / avg_s / 1e9 if avg_s > 0 else None
        kernels[n] = {"avg_launch_us": round(avg_s * 1e6, 3), "launches": launches,
                      "algorithmic_bytes_per_launch": int(per_launch),
                      "achieved_GBs": round(gbs, 1) if gbs else None,
                      "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs else None}
    readback = C + 72 * (passing / kv)
    return {"what": f"8-bit occupancy image + cells with v^T S^-1 v >= {min_vel} at occupancy >= {min_occ}",
            "ms_per_step_with_extraction": round(ms_step, 4), "velocity_cells_per_vehicle": counts,
            "cells_at_or_above_occupancy": occupied, "kernels": kernels,
            "readback_bytes_per_vehicle": int(readback), "full_grid_readback_bytes": 64 * C,
            "readback_ratio": round(64 * C / readback, 1)}


def run_serving(args) -> int:
    """Serving form: a fleet of resident vehicles, 10 Hz arrivals per vehicle, continuous batching
    in virtual time (fleet.simulate); prints one JSON line."""
    import fleet as fl
    import scene
    from dogm_ctypes import DeviceBuffer, Params, Server, lib, vehicle_footprint

    say = lambda m: print(f"[serve] {m}", file=sys.stderr, flush=True)  # noqa: E731
    if args.serve_vehicle == "fleet":
        fleet = fl.make_fleet(args.fleet, seed=args.seed)
        t0 = time.perf_counter()
        sides = [v.side for v in fleet]
        bank = fl.ScanBank(sides) if args.serve_input == "scan" else fl.measurement_bank(sides)
        say(f"{args.serve_input} bank for sides {sorted(set(sides))} in {time.perf_counter() - t0:.1f} s")
        label = (f"{args.fleet} vehicles, grid sides drawn from {{128, 256, 512, 768, 1024}} at 0.1 m, N = 1.5 x cells, "
                 "B = N/10")
    else:
        w = scene.WORKLOADS["plain_full"]
        p_full = Params(**w["params"])
        fleet = fl.make_fleet(args.fleet, seed=args.seed, sides=(1200,))
        for v in fleet:
            v.params = p_full
        sc = scene.make_scene("plain_full", seed=args.seed)
        if args.serve_input == "scan":
            bank = fl.ScanBank([1200], scenes={1200: sc})
        else:
            bufs = []
            for f in range(4):
                m = np.ascontiguousarray(sc.measurement(f))
                d = DeviceBuffer(m.nbytes)
                d.upload(m)
                bufs.append(d)
            bank = {1200: bufs}
        label = f"{args.fleet} vehicles at 1200x1200 cells, 2e6 particles + 2e5 birth each"
    budget = int(args.budget_gb * (1 << 30))
    srv = Server(hbm_budget_bytes=budget, max_batch=args.max_batch)
    vids, rejected = [], 0
    for i, v in enumerate(fleet):
        vid = srv.admit(v.params, 123456, i)
        if vid is None:
            rejected += 1
        vids.append(vid)
    admitted = [i for i, v in enumerate(vids) if v is not None]
    fleet_adm = [fleet[i] for i in admitted]
    vids_adm = [vids[i] for i in admitted]
    say(f"admitted {len(admitted)} of {len(fleet)} vehicles, {srv.resident_bytes() / 2**30:.2f} GiB resident")
    # Warm-up: two frames per vehicle (the first initialises its particles), not measured.
    for k in range(2):
        for j, v in enumerate(fleet_adm):
            if isinstance(bank, fl.ScanBank):
                bank.submit(srv, vids_adm[j], v.side, k, v.pose(k), 0.1, 0.0)
            else:
                srv.submit(vids_adm[j], bank[v.side][k % len(bank[v.side])], v.pose(k), 0.1, 0.0)
        while srv.queued():
            srv.step(0.0)
        srv.poll(1 << 16)
    lib.dogm_device_synchronize()
    sweep = []
    for mult in [float(x) for x in args.sweep.split(",")]:
        secs = args.serve_seconds if mult == 1.0 else args.sweep_seconds
        ev = fl.arrivals(fleet_adm, secs, rate_multiplier=mult, seed=args.seed + int(10 * mult))
        ev = [(t, v, k + 2) for (t, v, k) in ev]
        say(f"x{mult}: {len(ev)} frames over {secs} s simulated")
        t0 = time.perf_counter()
        r = fl.simulate(srv, fleet_adm, vids_adm, bank, ev, progress=say)
        r["rate_multiplier"] = mult
        r["wall_seconds"] = round(time.perf_counter() - t0, 2)
        say(json.dumps(r))
        sweep.append(r)
    base = sweep[0]
    ok = [r for r in sweep if r["p99_ms"] < 10.0]
    best = max(ok, key=lambda r: r["frames_per_s"]) if ok else None
    out = {
        "metric": "serving: frames/s and p50/p99 frame latency (arrival to result), continuous batching",
        "value": base["frames_per_s"],
        "unit": "frames/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": label, "arrivals": "10 Hz per vehicle, +-20 ms jitter, random phase",
                   "input": ("lidar scans (720 or 1080 beams); measurement grids generated on the device inside each "
                             "step" if args.serve_input == "scan" else "device-resident measurement grids"),
                   "hbm_budget_gb": args.budget_gb, "max_batch": args.max_batch,
                   "clock": "virtual time; every step's duration is the measured GPU step"},
        "admitted": len(admitted), "rejected": rejected,
        "resident_gib": round(srv.resident_bytes() / 2**30, 3),
        "p50_ms": base["p50_ms"], "p99_ms": base["p99_ms"],
        "sweep": sweep,
        "max_frames_per_s_with_p99_under_10ms": best["frames_per_s"] if best else None,
    }
    if args.serve_wall_seconds > 0:
        # The nominal stream once more on the wall clock: the worker thread steps while arrivals are
        # submitted in real time.
        k0 = 2 + int(10 * max(args.serve_seconds, args.sweep_seconds)) + 2
        ev = [(t, v, k + k0) for (t, v, k) in fl.arrivals(fleet_adm, args.serve_wall_seconds, rate_multiplier=1.0,
                                                           seed=args.seed + 1000)]
        say(f"wall clock: {len(ev)} frames over {args.serve_wall_seconds} s")
        out["wall_clock"] = fl.run_wallclock(srv, fleet_adm, vids_adm, bank, ev)
        say(json.dumps(out["wall_clock"]))
    print(json.dumps(out))
    srv.close()
    if isinstance(bank, fl.ScanBank):
        bank.close()
    else:
        for bl in bank.values():
            for d in bl:
                d.free()
    return 0


def run_accuracy(args) -> int:
    """Tracking accuracy (accuracy.run) of the plain filter in each CDF mode, of a batch of vehicles
    and of the same vehicles served from scans, in both ego modes and sensor frames, against the
    scene's ground-truth boxes; one JSON line."""
    import accuracy
    say = lambda m: print(f"[accuracy] {m}", file=sys.stderr, flush=True)  # noqa: E731
    rows = []
    k = max(1, min(args.batch, 8))
    for form, mode, veh, ego, frame in (("plain", "exact", 1, "reference", "world"),
                                        ("plain", "fp64", 1, "reference", "world"),
                                        ("plain", "fp32c", 1, "reference", "world"),
                                        ("plain", "fp32", 1, "reference", "world"),
                                        ("batched", "exact", k, "reference", "world"),
                                        ("serving", "exact", k, "reference", "world"),
                                        ("plain", "exact", 1, "metric", "world"),
                                        ("plain", "exact", 1, "reference", "vehicle"),
                                        ("plain", "exact", 1, "metric", "vehicle")):
        t0 = time.perf_counter()
        r = accuracy.run(args.workload, args.accuracy_frames, form=form, cdf_mode=mode, vehicles=veh, ego_mode=ego,
                         sensor_frame=frame)
        r.pop("per_frame")
        r["vehicles"] = veh
        r["seconds"] = round(time.perf_counter() - t0, 2)
        say(json.dumps(r))
        rows.append(r)
    print(json.dumps({"metric": "tracking accuracy: MAE / RMSE of cluster position [m] and velocity [m/s] against "
                                "ground-truth boxes", "workload": WORKLOAD_TEXT[args.workload], "rows": rows}))
    return 0


# Every kernel a resampling stage launches (CDF tile totals and top scan, float-mode CDF, spacings, merge).
RESAMPLING_KERNELS = ("scan_tiles", "scan_top", "cdf_write", "exp_tiles", "rs_merge")

PRECISION_PARAMS = dict(size=204.8, resolution=0.1, particle_count=8_000_000, new_born_particle_count=800_000,
                        persistence_prob=0.99, stddev_process_noise_position=0.1,
                        stddev_process_noise_velocity=2.0, birth_prob=0.02, stddev_velocity=30.0,
                        init_max_velocity=15.0, freespace_discount=0.01)


def run_precision(args) -> int:
    """Precision form: 2048 x 2048 cells at 0.1 m, 8e6 + 8e5 particles, f32 state.  Every frame runs
    the stages up to the moments, then the joint weights are replaced by 10^(-9 U) (U from a Philox
    stream, so the CDF spans ~9 decades) and resampled with the exact, FP64 and FP32 CDF scans.
    Reported per mode: ancestor agreement with a sequential FP64 CDF of the same weights and draws,
    the relative error of the CDF total, and the resampling stage time.  On two frames the per-cell
    masses of the f32 path are compared with the f64 checker run from the same state."""
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle_ctypes  # noqa: E402  (checker: FP64 reference only)
    import scene
    from dogm_ctypes import DOGM, Params, lib

    say = lambda m: print(f"[precision] {m}", file=sys.stderr, flush=True)  # noqa: E731
    p = Params(**PRECISION_PARAMS)
    N, B = p.particle_count, p.new_born_particle_count
    cfg = scene.SceneConfig(grid_size=2048, resolution=0.1, n_boxes=60, n_beams=1440, max_range=100.0,
                            ego_speed=10.0, ego_yaw_rate=0.1, seed=args.seed)
    sc = scene.Scene(cfg)
    frames = args.precision_frames
    modes = ["exact", "fp64", "fp32c", "fp32"]
    handles = {}
    for mo in modes:
        g = DOGM(p, seed=123456, instance=0)
        g.set_cdf_mode(mo)
        for kn in RESAMPLING_KERNELS:
            g.set_kernel_timing(kn, True)
        handles[mo] = g
    rng = np.random.Generator(np.random.Philox(args.seed))
    agree = {mo: [] for mo in modes}
    tot_err = {mo: [] for mo in modes}
    rs_ms = {mo: [] for mo in modes}
    mass = []
    for f in range(frames):
        t0 = time.perf_counter()
        m = sc.measurement(f)
        x, y, yaw = sc.ego_pose(f)
        check_mass = f in (2, frames - 1)
        for mo, g in handles.items():
            g.update_measurement_grid(m)
            g.update_pose(x, y, yaw)
            g.particle_prediction(0.1)
            g.particle_assignment()
            twin = _fp64_twin(g, m, oracle_ctypes, p) if (check_mass and mo == "exact") else None
            g.grid_cell_occupancy_update(0.1)
            g.update_persistent_particles()
            if twin is not None:
                twin.grid_cell_occupancy_update(0.1)
                twin.update_persistent_particles()
                mass.append(_compare_masses(f, g.get_grid_cells(), twin.get_grid_cells()))
                say(f"frame {f} masses vs f64: " + ", ".join(f"{k} {mass[-1][k]['max_rel']:.2e}" for k in MASS_FIELDS))
                twin.close()
            g.initialize_new_particles()
            g.statistical_moments()
        w = (10.0 ** (-9.0 * rng.random(N))).astype(np.float32)
        bw = (10.0 ** (-9.0 * rng.random(B))).astype(np.float32)
        cdf = np.cumsum(np.concatenate([w, bw]).astype(np.float64))
        total = cdf[-1]
        for mo, g in handles.items():
            g.set_weight_array(w)
            g.set_birth_particles(weight=bw)
            g.reset_kernel_timing()
            epoch = _handle_epoch(g)
            g.resampling()
            ms = sum(g.kernel_timing(k)[0] for k in RESAMPLING_KERNELS)
            rs_ms[mo].append(ms)
            u = oracle_ctypes.spacing_fractions(123456, 0, epoch, N)
            ref = np.minimum(np.searchsorted(cdf, total * u, side="left"), N + B - 1)
            anc = g.get_resampled_indices()
            agree[mo].append(float(np.mean(anc == ref)))
            nw = g.get_particles()["weight"][0]
            tot_err[mo].append(abs(float(nw) * N - total) / total)
        say(f"frame {f}: agreement " + ", ".join(f"{mo} {agree[mo][-1]:.6f}" for mo in modes) +
            f" ({time.perf_counter() - t0:.1f} s)")
    out = {
        "metric": "precision: resampling ancestor agreement with a sequential FP64 CDF, 9-decade weights",
        "value": round(min(agree["exact"]), 6),
        "unit": "fraction of ancestors equal (exact mode, worst frame)",
        "n_gpus": 1,
        "dtype": "f32 state and weights; CDF exact / f64 / compensated f32 / f32",
        "data": "synthetic",
        "config": {"workload": "2048x2048 cells at 0.1 m, 8e6 particles + 8e5 birth, 60 boxes, 1440-beam 100 m lidar; "
                               "joint weights replaced by 10^(-9 U) before every resampling",
                   "frames": frames},
        "modes": {mo: {"agreement_mean": round(float(np.mean(agree[mo])), 6),
                       "agreement_min": round(float(np.min(agree[mo])), 6),
                       "new_weight_total_rel_error_max": float(np.max(tot_err[mo])),
                       "resampling_ms_mean": round(float(np.mean(rs_ms[mo])), 4)} for mo in modes},
        "masses_vs_fp64_checker": mass,
    }
    print(json.dumps(out))
    for g in handles.values():
        g.close()
    return 0


def _handle_epoch(g) -> int:
    """Philox epoch of the handle's next random stage: the resampling draws use it."""
    from dogm_ctypes import lib
    return int(lib.dogm_get_epoch(g.h))


def _fp64_twin(g, meas, oracle_ctypes, p):
    """The f64 checker loaded with the GPU handle's current state (after the assignment stage)."""
    o = oracle_ctypes.OracleDOGM(p, seed=123456, fp64=True)
    o.set_flags(True, True)
    o.update_measurement_grid(meas)
    parts = g.get_particles()
    o.set_particles(state=parts["state"], grid_cell_idx=parts["grid_cell_idx"], weight=parts["weight"],
                    associated=parts["associated"])
    o.set_grid_cells(g.get_grid_cells())
    o.set_weight_array(g.get_weight_array())
    return o


MASS_FIELDS = ("occ_mass", "free_mass", "pers_occ_mass", "new_born_occ_mass", "pred_occ_mass", "mu_A", "mu_UA")
MOMENT_FIELDS = ("mean_x_vel", "mean_y_vel", "var_x_vel", "var_y_vel", "covar_xy_vel")
BIRTH_WEIGHT_FIELDS = ("w_A", "w_UA")


def _compare_all(frame, gcells, ocells, g_birth_idx, o_birth_idx) -> dict:
    """Error of every per-cell output of the f32 path against the FP64 checker run from the same state:
    masses (relative), velocity moments (relative to the cell's second moment E[v^2] = var + mean^2:
    its square root for the means, sqrt(E[vx^2] E[vy^2]) for the covariance, over cells holding
    persistent mass) and w_A / w_UA (relative, over cells given the same number of birth particles;
    the counts come from f32 vs f64 prefixes of the born masses and may differ by one)."""
    out = _compare_masses(frame, gcells, ocells)
    b = ocells
    occ = (np.asarray(b["start_idx"]) != -1) & (np.asarray(b["pers_occ_mass"]) > 0)
    m2x = np.asarray(b["var_x_vel"], np.float64) + np.asarray(b["mean_x_vel"], np.float64) ** 2
    m2y = np.asarray(b["var_y_vel"], np.float64) + np.asarray(b["mean_y_vel"], np.float64) ** 2
    scale = {"mean_x_vel": np.sqrt(m2x), "mean_y_vel": np.sqrt(m2y), "var_x_vel": m2x, "var_y_vel": m2y,
             "covar_xy_vel": np.sqrt(m2x * m2y)}
    for k in MOMENT_FIELDS:
        err = np.abs(gcells[k].astype(np.float64) - np.asarray(b[k], np.float64))[occ]
        sc = np.maximum(scale[k][occ], 1e-30)
        rel = err / sc
        out[k] = {"max_rel_to_second_moment": float(rel.max()) if rel.size else 0.0,
                  "within_1e-5": float(np.mean(rel <= 1e-5)) if rel.size else 1.0, "cells": int(occ.sum())}
    ncells = gcells.size
    same = np.bincount(g_birth_idx, minlength=ncells) == np.bincount(o_birth_idx, minlength=ncells)
    out["birth_counts_equal"] = float(np.mean(same))
    for k in BIRTH_WEIGHT_FIELDS:
        ov = np.asarray(b[k], np.float64)[same]
        rel = np.abs(gcells[k].astype(np.float64)[same] - ov) / np.maximum(np.abs(ov), 1e-30)
        out[k] = {"max_rel": float(rel.max()) if rel.size else 0.0, "within_1e-5": float(np.mean(rel <= 1e-5))}
    return out


def run_fp64_errors(p, meas, poses, dt, frames=(2, 3)) -> dict:
    """Config-2 frames of the plain f32 GPU path against the FP64 checker: after the assignment stage
    the checker takes the GPU handle's state (particles, grid, weights, RNG epoch) and both run the
    occupancy, persistent, birth and moment stages (statistical_moments.cu:58-106 for the moments)."""
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle_ctypes  # noqa: E402  (checker: FP64 reference only)
    from dogm_ctypes import DOGM
    g = DOGM(p, seed=123456)
    rows = []
    for f in range(max(frames) + 1):
        m = meas[f]
        x, y, yaw = poses[f]
        if f not in frames:
            g.update_grid(m, x, y, yaw, dt)
            continue
        g.update_measurement_grid(m)
        g.update_pose(x, y, yaw)
        g.particle_prediction(dt)
        g.particle_assignment()
        twin = _fp64_twin(g, m, oracle_ctypes, p)
        twin.epoch = _handle_epoch(g)
        for o in (g, twin):
            o.grid_cell_occupancy_update(dt)
            o.update_persistent_particles()
            o.initialize_new_particles()
            o.statistical_moments()
        rows.append(_compare_all(f, g.get_grid_cells(), twin.get_grid_cells(),
                                 g.get_birth_particles()["grid_cell_idx"], twin.get_birth_particles()["grid_cell_idx"]))
        twin.close()
        g.resampling()
    g.close()
    worst = {}
    for k in MASS_FIELDS + BIRTH_WEIGHT_FIELDS:
        worst[k] = max(r[k]["max_rel"] for r in rows)
    for k in MOMENT_FIELDS:
        worst[k] = max(r[k]["max_rel_to_second_moment"] for r in rows)
    return {"frames": list(frames), "tolerance": "1e-5: relative for masses, w_A and w_UA; relative to the cell's "
                                                 "second moment for the velocity moments",
            "max_error": worst, "all_within_tolerance": bool(all(v <= 1e-5 for v in worst.values())),
            "per_frame": rows}


def run_plain_small(args) -> dict:
    """Config 1 (250 x 250 cells at 0.2 m, 3e5 + 3e4 particles, stationary ego): plain per-frame
    updateGrid with resident grids, frames/s, p50/p99 and the frame's HBM fraction."""
    import scene
    from dogm_ctypes import DOGM, DeviceBuffer, Params
    w = scene.WORKLOADS["plain_small"]
    p = Params(**w["params"])
    sc = scene.make_scene("plain_small", seed=args.seed)
    nf = w["frames"]
    bufs = []
    for f in range(nf):
        d = DeviceBuffer(sc.measurement(f).nbytes)
        d.upload(sc.measurement(f))
        bufs.append(d)
    poses = [sc.ego_pose(f) for f in range(nf)]
    g = DOGM(p, seed=123456)
    lat = []
    frames = max(200, nf)
    for f in range(frames + args.warmup):
        k = f % nf
        t1 = time.perf_counter()
        g.update_grid(bufs[k], *poses[k], float(w["dt"]))
        lat.append(time.perf_counter() - t1)
    g.close()
    for d in bufs:
        d.free()
    lat = lat[args.warmup:]
    gs = 250
    fps = len(lat) / float(np.sum(lat))
    fb = frame_bytes(p.particle_count, p.new_born_particle_count, gs * gs, False)
    return {"workload": WORKLOAD_TEXT["plain_small"], "frames": len(lat), "frames_per_s": round(fps, 1),
            **{k: v for k, v in _percentiles(lat).items() if k != "samples"},
            "frame_hbm_fraction": round(fb * fps / (HBM_PEAK_GBS * 1e9), 4)}


def run_precision_leg(args) -> dict:
    """Config 3, short: run_precision on 5 frames (ancestor agreement per CDF mode)."""
    import contextlib
    import io
    a = argparse.Namespace(**vars(args))
    a.precision_frames = 5
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        run_precision(a)
    r = json.loads(buf.getvalue().strip().splitlines()[-1])
    return {"workload": r["config"]["workload"], "frames": 5, "modes": r["modes"]}


def run_serving_leg(args) -> dict:
    """Config 5, short: 256 vehicles of mixed size (scans in, grids generated inside each step), 10 Hz
    per vehicle with +-20 ms jitter, 200 GB admission budget.  A wall-clock run (the worker thread
    steps while this thread submits every frame at its arrival time; latency = result time - arrival
    time on the server's clock) and a virtual-time run of the same stream."""
    import fleet as fl
    from dogm_ctypes import Server, lib
    vehicles = fl.make_fleet(256, seed=args.seed)
    bank = fl.ScanBank([v.side for v in vehicles])
    srv = Server(hbm_budget_bytes=200 * 10**9)
    vids = [srv.admit(v.params, 123456, i) for i, v in enumerate(vehicles)]
    assert all(v is not None for v in vids), "the config-5 fleet fits the 200 GB budget"
    for k in range(2):
        for j, v in enumerate(vehicles):
            bank.submit(srv, vids[j], v.side, k, v.pose(k), 0.1, 0.0)
        while srv.queued():
            srv.step(0.0)
        srv.poll(1 << 16)
    lib.dogm_device_synchronize()
    secs = args.serve_wall_seconds
    ev = [(t, v, k + 2) for (t, v, k) in fl.arrivals(vehicles, secs, rate_multiplier=1.0, seed=args.seed)]
    wall = fl.run_wallclock(srv, vehicles, vids, bank, ev)
    ev2 = [(t, v, k + 2 + int(10 * secs) + 2) for (t, v, k) in fl.arrivals(vehicles, secs, rate_multiplier=1.0,
                                                                             seed=args.seed + 1)]
    virt = fl.simulate(srv, vehicles, vids, bank, ev2)
    srv.close()
    bank.close()
    return {"workload": "256 vehicles, grid sides from {128, 256, 512, 768, 1024} at 0.1 m, N = 1.5 x cells, B = N/10, "
                        "lidar scans in (grids generated on the device inside each step), 200 GB budget",
            "arrivals": "10 Hz per vehicle, +-20 ms jitter, random phase",
            "wall_clock": wall, "virtual_time": virt}




def _compare_masses(frame, gcells, ocells):
    """Relative error of the f32 path against the f64 checker per mass field: |g - o| / max(|o|, 1e-6)."""
    out = {"frame": frame}
    for k in MASS_FIELDS:
        gv = gcells[k].astype(np.float64)
        ov = np.asarray(ocells[k], np.float64)
        rel = np.abs(gv - ov) / np.maximum(np.abs(ov), 1e-6)
        out[k] = {"max_rel": float(rel.max()), "within_1e-5": float(np.mean(rel <= 1e-5))}
    return out


def _tuning_report(N: int, kveh: int, uneven: dict | None) -> dict:
    from dogm_ctypes import lib, tuning_lookup
    rep = {"table_entries": int(lib.dogm_tuning_entries()), "plain": tuning_lookup(1, N)}
    if kveh:
        rep["batched"] = tuning_lookup(kveh, kveh * N)
    if uneven:
        rep["batched_uneven"] = tuning_lookup(uneven["vehicles"], uneven["particles"])
    return rep


def _summary(r):
    fps = r["frames"] / r["elapsed"]
    return {"frames_per_s": round(fps, 3), "ms_per_step": round(1e3 * r["elapsed"] * len(r["lat"]) ** -1, 4),
            "p50_ms": round(1e3 * float(np.percentile(r["lat"], 50)), 4),
            "p99_ms": round(1e3 * float(np.percentile(r["lat"], 99)), 4)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="plain_full", choices=sorted(WORKLOAD_TEXT))
    ap.add_argument("--batch", type=int, default=32,
                    help="vehicles per batched step (0: plain form only); the batched form is the headline")
    ap.add_argument("--cpu-frames", type=int, default=24, help="frames of the host-core baseline (0: skip)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--precision-frames", type=int, default=30, help="frames of the precision form")
    ap.add_argument("--accuracy-frames", type=int, default=None, help="frames of the accuracy form (default: workload's)")
    ap.add_argument("--form", default="both", choices=["both", "plain", "serving", "precision", "accuracy"],
                    help="both: plain + batched (default); serving: continuous batching of a vehicle fleet")
    ap.add_argument("--serve-vehicle", default="fleet", choices=["fleet", "full"],
                    help="fleet: mixed grid sides; full: every vehicle at 1200x1200 / 2e6 particles")
    ap.add_argument("--fleet", type=int, default=256, help="vehicles of the serving fleet")
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="simulated seconds at the nominal rate")
    ap.add_argument("--sweep", default="1,2,4,8", help="arrival-rate multipliers (the first is the nominal run)")
    ap.add_argument("--sweep-seconds", type=float, default=10.0, help="simulated seconds per sweep point")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="device-memory admission budget")
    ap.add_argument("--serve-input", default="scan", choices=["scan", "grid"],
                    help="serving frames as lidar scans (grids generated on the device) or as ready grids")
    ap.add_argument("--max-batch", type=int, default=0, help="frames per serving step (0: no limit)")
    ap.add_argument("--uneven", type=int, default=64, help="grids of uneven size in the config-4 leg (0: skip)")
    ap.add_argument("--uneven-steps", type=int, default=100, help="timed steps of the config-4 leg")
    ap.add_argument("--generation-steps", type=int, default=20,
                    help="device measurement-grid generations for the batch after the timed frames (0: skip)")
    ap.add_argument("--extract-steps", type=int, default=5,
                    help="batched steps with result extraction after the timed frames (0: skip)")
    ap.add_argument("--latency-frames", type=int, default=1000,
                    help="distinct frames of the latency-percentile leg after the timed region (0: skip)")
    ap.add_argument("--config-legs", type=int, default=1,
                    help="short legs of configs 1, 3 and 5 and the FP64 error report in the default line (0: skip)")
    ap.add_argument("--serve-wall-seconds", type=float, default=10.0, help="wall-clock seconds of the config-5 leg")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="one device-to-device copy of this many MiB before the frames (scales HBM counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}): every box in this setup has a single "
              "MI355X. Filter instances are independent, so more GPUs would each run this engine on their own "
              "share of vehicles.", file=sys.stderr)
        return 2
    assert args.warmup >= 2, "--warmup must be at least 2"
    if args.form == "serving":
        return run_serving(args)
    if args.form == "precision":
        return run_precision(args)
    if args.form == "accuracy":
        return run_accuracy(args)
    if args.form == "plain":
        args.batch = 0

    import scene
    from dogm_ctypes import DOGM, DeviceBuffer, Params, lib

    w = scene.WORKLOADS[args.workload]
    p = Params(**w["params"])
    dt = float(w["dt"])
    sc = scene.make_scene(args.workload, seed=args.seed)
    total = args.warmup + args.steps
    t0 = time.perf_counter()
    meas = [sc.measurement(f) for f in range(total)]
    poses = [sc.ego_pose(f) for f in range(total)]
    print(f"[bench] generated {total} measurement grids in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)
    bufs = []
    for m in meas:
        d = DeviceBuffer(m.nbytes)
        d.upload(m)
        bufs.append(d)

    if args.copy_calibration_mb > 0:
        # One copy kernel dispatch that reads and writes exactly this many bytes (the reference for
        # scaling the FETCH_SIZE / WRITE_SIZE counters of a profiling run).
        nb = args.copy_calibration_mb << 20
        src, dst = DeviceBuffer(nb), DeviceBuffer(nb)
        assert lib.dogm_device_copy(C_void(dst.ptr), C_void(src.ptr), nb) == 0
        src.free()
        dst.free()

    gs = int(np.float32(p.size) / np.float32(p.resolution))
    N, B, C = p.particle_count, p.new_born_particle_count, gs * gs

    plain = run_plain(args, p, bufs, poses, dt)
    print(f"[bench] plain: {_summary(plain)}", file=sys.stderr, flush=True)
    head = plain
    batched = None
    if args.batch > 0:
        batched = run_batched(args, args.batch, p, bufs, poses, dt, sc)
        print(f"[bench] batched x{args.batch}: {_summary(batched)}", file=sys.stderr, flush=True)
        head = batched

    uneven = None
    if args.batch > 0 and args.uneven > 0:
        uneven = run_uneven(args)
        print(f"[bench] uneven x{args.uneven}: {json.dumps(uneven)}", file=sys.stderr, flush=True)

    legs = {}
    if args.config_legs and args.form == "both":
        t0 = time.perf_counter()
        legs["fp64_errors"] = run_fp64_errors(p, meas, poses, dt) if args.workload == "plain_full" else None
        print(f"[bench] fp64 errors ({time.perf_counter() - t0:.1f} s): "
              f"{json.dumps(legs['fp64_errors']['max_error']) if legs['fp64_errors'] else None}", file=sys.stderr,
              flush=True)
        t0 = time.perf_counter()
        legs["plain_small"] = run_plain_small(args)
        print(f"[bench] config 1 ({time.perf_counter() - t0:.1f} s): {json.dumps(legs['plain_small'])}",
              file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        legs["precision"] = run_precision_leg(args)
        print(f"[bench] config 3 ({time.perf_counter() - t0:.1f} s): {json.dumps(legs['precision'])}",
              file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        legs["serving"] = run_serving_leg(args)
        print(f"[bench] config 5 ({time.perf_counter() - t0:.1f} s): {json.dumps(legs['serving'])}",
              file=sys.stderr, flush=True)

    latency = None
    if args.latency_frames > args.warmup:
        latency = run_latency(args, p, sc, dt, bufs)
        print(f"[bench] latency: {json.dumps(latency)}", file=sys.stderr, flush=True)

    for tag, r in (("plain", plain), ("batched", batched)):
        if r is None:
            continue
        print(f"[bench] {tag}: per-kernel ms per warm-up step (launches):", file=sys.stderr)
        for k, (ms, n) in sorted(r["per_frame"].items(), key=lambda kv: -kv[1][0]):
            print(f"  {k:18s} {ms:8.3f} ms  ({n} launches)", file=sys.stderr)

    kveh = head.get("vehicles", 1)
    fps = head["frames"] / head["elapsed"]
    dominant = head["dominant"]
    roof = head.get("roofline_leg") or {"dom_avg_s": head["dom_avg_s"], "launches_per_step": head["launches_per_step"],
                                        "steps": args.steps, "lanes": 1}
    dom_avg_s = roof["dom_avg_s"]
    passes = max(1, (int(np.ceil(np.log2(max(2, C)))) + 7) // 8)
    kb = kernel_bytes(dominant, N, B, C, passes)
    # kernel_bytes is per launch and vehicle; the roofline leg runs one lane, so a launch covers all
    # vehicles of the step.
    dbytes = kb * kveh if kb else None
    achieved = (dbytes / dom_avg_s / 1e9) if (dbytes and dom_avg_s > 0) else None

    shifted = args.workload == "plain_full"
    fb = frame_bytes(N, B, C, shifted)
    traffic, traffic_src = measured_traffic(dominant, args.workload, kveh, kveh * N)
    cpu = None
    if args.cpu_frames > 0:
        cpu = cpu_baseline(args.workload, min(args.cpu_frames, total), meas, poses, dt)

    hs = _summary(head)
    out = {
        "metric": METRIC,
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": hs["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(fps / PUBLISHED_FPS, 3) if args.workload == "plain_full" else None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": WORKLOAD_TEXT[args.workload],
            "form": (f"batched: {kveh} vehicles of this workload per dogm_batch_update step (blocking), each with "
                     "its own Philox instance, split over two lanes (HIP streams)" if batched
                     else "plain (one updateGrid call per frame, blocking)"),
            "particles_per_vehicle": N, "birth_particles_per_vehicle": B, "cells_per_vehicle": C,
            "measurement_in_hbm": True,
            "baseline_source": "reference README: 50 ms/frame at 2e6/2e5/1.44e6 (GTX 980, Nuss et al.)",
        },
        # Step latency percentiles from the latency leg (>= 1000 distinct frames) when it ran, else from
        # the timed steps.
        "p50_ms": (latency["batched" if batched else "plain"]["p50_ms"] if latency else hs["p50_ms"]),
        "p99_ms": (latency["batched" if batched else "plain"]["p99_ms"] if latency else hs["p99_ms"]),
        "latency": latency,
        "plain": _summary(plain),
        "batched": _summary(batched) if batched else None,
        "batched_vs_plain": round(fps / (plain["frames"] / plain["elapsed"]), 3) if batched else None,
        "extraction": batched.get("extraction") if batched else None,
        "measurement_generation": batched.get("generation") if batched else None,
        "batched_uneven": uneven,
        # Short legs of the other configurations and the f32-vs-FP64 error report (config 2 frames).
        "fp64_errors": legs.get("fp64_errors"),
        "config_plain_small": legs.get("plain_small"),
        "config_precision": legs.get("precision"),
        "config_serving": legs.get("serving"),
        "particle_updates_per_s": round((N + B) * fps, 1),
        "frame_algorithmic_bytes": fb,
        "frame_hbm_fraction": round(fb * fps / (HBM_PEAK_GBS * 1e9), 5),
        # Launch choices the tuning table gave each form (dynamic-occupancy-grid-map_amd/tuning_table.json,
        # written by autotune.py on the device; choices never change results).
        "tuning": _tuning_report(N, kveh if batched else 0, uneven),
        "roofline": {
            "kernel": dominant,
            "bound": "hbm",
            "achieved": round(achieved, 2) if achieved else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "valu_issue": valu_issue(dominant, args.workload, kveh, kveh * N, dom_avg_s),
            "avg_launch_us": round(dom_avg_s * 1e6, 3),
            "algorithmic_bytes_per_launch": dbytes,
            "measured_on": (f"{roof['steps']} steps of the same {kveh}-vehicle batch on one lane after the timed "
                            "region (launches do not overlap)" if batched else "the timed region"),
            "avg_launch_us_two_lanes": round(head["dom_avg_s"] * 1e6, 3) if batched else None,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    for d in bufs:
        d.free()
    return 0


if __name__ == "__main__":
    sys.exit(main())
