"""Throughput benchmark of the explicit 2-D Saint-Venant time step on one MI355X.

Workload (the headline configuration): one 2048 x 2048 catchment, dx = 2 m,
tilted terrain with 30 Gaussian hills and 1 m fBm on a +300 m base, Manning n
from 4 Voronoi land-use classes in [0.02, 0.08], a drifting Gaussian storm
(80 mm/h peak, sigma 400 m, 5 m/s), Heun-2, CFL 0.5, free outflow on the
x = 0 edge, no velocity in cells shallower than 1 mm, velocity components
within 10 m/s (extension flags E1 E2 E4 E7 E8 E9), dry start.

A "step" is one accepted-or-rejected timestep of the whole catchment: both
RK stages (fused stage kernels) plus the device controller.  The run starts
from the dry initial state and is spun up untimed to --spinup simulated
seconds (600 s by default), so that the timed window sees the storm's
established flow rather than the start-up transient whatever K and W are;
then W warmup timesteps are untimed and exactly K timesteps are timed
between device synchronisations.  Inputs are resident in HBM before timing
starts.  The accepted fraction and the wet fraction of the timed window are
reported beside the value.

value = accepted cell-updates per second of the FP32 precision form (FP64
time, controller and mass balance); the FP64 plain method is measured the
same way and reported beside it, with the FP32-vs-FP64 depth, velocity and
mass errors after a common simulated horizon.

roofline: the dominant kernel is the stage kernel.  Algorithmic bytes per
cell-update B = s*(15 + 2m) with m = 2 streamed static fields (bed, n):
152 B in FP64, 76 B in FP32; one stage launch processes the N cells of the
grid, so achieved = (B/2)*N / (mean stage-launch duration from HIP events on
the library's stream).

cpu_baseline: the CPU oracle (a C restatement of the reference's kernel
sequence, FP64, OpenMP) on a bounded sample of the same workload: it
continues from the GPU's FP64 state at the end of the timed window, median
of 5 consecutive windows.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s & HBM GB/s, FP32-batched vs FP64 plain; p50/p99 scenario latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_scenario(n):
    from mfcm_amd import scenarios as S
    sc = S.storm_catchment(seed=1, n=n, dx=2.0, t_end=7200.0)
    sc.flags |= S.E7_Y_ORIENT
    return sc


def device_sync():
    import torch
    torch.cuda.synchronize()


def spin_up(cat, t_sim, chunk=512):
    """Advance a begun catchment, untimed, to at least t_sim simulated
    seconds (whole graph chunks; the time reached is reported)."""
    st = cat.stats()
    while st["time"] < t_sim and not st["done"]:
        st = cat.run(chunk, graph=True)
    return st


def time_precision(sc, prec, steps, warmup, stage_samples, rows=0, spinup_s=0.0,
                   keep_state=False):
    """Spin up untimed to spinup_s simulated seconds, then W warmup and K
    timed timesteps (hipGraph replay; a remainder below the 32-timestep
    chunk is launched eagerly), then stage-kernel launch durations from HIP
    events on the library's stream."""
    from mfcm_amd import mfsv
    from mfcm_amd.autotune import TileTuner
    timings = None
    if not rows:
        tuner = TileTuner()
        rows = tuner.tune(sc, prec, spinup_s=spinup_s)
        timings = {k: round(v, 1) for k, v in tuner.last_timings.items()}
        log(f"  tile rows {rows}: us per timestep {timings}")
    cat = mfsv.Catchment(sc, precision=prec)
    if rows:
        cat.set_tile_rows(rows)
    cat.run_begin()
    sp = spin_up(cat, spinup_s) if spinup_s > 0 else cat.stats()
    if warmup:
        cat.run(warmup, graph=True)
    device_sync()
    s0 = cat.stats()
    t0 = time.perf_counter()
    cat.advance(steps, graph=True)
    device_sync()
    wall = time.perf_counter() - t0
    ev_ms = cat.last_run_ms()
    s1 = cat.stats()
    acc = s1["accepted"] - s0["accepted"]
    rej = s1["rejected"] - s0["rejected"]
    h = cat.download("h")
    wet = float(np.mean(h > 1e-6))
    state = None
    if keep_state:
        state = {"h": h, "u": cat.download("u"), "v": cat.download("v"), "stats": s1}
    # per-launch stage-kernel durations from HIP events (eager launches)
    cat.run(stage_samples, graph=2)
    st_ms, launches = cat.last_stage_ms()
    s2 = cat.stats()
    out = {
        "wall_s": wall, "event_ms": ev_ms, "accepted": acc, "rejected": rej,
        "attempted": acc + rej, "cell_updates": acc * sc.ncells,
        "cell_updates_per_s": acc * sc.ncells / wall,
        "accepted_fraction": acc / max(acc + rej, 1),
        "wet_fraction": wet,
        "spinup_to_s": sp["time"], "spinup_timesteps": sp["accepted"] + sp["rejected"],
        "window_start_s": s0["time"], "window_end_s": s1["time"],
        "stage_ms_mean": st_ms / max(launches, 1), "stage_launches": launches,
        "sim_time_s": s1["time"], "done": s2["done"],
        "device_bytes": cat.device_bytes(),
        "tile_rows": cat.tile_rows(), "tile_timings_us_per_timestep": timings,
    }
    cat.close()
    return out, state


NOISE_M = 1e-7   # bed perturbation of the FP64 sensitivity baseline (m)


def bed_variants(sc):
    """The bed the FP32 form computes with (FP32 above its minimum), and the
    bed plus uniform noise of +-NOISE_M (a twentieth of the FP32 spacing of
    a 30 m relief): the FP64 solution's sensitivity to perturbations below
    FP32 resolution."""
    import copy
    zmin = float(np.nanmin(sc.zb))
    sc_r = copy.copy(sc)
    sc_r.zb = (sc.zb - zmin).astype(np.float32).astype(np.float64) + zmin
    sc_n = copy.copy(sc)
    sc_n.zb = sc.zb + np.random.default_rng(12345).uniform(-NOISE_M, NOISE_M, sc.zb.shape)
    return sc_r, sc_n


def diff_states(ref, other, dx, dy):
    """Depth, velocity and volume differences of `other` from `ref` (each
    (h, u, v, stats))."""
    h0, u0, v0, _ = ref
    h, u, v, _ = other
    d = np.abs(h - h0)
    wet = h0 > 1e-3
    anyw = (h0 > 1e-6) | (h > 1e-6)
    V0, V = float(np.sum(h0)) * dx * dy, float(np.sum(h)) * dx * dy
    return {
        "max_depth_err_m": float(d.max()),
        "p99_depth_err_m": float(np.quantile(d[anyw], 0.99)) if anyw.any() else 0.0,
        "p999_depth_err_m": float(np.quantile(d[anyw], 0.999)) if anyw.any() else 0.0,
        "cells_over_1mm": int((d > 1e-3).sum()),
        "wet_cells_differing": int(np.sum((h0 > 1e-6) != (h > 1e-6))),
        "max_velocity_err_m_s_wet": float(max(np.abs(u - u0)[wet].max(initial=0.0),
                                              np.abs(v - v0)[wet].max(initial=0.0))),
        "rel_volume_err": abs(V - V0) / max(V0, 1e-300),
    }


def mass_residual(st):
    return (st["volume"] - st["volume0"] - st["source_volume"]) / max(abs(st["volume"]), 1e-300)


def precision_study(sc0, horizon, rows=None, variants=("f32", "f64_zb_f32", "f64_bed_noise")):
    """FP32 form vs FP64 plain after `horizon` simulated seconds, beside the
    FP64 solution's own sensitivity: FP64 on the bed the FP32 form computes
    with, and FP64 on the bed plus +-NOISE_M noise.  rows: tile heights per
    precision (None = the library default)."""
    import copy

    from mfcm_amd import mfsv
    sc = copy.copy(sc0)
    sc.t_end = horizon
    sc_r, sc_n = bed_variants(sc)
    runs = {"f64": (mfsv.F64, sc), "f32": (mfsv.F32, sc), "f64_zb_f32": (mfsv.F64, sc_r),
            "f64_bed_noise": (mfsv.F64, sc_n)}
    res, info = {}, {}
    for name in ("f64",) + tuple(variants):
        prec, s_ = runs[name]
        r = (rows or {}).get("f32" if prec == mfsv.F32 else "f64", 0)
        h, u, v, st, wall = run_to_end(s_, prec, r)
        res[name] = (h, u, v, st)
        info[name] = {"wall_s": wall, "accepted": st["accepted"], "rejected": st["rejected"],
                      "cell_updates_per_s": st["accepted"] * sc.ncells / wall,
                      "mass_balance_residual": mass_residual(st)}
        log(f"  {name}: t = {st['time']:.1f} s, {st['accepted']} accepted, "
            f"{st['rejected']} rejected, {wall:.1f} s")
    out = {"horizon_s": horizon, "wet_cells_f64": int(np.sum(res["f64"][0] > 1e-6)),
           "mass_balance_residual_f64": mass_residual(res["f64"][3]),
           "timesteps_f64": res["f64"][3]["accepted"] + res["f64"][3]["rejected"],
           "runs": info}
    if "f32" in res:
        out["f32_vs_f64"] = diff_states(res["f64"], res["f32"], sc.dx, sc.dy)
        out["mass_balance_residual_f32"] = mass_residual(res["f32"][3])
        out["timesteps_f32"] = res["f32"][3]["accepted"] + res["f32"][3]["rejected"]
    sens = {}
    if "f64_zb_f32" in res:
        sens["bed_as_the_f32_form_holds_it"] = diff_states(res["f64"], res["f64_zb_f32"], sc.dx, sc.dy)
    if "f64_bed_noise" in res:
        sens[f"bed_noise_{NOISE_M:g}_m"] = diff_states(res["f64"], res["f64_bed_noise"], sc.dx, sc.dy)
        out["timesteps_f64_bed_noise"] = res["f64_bed_noise"][3]["accepted"] + \
            res["f64_bed_noise"][3]["rejected"]
    out["f64_sensitivity"] = sens
    return out, res


def _batched(scs, prec, rows, steps, warmup, spinup_s=0.0):
    """Attempted timesteps of every catchment advanced together (one
    stage launch per RK stage, hipGraph chunks), after an untimed
    spin-up until every member has reached spinup_s simulated seconds;
    accepted cell-updates per second over the timed part."""
    from mfcm_amd import mfsv
    from mfcm_amd.batch import Batch
    cats = []
    b = Batch(prec)
    for sc in scs:
        c = mfsv.Catchment(sc, precision=prec)
        c.set_tile_rows(rows)
        c.run_begin()
        b.add(c)
        cats.append(c)
    while spinup_s > 0 and any(not st["done"] and st["time"] < spinup_s
                               for st in (c.stats() for c in cats)):
        b.run(512)
    b.run(warmup)
    device_sync()
    a0 = [c.stats()["accepted"] for c in cats]
    t0 = time.perf_counter()
    b.run(steps)
    device_sync()
    wall = time.perf_counter() - t0
    a1 = [c.stats()["accepted"] for c in cats]
    cu = sum((y - x) * sc.ncells for x, y, sc in zip(a0, a1, scs))
    acc_frac = sum(y - x for x, y in zip(a0, a1)) / float(steps * len(cats))
    b.close()
    for c in cats:
        c.close()
    return cu / wall, wall, acc_frac


def ensemble(n, steps, warmup, horizon_members, horizon, spinup_s=300.0):
    """Batched ensemble: n catchments (sides log-uniform in [128, 1024], dx
    U[2, 10] m, rain U[10, 100] mm/h for U[15, 60] min, n U[0.02, 0.06], 1 h
    horizon each) advanced together by the tile-table kernel with
    an autotuned tile height and graph-captured timesteps, in FP32 and in
    FP64, against the same catchments solved one after another by the plain
    FP64 path.  Every leg spins each catchment up untimed from its dry start
    to spinup_s simulated seconds, then times `steps` attempted timesteps per
    catchment after `warmup` untimed ones; cell-updates count accepted
    timesteps only.
    The FP32-batched over FP64-sequential ratio splits into a batching part
    (FP64 batched over FP64 sequential) and a precision part (FP32 batched
    over FP64 batched)."""
    from mfcm_amd import mfsv
    from mfcm_amd import scenarios as S
    from mfcm_amd.autotune import TileTuner, model_rows_batch
    scs = [S.ensemble_member(3000 + k, 128, 1024, 3600.0) for k in range(n)]
    shapes = [(sc.nx, sc.ny) for sc in scs]
    cells = sum(sc.ncells for sc in scs)
    log(f"ensemble: {n} catchments, {cells / 1e6:.2f} M cells; tuning the tile heights")
    rows, timings = {}, {}
    for name, prec in (("f32", mfsv.F32), ("f64", mfsv.F64)):
        tuner = TileTuner()
        rows[name] = tuner.tune_batch(scs, prec, spinup_s=spinup_s)
        timings[name] = {k: round(v, 1) for k, v in tuner.last_timings.items()}
        log(f"  {name} tile rows {rows[name]} (model {model_rows_batch(shapes, prec)}), "
            f"us/timestep {timings[name]}")
    v32, wall32, af32 = _batched(scs, mfsv.F32, rows["f32"], steps, warmup, spinup_s)
    log(f"  FP32 batched: {v32:.4g} cell-updates/s over {wall32:.2f} s, accepted {af32:.3f}")
    v64b, wall64b, af64 = _batched(scs, mfsv.F64, rows["f64"], steps, warmup, spinup_s)
    log(f"  FP64 batched: {v64b:.4g} cell-updates/s over {wall64b:.2f} s, accepted {af64:.3f}")
    wall64, cu64 = 0.0, 0
    for i, sc in enumerate(scs):
        if i % 16 == 0:
            log(f"  FP64 plain sequential: catchment {i}/{n}")
        c = mfsv.Catchment(sc, precision=mfsv.F64)
        c.run_begin()
        spin_up(c, spinup_s)
        c.run(warmup, graph=True)
        device_sync()
        x = c.stats()["accepted"]
        t0 = time.perf_counter()
        c.run(steps, graph=True)
        device_sync()
        wall64 += time.perf_counter() - t0
        cu64 += (c.stats()["accepted"] - x) * sc.ncells
        c.close()
    v64 = cu64 / wall64
    log(f"  FP64 plain sequential: {v64:.4g} cell-updates/s over {wall64:.2f} s")
    B = {"f32": 4 * (15 + 2), "f64": 8 * (15 + 2)}   # m = 1: the bed (n per catchment)
    out = {
        "catchments": n, "cells": cells, "tile_rows": rows,
        "tile_timings_us_per_timestep": timings,
        "timed_timesteps_per_catchment": steps, "spinup_s": spinup_s,
        "f32_batched": {"value": v32, "unit": "cell-updates/s", "wall_s": wall32,
                        "accepted_fraction": af32,
                        "alg_GBs": v32 * B["f32"] / 1e9,
                        "hbm_frac": v32 * B["f32"] / 1e9 / HBM_PEAK_GBS},
        "f64_batched": {"value": v64b, "unit": "cell-updates/s", "wall_s": wall64b,
                        "accepted_fraction": af64,
                        "alg_GBs": v64b * B["f64"] / 1e9,
                        "hbm_frac": v64b * B["f64"] / 1e9 / HBM_PEAK_GBS},
        "f64_plain_sequential": {"value": v64, "unit": "cell-updates/s", "wall_s": wall64},
        "ratio": v32 / v64,
        "ratio_batching_f64": v64b / v64,
        "ratio_precision_batched": v32 / v64b,
    }
    if horizon_members:
        studies = []
        for k in range(min(horizon_members, n)):
            st, _ = precision_study(S.ensemble_member(3000 + k, 128, 1024, horizon), horizon,
                                    variants=("f32", "f64_bed_noise"))
            studies.append(st)
        key = f"bed_noise_{NOISE_M:g}_m"
        out["accuracy"] = {
            "members": horizon_members, "horizon_s": horizon,
            "f32_vs_f64": {k: max(e["f32_vs_f64"][k] for e in studies)
                           for k in ("max_depth_err_m", "p99_depth_err_m",
                                     "max_velocity_err_m_s_wet", "rel_volume_err")},
            "f64_bed_noise_vs_f64": {k: max(e["f64_sensitivity"][key][k] for e in studies)
                                     for k in ("max_depth_err_m", "p99_depth_err_m",
                                               "max_velocity_err_m_s_wet", "rel_volume_err")},
            "max_mass_balance_residual_f32": max(abs(e["mass_balance_residual_f32"]) for e in studies),
            "max_mass_balance_residual_f64": max(abs(e["mass_balance_residual_f64"]) for e in studies),
            "per_member": [{"f32_rel_volume_err": e["f32_vs_f64"]["rel_volume_err"],
                            "f64_noise_rel_volume_err": e["f64_sensitivity"][key]["rel_volume_err"],
                            "f32_max_depth_err_m": e["f32_vs_f64"]["max_depth_err_m"],
                            "f64_noise_max_depth_err_m": e["f64_sensitivity"][key]["max_depth_err_m"]}
                           for e in studies],
        }
    return out


def _serve_once(reqs, budget_gb, label):
    from mfcm_amd import mfsv
    from mfcm_amd.serving import FloodServer, summarize
    srv = FloodServer(mfsv.F32, hbm_budget=budget_gb * 1e9, chunk=32)
    done = srv.serve(reqs, log=log)
    gpu_s = srv.gpu_s
    srv.close()
    s = summarize(done, srv.elapsed)
    s["batched_device_cell_updates_per_s"] = s["cell_updates"] / max(gpu_s, 1e-9)
    s["device_busy_s"] = gpu_s
    s["hbm_budget_gb"] = budget_gb
    s["peak_resident_gb"] = srv.peak_used / 1e9
    s["peak_resident_requests"] = srv.peak_active
    s["requests_held_back_by_budget"] = len(srv.held_back)
    log(f"  {label}: served {s['completed']}: p50 {s['p50_s']:.2f} s, p99 {s['p99_s']:.2f} s, "
        f"{s['cell_updates_per_s']:.4g} cell-updates/s over the stream, "
        f"{s['batched_device_cell_updates_per_s']:.4g} while launching, "
        f"peak {s['peak_resident_gb']:.2f} GB / {s['peak_resident_requests']} requests resident")
    return s, done


def serving(rate, duration, sides, horizons, budget_gb, tight_gb, seq_sample):
    """A Poisson stream of flood-scenario requests (sides log-uniform in
    `sides`, horizons uniform in `horizons` s, the rest drawn as the
    ensemble's) served by continuous batching (FP32 form, inputs prepared on
    host threads while the GPU works): once under the full HBM budget and
    once under a budget small enough that admission holds requests back.
    Beside it the plain FP64 path solves a sample of the same requests one
    after another (inputs prepared beforehand; the timed part is device
    upload plus the solve)."""
    from mfcm_amd import mfsv
    from mfcm_amd.serving import poisson_stream
    reqs = poisson_stream(rate, duration, seed=7, side_lo=sides[0], side_hi=sides[1],
                          horizon_lo=horizons[0], horizon_hi=horizons[1])
    log(f"serving: {len(reqs)} requests over {duration} s at {rate}/s, sides {sides}, "
        f"horizons {horizons} s, budgets {budget_gb} GB and {tight_gb} GB")
    full, done = _serve_once(reqs, budget_gb, f"budget {budget_gb} GB")
    tight, _ = _serve_once(reqs, tight_gb, f"budget {tight_gb} GB")
    # offered load: the cell-updates the stream asks for, per second of arrivals
    offered = full["cell_updates"] / duration
    seq, cu64, n64 = 0.0, 0, 0
    for i, r in enumerate(reqs[:seq_sample]):
        sc = r.scenario()
        t0 = time.perf_counter()
        c = mfsv.Catchment(sc, precision=mfsv.F64)
        c.run_begin()
        st = c.run(graph=True)
        c.close()
        seq += time.perf_counter() - t0
        cu64 += st["accepted"] * sc.ncells
        n64 += 1
    log(f"  sequential FP64 ({n64} requests): {cu64 / max(seq, 1e-9):.4g} cell-updates/s "
        f"over {seq:.1f} s")
    sizes = [(r.result["nx"], r.result["ny"]) for r in done]
    out = {
        "window": f"{duration:g} s of Poisson arrivals at {rate:g}/s ({len(reqs)} requests)",
        "requests": len(reqs), "rate_per_s": rate, "stream_s": duration,
        "sides": list(sides), "horizons_s": list(horizons),
        "precision": "f32 (f64 time, controller and mass)",
        "grid_cells_mean": float(np.mean([a * b for a, b in sizes])) if sizes else 0.0,
        "offered_cell_updates_per_s": offered,
        "offered_over_capacity": offered / max(full["batched_device_cell_updates_per_s"], 1e-9),
        "full_budget": full, "tight_budget": tight,
        "p50_s": full["p50_s"], "p99_s": full["p99_s"],
        "sequential_f64_plain": {"value": cu64 / max(seq, 1e-9), "unit": "cell-updates/s",
                                 "wall_s": seq, "requests": n64},
        "batched_f32_over_sequential_f64": full["batched_device_cell_updates_per_s"] /
        max(cu64 / max(seq, 1e-9), 1e-9),
    }
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "stage_traffic_storm2048.json")


def traffic_bytes(name, rows, size):
    """Measured HBM bytes per stage launch (rocprofv3 FETCH_SIZE and
    WRITE_SIZE, each calibrated on a plain device copy of known size) from
    the committed profile of this workload (its tile height is recorded
    there); None when no profile matches the grid."""
    try:
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)[name]
    except (OSError, KeyError, ValueError):
        return None
    if t.get("size") != size:
        return None
    return t["bytes_per_launch"]


CU_COUNT, ENGINE_CLOCK_HZ = 256, 2.4e9


def valu_issue(name, size, stage_us):
    """The stage kernel against its instruction-issue limit: VALU
    instructions per launch (rocprofv3 SQ_INSTS_VALU in the committed profile
    of this workload) at one wave-instruction per clock per CU, over the
    launch time measured here; None when no profile matches the grid."""
    try:
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)[name]
        v = t["valu_insts_per_launch"]
    except (OSError, KeyError, ValueError):
        return None
    if t.get("size") != size:
        return None
    issue_us = v / (CU_COUNT * ENGINE_CLOCK_HZ) * 1e6
    return {"valu_insts_per_launch": v, "issue_us": issue_us, "frac": issue_us / stage_us,
            "clock_ghz": ENGINE_CLOCK_HZ / 1e9, "source": os.path.relpath(TRAFFIC_FILE, ROOT)}


PROBE = os.path.join(ROOT, "mfcm_amd", "probe_march")


def memory_floor(size, rows32, rows64):
    """The stage kernels' access pattern without their arithmetic
    (mfcm_amd/probe_march, built beside the library): launch times of
    memory-only kernels that move exactly one stage launch's bytes the same
    way, at the tile heights timed here, plus a device copy and a read-only
    stream.  None when the probe is not built."""
    import subprocess
    if not os.path.exists(PROBE):
        return None
    try:
        out = subprocess.run([PROBE, str(size), str(size), str(rows32), str(rows64)],
                             capture_output=True, text=True, timeout=120)
    except (OSError, subprocess.TimeoutExpired) as e:
        log(f"memory-pattern probe failed: {e}")
        return None
    if out.returncode != 0:
        log(f"memory-pattern probe failed: {out.stderr.strip()[-300:]}")
        return None
    vals = {}
    for line in out.stdout.splitlines():
        if "|" not in line:
            continue
        tok = line.split("|", 1)[1].split()
        if len(tok) == 2:
            vals[tok[0]] = float(tok[1])
        elif len(tok) == 5 and tok[1] == "rows" and tok[3] == "us":
            vals[tok[0]] = float(tok[4])
    return vals


def run_to_end(sc, prec, rows=0):
    from mfcm_amd import mfsv
    cat = mfsv.Catchment(sc, precision=prec)
    if rows:
        cat.set_tile_rows(rows)
    cat.run_begin()
    device_sync()
    t0 = time.perf_counter()
    st = cat.run(graph=True)
    device_sync()
    wall = time.perf_counter() - t0
    h, u, v = cat.state()
    cat.close()
    return h, u, v, st, wall


def hillslope_section():
    """The CPU-runnable configuration: 256^2, dx 5 m, tilted plane + 8 hills +
    fBm 0.5 m, dry start, 50 mm/h for 30 min, n 0.03, free outflow at x = 0,
    Heun-2, CFL 0.5, 1 h, physics flags.  FP64 and FP32 on the GPU for the
    whole hour with the FP64 sensitivity baselines (parity with the CPU
    oracle at this size is a GPU test)."""
    from mfcm_amd import scenarios as S
    sc = S.hillslope_rain(seed=0, n=256, dx=5.0, physics=True, t_end=3600.0)
    log("hillslope 256^2, 1 h: FP64, FP32 and the FP64 sensitivity runs on the GPU")
    study, _ = precision_study(sc, 3600.0)
    out = {"grid": "256x256", "dx": 5.0, "horizon_s": 3600.0, "flags": int(sc.flags)}
    out.update(study)
    return out


def breach_section(horizon):
    """Wet/dry front stress case: 4096^2, dx 1 m, flat bed + fBm 0.2 m, a wall
    column at x = 1000 with a 50 m breach, 5 m upstream, dry downstream,
    n 0.025, no rain, run to `horizon` s in FP64 and FP32, and in FP64 on the
    bed plus +-NOISE_M noise (the FP64 solution's sensitivity)."""
    from mfcm_amd import mfsv
    from mfcm_amd import scenarios as S
    from mfcm_amd.autotune import TileTuner
    sc = S.breach_front(seed=2, n=4096, dx=1.0, t_end=horizon)
    out = {"grid": "4096x4096", "dx": 1.0, "horizon_s": horizon, "flags": int(sc.flags)}
    rows = {}
    for name, prec in (("f32", mfsv.F32), ("f64", mfsv.F64)):
        rows[name] = TileTuner().tune(sc, prec, warmup=4, samples=8, repeats=2)
    log(f"breach 4096^2: tile rows {rows}, running {horizon} s")
    study, _ = precision_study(sc, horizon, rows=rows, variants=("f32", "f64_bed_noise"))
    out["tile_rows"] = rows
    out.update(study)
    for name in ("f32", "f64"):
        r = study["runs"][name]
        log(f"  {name}: {r['cell_updates_per_s']:.4g} cell-updates/s, "
            f"{r['accepted']} timesteps, {r['wall_s']:.1f} s")
    return out


def sources_section(size, steps, warmup):
    """The headline catchment with the further source terms on: energy loss
    (per-cell fx, fy), infiltration (per-cell rate and capacity, consumed at
    every stage) and Smagorinsky eddy viscosity.  Algorithmic bytes per
    cell-update: s * (15 + 2m + 4) with m = 5 static fields (bed, n, fx, fy,
    rate) and the capacity read and written in both stages."""
    import copy

    from mfcm_amd import mfsv
    rng = np.random.default_rng(11)
    sc = copy.copy(make_scenario(size))
    sc.energy_loss = (rng.uniform(0.0, 0.02, (size, size)), rng.uniform(0.0, 0.02, (size, size)))
    sc.infiltration = (np.full((size, size), 2e-6), np.full((size, size), 0.05))
    sc.eddy = {"smagorinsky": (0.05, 0.5)}
    out = {"workload": f"storm catchment {size}x{size} + energy loss + infiltration + "
                       f"Smagorinsky eddy viscosity", "flags": int(sc.flags)}
    for name, prec, s_ in (("f32", mfsv.F32, 4), ("f64", mfsv.F64, 8)):
        r, _ = time_precision(sc, prec, steps, warmup, 40)
        B = s_ * (15 + 2 * 5 + 4)
        ach = (B / 2.0) * sc.ncells / (r["stage_ms_mean"] * 1e-3) / 1e9
        out[name] = {"cell_updates_per_s": r["cell_updates_per_s"], "tile_rows": r["tile_rows"],
                     "stage_us_mean": r["stage_ms_mean"] * 1e3,
                     "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                                  "bytes_per_cell_update": B}}
        log(f"  sources {name}: {r['cell_updates_per_s']:.4g} cell-updates/s, "
            f"stage {r['stage_ms_mean'] * 1e3:.1f} us")
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_timesteps(sim, sc, step, t_local, dt, budget_s):
    """Timesteps of the oracle (TemporalScheme.hpp:112-149 and :244-276)
    from outer step `step`, local time t_local and timestep dt, for about
    budget_s seconds; returns the rate and where it stopped."""
    n_acc = n_att = 0
    t0 = time.perf_counter()
    while True:
        new_dt, rep = sim.do_timestep(t_local, dt)
        n_att += 1
        if not rep:
            sim.accept()
            t_local += dt
            n_acc += 1
        dt = new_dt
        el = time.perf_counter() - t0
        if t_local >= sc.step_duration:
            step += 1
            t_local = 0.0
            sim.start_new_step(step * sc.step_duration)
        elif t_local + dt > sc.step_duration:
            dt = sc.step_duration - t_local
        elif t_local + 2 * dt > sc.step_duration:
            dt = (sc.step_duration - t_local) * 0.4
        if (el > budget_s and n_acc >= 1) or n_att >= 100000:
            break
    return n_acc * sc.ncells / el, n_acc, n_att, el, step, t_local, dt


def cpu_baseline(sc, state, budget_s, samples=5):
    """The CPU oracle on a bounded sample of the same workload: it continues
    from the GPU's FP64 state at the end of the timed window (same simulated
    time, outer step, local time and timestep), for `samples` consecutive
    windows of budget_s / samples seconds each; the median window rate is
    the value."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    sim = O.sim_from_scenario(sc)
    st = state["stats"]
    sim.set_state(state["h"], state["u"], state["v"])
    step, t_local, dt = int(st["step"]), float(st["t_local"]), float(st["dt"])
    sim.start_new_step(step * sc.step_duration)
    rates, acc, att, el_tot = [], 0, 0, 0.0
    for _ in range(samples):
        r, a, n, el, step, t_local, dt = _oracle_timesteps(sim, sc, step, t_local, dt,
                                                           budget_s / samples)
        rates.append(r)
        acc += a
        att += n
        el_tot += el
    return {
        "value": float(np.median(rates)),
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": cpu_model(),
        "sample": f"{samples} consecutive windows ({att} timesteps, {acc} accepted, "
                  f"{el_tot:.1f} s of CPU time) of the same {sc.nx}x{sc.ny} workload "
                  f"continuing from the GPU's FP64 state at t = {st['time']:.1f} s; median "
                  f"window rate; FP64 C oracle (reference-semantics CPU restatement, not "
                  f"the reference binary) with OpenMP over rows",
        "window_rates": rates,
    }


def cpu_config0(runs):
    """The CPU-runnable configuration (256^2 hillslope, 1 h, physics flags)
    run in full by the CPU oracle `runs` times; the median run."""
    from mfcm_amd import scenarios as S
    from oracle import oracle as O
    sc = S.hillslope_rain(seed=0, n=256, dx=5.0, physics=True, t_end=3600.0)
    walls, acc = [], None
    for i in range(runs):
        t0 = time.perf_counter()
        sim = O.sim_from_scenario(sc)
        sim.solve()
        walls.append(time.perf_counter() - t0)
        acc = sim.stats()
        log(f"  config 0 on the CPU oracle, run {i + 1}/{runs}: {walls[-1]:.1f} s, {acc}")
    w = float(np.median(walls))
    return {"workload": "hillslope 256x256 dx 5 m, 1 h, physics flags, FP64",
            "runs": runs, "wall_s_median": w, "wall_s": walls,
            "accepted": acc[0], "rejected": acc[1],
            "cell_updates_per_s": acc[0] * sc.ncells / w,
            "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
            "cpu_model": cpu_model()}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000, help="timed timesteps")
    ap.add_argument("--warmup", type=int, default=200, help="untimed timesteps")
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--stage-samples", type=int, default=200)
    ap.add_argument("--spinup", type=float, default=600.0,
                    help="simulated seconds run untimed from the dry start before warmup")
    ap.add_argument("--accuracy-horizon", type=float, default=300.0)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--cpu-config0", type=int, default=0,
                    help="runs of the 256^2 hillslope in full on the CPU oracle (0 = none)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--ensemble", type=int, default=64, help="catchments in the batched ensemble")
    ap.add_argument("--ensemble-steps", type=int, default=256)
    ap.add_argument("--ensemble-accuracy", type=int, default=8,
                    help="members compared FP32 vs FP64 to a common horizon")
    ap.add_argument("--ensemble-horizon", type=float, default=600.0)
    ap.add_argument("--serve-rate", type=float, default=20.0, help="requests per second")
    ap.add_argument("--serve-duration", type=float, default=2.0, help="seconds of arrivals")
    ap.add_argument("--serve-sides", type=str, default="256,2048")
    ap.add_argument("--serve-horizons", type=str, default="1800,10800")
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-tight-gb", type=float, default=1.0,
                    help="a second serving run under this HBM budget (admission binds)")
    ap.add_argument("--serve-seq-sample", type=int, default=6,
                    help="requests solved one after another in FP64 for the comparison")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--rows-f32", type=int, default=0, help="stage tile rows (0 = autotune)")
    ap.add_argument("--rows-f64", type=int, default=0, help="stage tile rows (0 = autotune)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 256^2 hillslope and 4096^2 breach-front sections")
    ap.add_argument("--breach-horizon", type=float, default=600.0)
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="MiB of a plain device copy to run first (counter calibration)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(asked for --gpus {args.gpus}, WORLD_SIZE {world}); the catchment "
              f"would shard into row strips with a 2-row halo per RK stage.", file=sys.stderr)
        sys.exit(2)

    # torch first: its HIP runtime must be the one in the process before
    # libmfsv_hip.so binds libamdhip64.so.7 (two runtimes cannot share a GPU).
    import torch
    torch.cuda.init()
    from mfcm_amd import mfsv
    mfsv.lib()
    if mfsv.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        sys.exit(1)

    if args.copy_probe:
        # A plain copy of known size: the calibration kernel for HBM counters.
        src = torch.ones(args.copy_probe * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: 5 x {args.copy_probe} MiB read + {args.copy_probe} MiB written")
        del src, dst

    log(f"generating scenario {args.size}x{args.size}")
    sc = make_scenario(args.size)
    N = sc.ncells
    res, state64 = {}, None
    for name, prec in (("f32", mfsv.F32), ("f64", mfsv.F64)):
        log(f"timing {name}: spin-up to {args.spinup} s, warmup {args.warmup}, steps {args.steps}")
        rows = args.rows_f32 if name == "f32" else args.rows_f64
        res[name], st_ = time_precision(sc, prec, args.steps, args.warmup, args.stage_samples,
                                        rows, spinup_s=args.spinup, keep_state=name == "f64")
        if name == "f64":
            state64 = st_
        r = res[name]
        log(f"  {name}: {r['cell_updates_per_s']:.4g} cell-updates/s, stage "
            f"{r['stage_ms_mean'] * 1e3:.1f} us, window t = {r['window_start_s']:.1f}.."
            f"{r['window_end_s']:.1f} s, accepted {r['accepted']}/{r['attempted']}, "
            f"wet {r['wet_fraction']:.3f}")
    acc = None
    if not args.no_accuracy:
        log(f"accuracy FP32 vs FP64 over {args.accuracy_horizon} s")
        acc, _ = precision_study(sc, args.accuracy_horizon,
                                 rows={"f32": res["f32"]["tile_rows"], "f64": res["f64"]["tile_rows"]})
    ens = None
    if not args.no_ensemble:
        ens = ensemble(args.ensemble, args.ensemble_steps, 32, args.ensemble_accuracy,
                       args.ensemble_horizon)
    srv = None
    if not args.no_serve:
        srv = serving(args.serve_rate, args.serve_duration,
                      tuple(int(x) for x in args.serve_sides.split(",")),
                      tuple(float(x) for x in args.serve_horizons.split(",")),
                      args.serve_budget_gb, args.serve_tight_gb, args.serve_seq_sample)
    cfgs = None
    if not args.no_configs:
        cfgs = {"hillslope_256": hillslope_section(), "breach_4096": breach_section(args.breach_horizon),
                "further_sources_2048": sources_section(args.size, 400, 50)}
    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(sc, state64, args.cpu_budget)
    cpu0 = None
    if args.cpu_config0:
        log(f"config 0 in full on the CPU oracle, {args.cpu_config0} runs")
        cpu0 = cpu_config0(args.cpu_config0)

    B = {"f64": 8 * (15 + 2 * 2), "f32": 4 * (15 + 2 * 2)}
    log("memory-pattern probe")
    floor = memory_floor(args.size, res["f32"]["tile_rows"], res["f64"]["tile_rows"])
    roof = {}
    for name in ("f32", "f64"):
        r = res[name]
        ach = (B[name] / 2.0) * N / (r["stage_ms_mean"] * 1e-3) / 1e9
        roof[name] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": ach / HBM_PEAK_GBS,
                      "traffic": traffic_bytes(name, r["tile_rows"], args.size),
                      "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT),
                      "kernel": f"mfsv::k_stage<{'float' if name == 'f32' else 'double'}>",
                      "stage_us_mean": r["stage_ms_mean"] * 1e3,
                      "bytes_per_cell_update": B[name],
                      "valu_issue": valu_issue(name, args.size, r["stage_ms_mean"] * 1e3)}
        if floor and f"{name}_s1" in floor and f"{name}_s2" in floor:
            # what this access pattern allows with free arithmetic
            fl_us = 0.5 * (floor[f"{name}_s1"] + floor[f"{name}_s2"])
            fl_ach = (B[name] / 2.0) * N / (fl_us * 1e-6) / 1e9
            roof[name]["memory_floor"] = {
                "stage_pattern_us": [floor[f"{name}_s1"], floor[f"{name}_s2"]],
                "stage_us_mean": fl_us, "achieved": fl_ach, "frac": fl_ach / HBM_PEAK_GBS,
                "kernel_over_floor": r["stage_ms_mean"] * 1e3 / fl_us,
                "device_copy_GBs": floor.get("copy_GBs"), "read_only_GBs": floor.get("read_GBs"),
                "tile_rows": r["tile_rows"], "source": os.path.relpath(PROBE, ROOT)}
    f32, f64 = res["f32"], res["f64"]
    line = {
        "metric": METRIC,
        "value": f32["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f32["wall_s"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 time, controller and mass)",
        "data": "synthetic",
        "config": {"workload": f"storm_catchment: single catchment {args.size}x{args.size} dx=2m, "
                               f"30 hills + fBm 1m on +300m, Voronoi n, drifting storm, "
                               f"heun2, CFL 0.5, flags E1|E2|E4|E7|E8|E9, from a dry start "
                               f"spun up untimed to t = {args.spinup:g} s simulated",
                   "cells": N, "precision_form": "f32",
                   "timed_window_s": [f32["window_start_s"], f32["window_end_s"]],
                   "timed_accepted_fraction": f32["accepted_fraction"],
                   "timed_wet_fraction": f32["wet_fraction"]},
        "roofline": roof["f32"],
        "cpu_baseline": cpu,
        "fp64_plain": {"value": f64["cell_updates_per_s"], "unit": "cell-updates/s",
                       "ms_per_step": f64["wall_s"] * 1e3 / args.steps,
                       "roofline": roof["f64"]},
        "f32_over_f64": f32["cell_updates_per_s"] / f64["cell_updates_per_s"],
        "timesteps": {k: {"accepted": res[k]["accepted"], "rejected": res[k]["rejected"],
                          "accepted_fraction": res[k]["accepted_fraction"],
                          "wet_fraction": res[k]["wet_fraction"],
                          "spinup_to_s": res[k]["spinup_to_s"],
                          "spinup_timesteps": res[k]["spinup_timesteps"],
                          "window_s": [res[k]["window_start_s"], res[k]["window_end_s"]],
                          "tile_rows": res[k]["tile_rows"]} for k in res},
        "accuracy": acc,
        "ensemble": ens,
        "serving": srv,
        "configs": cfgs,
        "cpu_config0": cpu0,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
