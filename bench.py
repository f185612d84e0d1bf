"""Throughput benchmark of the explicit 2-D Saint-Venant time step on one MI355X.

Workload (the headline configuration): one 2048 x 2048 catchment, dx = 2 m,
tilted terrain with 30 Gaussian hills and 1 m fBm on a +300 m base, Manning n
from 4 Voronoi land-use classes in [0.02, 0.08], a drifting Gaussian storm
(80 mm/h peak, sigma 400 m, 5 m/s), Heun-2, CFL 0.5, free outflow on the
x = 0 edge, no velocity in cells shallower than 1 mm, velocity components
within 10 m/s (extension flags E1 E2 E4 E7 E8 E9), dry start.

A "step" is one accepted-or-rejected timestep of the whole catchment: both
RK stages (fused stage kernels) plus the device controller.  The run starts
from the dry initial state; W warmup timesteps are untimed, then exactly K
timesteps are timed between device synchronisations.  Inputs are resident
in HBM before timing starts.

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
sequence, FP64, OpenMP) on a bounded sample of the same workload.
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


def time_precision(sc, prec, steps, warmup, stage_samples, rows=0):
    from mfcm_amd import mfsv
    from mfcm_amd.autotune import TileTuner
    timings = None
    if not rows:
        tuner = TileTuner()
        rows = tuner.tune(sc, prec)
        timings = {k: round(v, 1) for k, v in tuner.last_timings.items()}
        log(f"  tile rows {rows}: stage us {timings}")
    cat = mfsv.Catchment(sc, precision=prec)
    if rows:
        cat.set_tile_rows(rows)
    cat.run_begin()
    if warmup:
        cat.run(warmup, graph=True)
    device_sync()
    s0 = cat.stats()
    t0 = time.perf_counter()
    cat.run(steps, graph=True)
    device_sync()
    wall = time.perf_counter() - t0
    ev_ms = cat.last_run_ms()
    s1 = cat.stats()
    acc = s1["accepted"] - s0["accepted"]
    rej = s1["rejected"] - s0["rejected"]
    # per-launch stage-kernel durations from HIP events (eager launches)
    cat.run(stage_samples, graph=2)
    st_ms, launches = cat.last_stage_ms()
    s2 = cat.stats()
    out = {
        "wall_s": wall, "event_ms": ev_ms, "accepted": acc, "rejected": rej,
        "attempted": acc + rej, "cell_updates": acc * sc.ncells,
        "cell_updates_per_s": acc * sc.ncells / wall,
        "stage_ms_mean": st_ms / max(launches, 1), "stage_launches": launches,
        "sim_time_s": s1["time"], "done": s2["done"],
        "device_bytes": cat.device_bytes(),
        "tile_rows": cat.tile_rows(), "tile_timings_us": timings,
    }
    cat.close()
    return out


def accuracy(sc0, horizon):
    """FP32 form vs FP64 plain after `horizon` simulated seconds, beside the
    FP64 solution's own sensitivity: the FP64 path rerun on the bed rounded
    to FP32 precision (the input rounding the FP32 form starts from) and on
    the bed scaled by 1 + 1e-9."""
    import copy

    from mfcm_amd import mfsv
    sc = copy.copy(sc0)
    sc.t_end = horizon
    zmin = float(sc.zb.min())
    sc_r = copy.copy(sc)
    sc_r.zb = (sc.zb - zmin).astype(np.float32).astype(np.float64) + zmin
    sc_p = copy.copy(sc)
    sc_p.zb = sc.zb * (1.0 + 1e-9)
    res = {}
    for name, prec, s in (("f64", mfsv.F64, sc), ("f32", mfsv.F32, sc),
                          ("f64_zb_f32", mfsv.F64, sc_r), ("f64_zb_1e-9", mfsv.F64, sc_p)):
        cat = mfsv.Catchment(s, precision=prec)
        cat.run_begin()
        st = cat.stats()
        while not st["done"]:
            st = cat.run(4000, graph=True)
        log(f"  accuracy {name}: t = {st['time']:.1f} s, {st['accepted']} accepted, "
            f"{st['rejected']} rejected")
        h, u, v = cat.state()
        res[name] = (h, u, v, st)
        cat.close()
    h64, u64, v64, st64 = res["f64"]
    wet = h64 > 1e-3
    V64 = float(np.sum(h64)) * sc.dx * sc.dy

    def diff(name):
        h, u, v, _ = res[name]
        d = np.abs(h - h64)
        V = float(np.sum(h)) * sc.dx * sc.dy
        return {
            "max_depth_err_m": float(d.max()),
            "p99_depth_err_m": float(np.quantile(d, 0.99)),
            "p999_depth_err_m": float(np.quantile(d, 0.999)),
            "mean_depth_err_m": float(d.mean()),
            "cells_over_1mm": int((d > 1e-3).sum()),
            "max_velocity_err_m_s_wet": float(max(np.abs(u - u64)[wet].max(initial=0.0),
                                                  np.abs(v - v64)[wet].max(initial=0.0))),
            "rel_mass_err_vs_f64": abs(V - V64) / max(V64, 1e-300),
        }

    mb = lambda s: (s["volume"] - s["volume0"] - s["source_volume"]) / max(abs(s["volume"]), 1e-300)
    out = {"horizon_s": horizon, "wet_cells_f64": int(wet.sum()), "volume_m3_f64": V64}
    out.update(diff("f32"))
    out["mass_balance_residual_f64"] = mb(st64)
    out["mass_balance_residual_f32"] = mb(res["f32"][3])
    out["timesteps_f64"] = st64["accepted"] + st64["rejected"]
    out["timesteps_f32"] = res["f32"][3]["accepted"] + res["f32"][3]["rejected"]
    out["f64_sensitivity"] = {"bed_rounded_to_f32": diff("f64_zb_f32"),
                              "bed_scaled_1e-9": diff("f64_zb_1e-9")}
    return out


def ensemble(n, steps, warmup, horizon_members, horizon):
    """Batched ensemble: n catchments (sides log-uniform in [128, 1024], dx
    U[2, 10] m, rain U[10, 100] mm/h for U[15, 60] min, n U[0.02, 0.06], 1 h
    horizon each) advanced together in FP32 by the persistent tile-table
    kernel with an autotuned tile height and graph-captured timesteps,
    against the same catchments solved one after another by the plain FP64
    path.  Both sides time `steps` attempted timesteps per catchment after
    `warmup` untimed ones; cell-updates count accepted timesteps only."""
    from mfcm_amd import mfsv
    from mfcm_amd import scenarios as S
    from mfcm_amd.autotune import TileTuner, model_rows_batch
    from mfcm_amd.batch import Batch
    scs = [S.ensemble_member(3000 + k, 128, 1024, 3600.0) for k in range(n)]
    shapes = [(sc.nx, sc.ny) for sc in scs]
    cells = sum(sc.ncells for sc in scs)
    log(f"ensemble: {n} catchments, {cells / 1e6:.2f} M cells; tuning the tile height")
    tuner = TileTuner()
    rows = tuner.tune_batch(scs, mfsv.F32)
    log(f"  tile rows {rows} (model {model_rows_batch(shapes, mfsv.F32)}), "
        f"us/timestep {({k: round(v, 1) for k, v in tuner.last_timings.items()})}")
    cats = []
    b = Batch(mfsv.F32)
    for sc in scs:
        c = mfsv.Catchment(sc, precision=mfsv.F32)
        c.set_tile_rows(rows)
        c.run_begin()
        b.add(c)
        cats.append(c)
    b.run(warmup)
    device_sync()
    a0 = [c.stats()["accepted"] for c in cats]
    t0 = time.perf_counter()
    b.run(steps)
    device_sync()
    wall32 = time.perf_counter() - t0
    a1 = [c.stats()["accepted"] for c in cats]
    cu32 = sum((y - x) * sc.ncells for x, y, sc in zip(a0, a1, scs))
    b.close()
    for c in cats:
        c.close()
    log(f"  FP32 batched: {cu32 / wall32:.4g} cell-updates/s over {wall32:.2f} s")
    wall64, cu64 = 0.0, 0
    for i, sc in enumerate(scs):
        if i % 16 == 0:
            log(f"  FP64 plain sequential: catchment {i}/{n}")
        c = mfsv.Catchment(sc, precision=mfsv.F64)
        c.run_begin()
        c.run(warmup, graph=True)
        device_sync()
        x = c.stats()["accepted"]
        t0 = time.perf_counter()
        c.run(steps, graph=True)
        device_sync()
        wall64 += time.perf_counter() - t0
        cu64 += (c.stats()["accepted"] - x) * sc.ncells
        c.close()
    log(f"  FP64 plain sequential: {cu64 / wall64:.4g} cell-updates/s over {wall64:.2f} s")
    out = {
        "catchments": n, "cells": cells, "tile_rows": rows,
        "tile_timings_us_per_timestep": tuner.last_timings,
        "timed_timesteps_per_catchment": steps,
        "f32_batched": {"value": cu32 / wall32, "unit": "cell-updates/s", "wall_s": wall32},
        "f64_plain_sequential": {"value": cu64 / wall64, "unit": "cell-updates/s",
                                 "wall_s": wall64},
        "ratio": (cu32 / wall32) / (cu64 / wall64),
    }
    if horizon_members:
        errs = []
        for k in range(min(horizon_members, n)):
            errs.append(compare_precisions(S.ensemble_member(3000 + k, 128, 1024, horizon)))
        out["accuracy"] = {
            "members": horizon_members, "horizon_s": horizon,
            "max_depth_err_m": max(e["max_depth_err_m"] for e in errs),
            "max_velocity_err_m_s_wet": max(e["max_velocity_err_m_s_wet"] for e in errs),
            "max_rel_mass_err_vs_f64": max(e["rel_mass_err_vs_f64"] for e in errs),
            "max_mass_balance_residual_f32": max(abs(e["mass_balance_residual_f32"]) for e in errs),
        }
    return out


def compare_precisions(sc):
    """Run a scenario to its end in FP64 and in FP32; depth, velocity and mass
    differences of the FP32 form."""
    from mfcm_amd import mfsv
    res = {}
    for name, prec in (("f64", mfsv.F64), ("f32", mfsv.F32)):
        cat = mfsv.Catchment(sc, precision=prec)
        cat.run_begin()
        st = cat.run(graph=True)
        res[name] = cat.state() + (st,)
        cat.close()
    h64, u64, v64, st64 = res["f64"]
    h32, u32, v32, st32 = res["f32"]
    wet = h64 > 1e-3
    V64 = float(np.sum(h64))
    V32 = float(np.sum(h32))
    mb = lambda s: (s["volume"] - s["volume0"] - s["source_volume"]) / max(abs(s["volume"]), 1e-300)
    return {
        "max_depth_err_m": float(np.abs(h32 - h64).max()),
        "max_velocity_err_m_s_wet": float(max(np.abs(u32 - u64)[wet].max(initial=0.0),
                                              np.abs(v32 - v64)[wet].max(initial=0.0))),
        "rel_mass_err_vs_f64": abs(V32 - V64) / max(V64, 1e-300),
        "mass_balance_residual_f32": mb(st32),
    }


def serving(rate, duration, sides, horizons, budget_gb):
    """A Poisson stream of flood-scenario requests served by continuous
    batching under an HBM budget (FP32 form, inputs prepared on host threads
    while the GPU works), then the same requests solved one after another by
    the plain FP64 path (inputs prepared beforehand; the timed part is device
    upload plus the solve)."""
    from mfcm_amd import mfsv
    from mfcm_amd.serving import FloodServer, poisson_stream, summarize
    reqs = poisson_stream(rate, duration, seed=7, side_lo=sides[0], side_hi=sides[1],
                          horizon_lo=horizons[0], horizon_hi=horizons[1])
    log(f"serving: {len(reqs)} requests over {duration} s at {rate}/s, sides {sides}, "
        f"horizons {horizons} s, budget {budget_gb} GB")
    srv = FloodServer(mfsv.F32, hbm_budget=budget_gb * 1e9, chunk=32)
    done = srv.serve(reqs, log=log)
    gpu_s = srv.gpu_s
    srv.close()
    s = summarize(done, srv.elapsed)
    s["batched_device_cell_updates_per_s"] = s["cell_updates"] / max(gpu_s, 1e-9)
    s["device_busy_s"] = gpu_s
    log(f"  served {s['completed']}: p50 {s['p50_s']:.2f} s, p99 {s['p99_s']:.2f} s, "
        f"{s['cell_updates_per_s']:.4g} cell-updates/s over the stream, "
        f"{s['batched_device_cell_updates_per_s']:.4g} while launching")
    seq, cu64 = 0.0, 0
    for i, r in enumerate(reqs):
        if i % 50 == 0:
            log(f"  sequential FP64: request {i}/{len(reqs)}, {seq:.1f} s of solves")
        sc = r.scenario()
        t0 = time.perf_counter()
        c = mfsv.Catchment(sc, precision=mfsv.F64)
        c.run_begin()
        st = c.run(graph=True)
        c.close()
        seq += time.perf_counter() - t0
        cu64 += st["accepted"] * sc.ncells
    log(f"  sequential FP64: {cu64 / seq:.4g} cell-updates/s over {seq:.1f} s")
    s.update({
        "requests": len(reqs), "rate_per_s": rate, "stream_s": duration,
        "sides": list(sides), "horizons_s": list(horizons), "hbm_budget_gb": budget_gb,
        "precision": "f32 (f64 time, controller and mass)",
        "sequential_f64_plain": {"value": cu64 / seq, "unit": "cell-updates/s", "wall_s": seq},
        "batched_f32_over_sequential_f64": s["batched_device_cell_updates_per_s"] / (cu64 / seq),
    })
    return s


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


def compare(a, b, dx, dy):
    h0, u0, v0, s0 = a[:4]
    h1, u1, v1, s1 = b[:4]
    d = np.abs(h1 - h0)
    wet = h0 > 1e-3
    V0, V1 = float(h0.sum()) * dx * dy, float(h1.sum()) * dx * dy
    return {"max_depth_err_m": float(d.max()), "p99_depth_err_m": float(np.quantile(d, 0.99)),
            "max_velocity_err_m_s_wet": float(max(np.abs(u1 - u0)[wet].max(initial=0.0),
                                                  np.abs(v1 - v0)[wet].max(initial=0.0))),
            "rel_mass_err": abs(V1 - V0) / max(V0, 1e-300),
            "wet_masks_identical": bool(np.array_equal(h0 > 1e-6, h1 > 1e-6))}


def hillslope_section():
    """The CPU-runnable configuration: 256^2, dx 5 m, tilted plane + 8 hills +
    fBm 0.5 m, dry start, 50 mm/h for 30 min, n 0.03, free outflow at x = 0,
    Heun-2, CFL 0.5, 1 h, physics flags.  FP64 and FP32 on the GPU for the
    whole hour (parity with the CPU oracle at this size is a GPU test)."""
    from mfcm_amd import mfsv
    from mfcm_amd import scenarios as S
    sc = S.hillslope_rain(seed=0, n=256, dx=5.0, physics=True, t_end=3600.0)
    log("hillslope 256^2, 1 h: FP64 and FP32 on the GPU")
    r64 = run_to_end(sc, mfsv.F64)
    r32 = run_to_end(sc, mfsv.F32)
    out = {"grid": "256x256", "dx": 5.0, "horizon_s": 3600.0, "flags": int(sc.flags)}
    for name, r in (("f64", r64), ("f32", r32)):
        st = r[3]
        out[name] = {"wall_s": r[4], "accepted": st["accepted"], "rejected": st["rejected"],
                     "cell_updates_per_s": st["accepted"] * sc.ncells / r[4],
                     "mass_balance_residual": (st["volume"] - st["volume0"] - st["source_volume"])
                     / max(abs(st["volume"]), 1e-300)}
    out["f32_vs_f64"] = compare(r64, r32, sc.dx, sc.dy)
    return out


def breach_section(horizon):
    """Wet/dry front stress case: 4096^2, dx 1 m, flat bed + fBm 0.2 m, a wall
    column at x = 1000 with a 50 m breach, 5 m upstream, dry downstream,
    n 0.025, no rain, run to `horizon` s in FP64 and FP32."""
    from mfcm_amd import mfsv
    from mfcm_amd import scenarios as S
    from mfcm_amd.autotune import TileTuner
    sc = S.breach_front(seed=2, n=4096, dx=1.0, t_end=horizon)
    out = {"grid": "4096x4096", "dx": 1.0, "horizon_s": horizon, "flags": int(sc.flags)}
    res = {}
    for name, prec in (("f32", mfsv.F32), ("f64", mfsv.F64)):
        rows = TileTuner().tune(sc, prec, warmup=4, samples=8)
        log(f"breach 4096^2 {name}: tile rows {rows}, running {horizon} s")
        r = run_to_end(sc, prec, rows)
        st = r[3]
        res[name] = r
        out[name] = {"wall_s": r[4], "accepted": st["accepted"], "rejected": st["rejected"],
                     "tile_rows": rows,
                     "cell_updates_per_s": st["accepted"] * sc.ncells / r[4],
                     "mass_balance_residual": (st["volume"] - st["volume0"] - st["source_volume"])
                     / max(abs(st["volume"]), 1e-300)}
        log(f"  {name}: {out[name]['cell_updates_per_s']:.4g} cell-updates/s, "
            f"{st['accepted']} timesteps, {r[4]:.1f} s")
    out["f32_vs_f64"] = compare(res["f64"], res["f32"], sc.dx, sc.dy)
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
        r = time_precision(sc, prec, steps, warmup, 40)
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


def cpu_baseline(sc, budget_s):
    """The CPU oracle on the first timesteps of the same workload."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    sim = O.sim_from_scenario(sc)
    step = 0
    sim.start_new_step(0.0)
    dt, t_local = sc.initial_dt, 0.0
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
        if (el > budget_s and n_acc >= 2) or n_att >= 100000:
            break
        if t_local >= sc.step_duration:  # next outer step (TemporalScheme.hpp:244-276)
            step += 1
            t_local = 0.0
            sim.start_new_step(step * sc.step_duration)
        elif t_local + dt > sc.step_duration:
            dt = sc.step_duration - t_local
        elif t_local + 2 * dt > sc.step_duration:
            dt = (sc.step_duration - t_local) * 0.4
    return {
        "value": n_acc * sc.ncells / el,
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {n_att} timesteps ({n_acc} accepted) of the same {sc.nx}x{sc.ny} "
                  f"workload from the dry start, FP64 C oracle with OpenMP over rows, "
                  f"{el:.1f} s of CPU time",
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000, help="timed timesteps")
    ap.add_argument("--warmup", type=int, default=200, help="untimed timesteps")
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--stage-samples", type=int, default=200)
    ap.add_argument("--accuracy-horizon", type=float, default=300.0)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--ensemble", type=int, default=64, help="catchments in the batched ensemble")
    ap.add_argument("--ensemble-steps", type=int, default=256)
    ap.add_argument("--ensemble-accuracy", type=int, default=8,
                    help="members compared FP32 vs FP64 to a common horizon")
    ap.add_argument("--ensemble-horizon", type=float, default=600.0)
    ap.add_argument("--serve-rate", type=float, default=20.0, help="requests per second")
    ap.add_argument("--serve-duration", type=float, default=10.0, help="seconds of arrivals")
    ap.add_argument("--serve-sides", type=str, default="256,1024")
    ap.add_argument("--serve-horizons", type=str, default="300,1800")
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
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
    res = {}
    for name, prec in (("f32", mfsv.F32), ("f64", mfsv.F64)):
        log(f"timing {name}: warmup {args.warmup}, steps {args.steps}")
        rows = args.rows_f32 if name == "f32" else args.rows_f64
        res[name] = time_precision(sc, prec, args.steps, args.warmup, args.stage_samples, rows)
        log(f"  {name}: {res[name]['cell_updates_per_s']:.4g} cell-updates/s, "
            f"stage {res[name]['stage_ms_mean'] * 1e3:.1f} us")
    acc = None
    if not args.no_accuracy:
        log(f"accuracy FP32 vs FP64 over {args.accuracy_horizon} s")
        acc = accuracy(sc, args.accuracy_horizon)
    ens = None
    if not args.no_ensemble:
        ens = ensemble(args.ensemble, args.ensemble_steps, 32, args.ensemble_accuracy,
                       args.ensemble_horizon)
    srv = None
    if not args.no_serve:
        srv = serving(args.serve_rate, args.serve_duration,
                      tuple(int(x) for x in args.serve_sides.split(",")),
                      tuple(float(x) for x in args.serve_horizons.split(",")),
                      args.serve_budget_gb)
    cfgs = None
    if not args.no_configs:
        cfgs = {"hillslope_256": hillslope_section(), "breach_4096": breach_section(args.breach_horizon),
                "further_sources_2048": sources_section(args.size, 400, 50)}
    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(sc, args.cpu_budget)

    B = {"f64": 8 * (15 + 2 * 2), "f32": 4 * (15 + 2 * 2)}
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
                               f"heun2, CFL 0.5, flags E1|E2|E4|E7|E8|E9, dry start",
                   "cells": N, "precision_form": "f32"},
        "roofline": roof["f32"],
        "cpu_baseline": cpu,
        "fp64_plain": {"value": f64["cell_updates_per_s"], "unit": "cell-updates/s",
                       "ms_per_step": f64["wall_s"] * 1e3 / args.steps,
                       "roofline": roof["f64"]},
        "f32_over_f64": f32["cell_updates_per_s"] / f64["cell_updates_per_s"],
        "timesteps": {k: {"accepted": res[k]["accepted"], "rejected": res[k]["rejected"],
                          "sim_time_s": res[k]["sim_time_s"]} for k in res},
        "accuracy": acc,
        "ensemble": ens,
        "serving": srv,
        "configs": cfgs,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
