"""Tile-height autotuner for the stage kernel.

The stage kernel gives each wavefront a strip of 60 interior columns and
``rows`` interior rows (plus a 2-row halo above and below).  Fewer rows mean
more halo work per cell; more rows mean fewer wavefronts, which can leave
SIMDs idle or unevenly loaded (4480 waves on 1024 SIMDs leave a tail).  The
best height depends on the grid shape and the precision, so it is measured:
``TileTuner.tune`` times each candidate on a scratch copy of the catchment
(HIP events around the stage launches) and caches the winner keyed by
(precision, nx, ny).  Results never depend on the choice: every cell's
arithmetic is the same for any tiling.

Without a measurement ``rows_for`` uses a model of the same trade-off.
"""
from __future__ import annotations

import copy
import json
import math
import os

CANDIDATES = (4, 6, 8, 12, 14, 15, 16, 18, 20, 24, 28, 32, 36, 40, 48)
STRIP = 60
SIMDS = 1024          # 256 CUs x 4 SIMDs on MI355X
# Waves per SIMD of the stage kernels at their register counts (F64, F32):
# the single-catchment FP64 kernel of power-of-two spacing is built for 3
# (<= 168 VGPRs), the batch kernel holds 2 FP64 / 4 FP32 waves.
OCCUPANCY = {0: 3, 1: 4}
OCCUPANCY_BATCH = {0: 2, 1: 4}
REPEATS = 3           # interleaved timing rounds per candidate (median)


def model_rows(nx: int, ny: int, precision: int) -> int:
    """Candidate with the lowest modelled time: halo overhead (rows + 2)/rows
    times the wave-count quantisation ceil(w)/w over full occupancy rounds."""
    occ = OCCUPANCY.get(precision, 2)
    strips = math.ceil(nx / STRIP)
    best, best_cost = CANDIDATES[0], float("inf")
    for r in CANDIDATES:
        waves = strips * math.ceil(ny / r)
        rounds = waves / (SIMDS * occ)
        quant = math.ceil(rounds) / rounds if rounds > 0 else 1.0
        cost = (r + 2.0) / r * quant
        if cost < best_cost - 1e-12:
            best, best_cost = r, cost
    return best


def model_rows_batch(shapes, precision: int) -> int:
    """One tile height for a batch of (nx, ny) grids sharing a stage
    launch: the same trade-off over the concatenated tile table, where the
    halo overhead is weighted by each grid's cells and the last strip of a
    grid wastes the rows beyond ny."""
    occ = OCCUPANCY_BATCH.get(precision, 2)
    cells = sum(nx * ny for nx, ny in shapes)
    if cells == 0:
        return CANDIDATES[0]
    best, best_cost = CANDIDATES[0], float("inf")
    for r in CANDIDATES:
        waves = sum(math.ceil(nx / STRIP) * math.ceil(ny / r) for nx, ny in shapes)
        work = sum(math.ceil(nx / STRIP) * STRIP * math.ceil(ny / r) * (r + 2)
                   for nx, ny in shapes)
        rounds = waves / (SIMDS * occ)
        quant = math.ceil(rounds) / rounds if rounds > 0 else 1.0
        cost = work / cells * quant
        if cost < best_cost - 1e-12:
            best, best_cost = r, cost
    return best


class TileTuner:
    def __init__(self, cache_path: str | None = None):
        self.cache_path = cache_path
        self.table: dict[str, int] = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as fh:
                self.table = {k: int(v) for k, v in json.load(fh).items()}

    @staticmethod
    def key(precision, nx, ny):
        return f"{int(precision)}:{int(nx)}x{int(ny)}"

    def rows_for(self, nx, ny, precision) -> int:
        return self.table.get(self.key(precision, nx, ny), model_rows(nx, ny, precision))

    def tune(self, scenario, precision, warmup=20, samples=64, candidates=CANDIDATES,
             repeats=REPEATS, spinup_s=0.0) -> int:
        """Time every candidate on a scratch catchment of the same scenario:
        run untimed to spinup_s simulated seconds (the flow the caller will
        time: a dry start's thin films make faces and friction cheaper than
        an established flow's), then the state after the warmup is restored
        before each timing, the candidates are timed in `repeats` interleaved
        rounds, and each one's median time per timestep decides.  A timing is
        the replay of `samples` timesteps from the timestep graph (the launches
        the caller's run makes), after an untimed chunk at the new height that
        captures that graph."""
        import statistics

        from .mfsv import Catchment
        sc = copy.copy(scenario)
        # the restored state does not rewind the clock: a horizon the tuning
        # cannot reach keeps the run from finishing between candidates
        sc.t_end = 4.0 * max(float(sc.t_end), float(spinup_s))
        cat = Catchment(sc, precision=precision)
        try:
            cat.run_begin()
            st = cat.stats()
            while st["time"] < spinup_s and not st["done"]:
                st = cat.run(512, graph=True)
            cat.run(warmup, graph=True)
            snap = cat.state()
            times = {r: [] for r in candidates}
            for k in range(repeats):
                for r in (candidates if k % 2 == 0 else tuple(reversed(candidates))):
                    cat.set_state(*snap)
                    cat.set_tile_rows(r)
                    cat.run(32, graph=True)
                    cat.run(samples, graph=True)
                    times[r].append(cat.last_run_ms() / samples * 1e3)
            timings = {r: statistics.median(v) for r, v in times.items() if v}
        finally:
            cat.close()
        best = min(timings, key=timings.get) if timings else model_rows(
            scenario.nx, scenario.ny, precision)
        self.table[self.key(precision, scenario.nx, scenario.ny)] = best
        self.last_timings = timings
        return best

    def rows_for_batch(self, shapes, precision) -> int:
        key = f"{int(precision)}:batch{len(shapes)}:{sum(nx * ny for nx, ny in shapes)}"
        return self.table.get(key, model_rows_batch(shapes, precision))

    def tune_batch(self, scenarios, precision, warmup=16, samples=64,
                   candidates=CANDIDATES, repeats=REPEATS, spinup_s=0.0) -> int:
        """Time a shared tile height for a batch: scratch copies of the members
        run untimed until every one has reached spinup_s simulated seconds,
        then `samples` timesteps per candidate (the batch's event-timed run,
        after two untimed ones at the new height), each timing from the
        members' states after the warmup, in `repeats`
        rounds that alternate their order (no candidate always runs first or
        last); the median time per timestep decides."""
        import statistics

        from .batch import Batch
        from .mfsv import Catchment
        # the restored states do not rewind the members' clocks: horizons the
        # tuning cannot reach keep members from finishing (and leaving the
        # batch) between candidates
        scratch = []
        for sc in scenarios:
            sc = copy.copy(sc)
            sc.t_end = 4.0 * max(float(sc.t_end), float(spinup_s))
            scratch.append(sc)
        cats = [Catchment(sc, precision=precision) for sc in scratch]
        b = Batch(precision)
        times = {r: [] for r in candidates}
        try:
            for c in cats:
                c.run_begin()
                b.add(c)
            while spinup_s > 0 and any(not st["done"] and st["time"] < spinup_s
                                       for st in (c.stats() for c in cats)):
                b.run(512)
            b.run(warmup)
            snaps = [c.state() for c in cats]
            for k in range(repeats):
                for r in (candidates if k % 2 == 0 else tuple(reversed(candidates))):
                    for c, sn in zip(cats, snaps):
                        c.set_state(*sn)
                        c.set_tile_rows(r)
                    b.run(2)
                    b.run(samples)
                    times[r].append(b.last_run_ms() / samples * 1e3)
        finally:
            b.close()
            for c in cats:
                c.close()
        timings = {r: statistics.median(v) for r, v in times.items() if v}
        shapes = [(sc.nx, sc.ny) for sc in scenarios]
        best = min(timings, key=timings.get) if timings else model_rows_batch(shapes, precision)
        key = f"{int(precision)}:batch{len(shapes)}:{sum(nx * ny for nx, ny in shapes)}"
        self.table[key] = best
        self.last_timings = timings
        return best

    def save(self):
        if self.cache_path:
            with open(self.cache_path, "w") as fh:
                json.dump(self.table, fh, indent=1, sort_keys=True)
