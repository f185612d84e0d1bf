"""Continuous-batching flood-scenario server on one GPU.

Requests (a catchment scenario with its horizon) arrive over time.  The
server keeps every admitted scenario's state resident in HBM and advances
all of them together, a chunk of timesteps per batch launch sequence:

* admission: FIFO, as soon as the request has arrived and its device
  footprint fits in the HBM budget next to the scenarios already resident;
* continuous batching: an admitted scenario joins the running batch at the
  next chunk boundary; a finished one leaves it (its fields are read back and
  its memory is freed) so that waiting requests can be admitted;
* every scenario's result is the one it would have alone (the batch runs the
  single-catchment kernel code per tile; see mfcm_amd.batch);
* request inputs (terrain, roughness, forcing) are prepared on host threads
  while the GPU advances the admitted scenarios, so input preparation
  overlaps device work instead of stalling it.

Time-to-result is measured from a request's arrival to the moment its result
is on the host.
"""
from __future__ import annotations

import concurrent.futures as cf
import dataclasses
import math
import time

import numpy as np

from . import scenarios as S
from .autotune import TileTuner, model_rows_batch
from .batch import Batch
from .mfsv import F32, Catchment


@dataclasses.dataclass
class Request:
    rid: int
    seed: int
    arrival: float            # seconds after the stream starts
    side_lo: int = 256
    side_hi: int = 2048
    horizon: float = 3600.0   # simulated seconds
    result: dict | None = None

    def scenario(self):
        sc = S.ensemble_member(self.seed, side_lo=self.side_lo, side_hi=self.side_hi,
                               t_end=self.horizon)
        return sc


def poisson_stream(rate_hz, duration_s, seed=0, side_lo=256, side_hi=2048,
                   horizon_lo=1800.0, horizon_hi=10800.0):
    """Requests with exponential inter-arrival times (a Poisson stream);
    sides log-uniform in [side_lo, side_hi], horizons uniform."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate_hz)
        if t >= duration_s:
            break
        out.append(Request(len(out), int(rng.integers(1 << 30)), t, side_lo, side_hi,
                           float(rng.uniform(horizon_lo, horizon_hi))))
    return out


def footprint_bytes(nx, ny, precision):
    """Device bytes of a heun2 scenario: two states and one accumulator row
    (3 fields each), bed and Manning n (plus, in the FP32 form, the
    low-order depth of both states), plus partial sums."""
    s = 4 if precision == F32 else 8
    fields = 13 if precision == F32 else 11
    return int(fields * nx * ny * s * 1.02) + (1 << 20)


class FloodServer:
    def __init__(self, precision=F32, hbm_budget=200e9, chunk=32, tuner=None, device=0,
                 prep_workers=4):
        self.precision = precision
        self.pool = cf.ThreadPoolExecutor(max_workers=max(1, int(prep_workers)))
        self.budget = float(hbm_budget)
        self.chunk = int(chunk)
        self.tuner = tuner or TileTuner()
        self.device = device
        self.batch = Batch(precision, device)
        self.used = 0.0
        self.peak_used = 0.0      # largest device footprint resident at once
        self.peak_active = 0
        self.held_back = set()    # requests that waited for HBM room
        self.active = {}          # rid -> (request, catchment, bytes)
        self.done = []
        self.refused = []
        self.prepared = {}
        self.gpu_s = 0.0          # wall time inside batch launches
        self.retile = True        # one tile height for the resident set, kept current
        self._members_gen = 0     # bumped by every admission and retirement
        self._tiled_gen = -1

    def _admit(self, req, now):
        fut = self.prepared.get(req.rid)
        if fut is not None and not fut.done():
            return False  # inputs still being prepared (FIFO: later requests wait)
        sc = fut.result() if fut is not None else req.scenario()
        need = footprint_bytes(sc.nx, sc.ny, self.precision)
        if need > self.budget:
            req.result = {"refused": "larger than the HBM budget"}
            self.refused.append(req)
            return True
        if self.used + need > self.budget:
            self.held_back.add(req.rid)   # arrived, ready, but no room yet
            return False
        cat = Catchment(sc, precision=self.precision, device=self.device)
        # the members share one stage launch: the tile height that suits
        # the resident set with the new member (fewer halo rows once the
        # batch fills the chip; results do not depend on it)
        shapes = [(c.nx, c.ny) for _, c, _, _, _ in self.active.values()] + [(sc.nx, sc.ny)]
        cat.set_tile_rows(model_rows_batch(shapes, self.precision))
        cat.run_begin()
        self.batch.add(cat)
        self.used += need
        self.peak_used = max(self.peak_used, self.used)
        self.peak_active = max(self.peak_active, len(self.active) + 1)
        self.prepared.pop(req.rid, None)
        self.active[req.rid] = (req, cat, need, now, sc.ncells)
        self._members_gen += 1
        return True

    def _retire(self, now):
        # the batch drops finished members from its active set at every chunk
        # boundary; when none was dropped there is no control block to read
        if self.batch.size()[1] == len(self.active):
            return
        for rid in list(self.active):
            req, cat, need, t_adm, ncells = self.active[rid]
            st = cat.stats()
            if not st["done"]:
                continue
            h = cat.download("h")
            t_res = now()
            req.result = {
                "arrival": req.arrival, "admitted": t_adm, "finished": t_res,
                "time_to_result": t_res - req.arrival, "nx": cat.nx, "ny": cat.ny,
                "accepted": st["accepted"], "rejected": st["rejected"],
                "cell_updates": st["accepted"] * ncells, "volume": st["volume"],
                "max_depth": float(h.max()), "h": h,
            }
            self.batch.remove(cat)
            cat.close()
            self.used -= need
            del self.active[rid]
            self.done.append(req)
            self._members_gen += 1

    def _retile(self):
        """One tile height for the whole resident set, chosen again whenever
        the set changes: a member admitted next to few others takes short
        tiles (more, smaller waves to fill the chip), which cost it halo rows
        for the rest of its run once the batch has grown.  Results do not
        depend on the tile height; the batch rebuilds its tile table once."""
        if not self.retile or self._tiled_gen == self._members_gen or not self.active:
            return
        self._tiled_gen = self._members_gen
        shapes = [(c.nx, c.ny) for _, c, _, _, _ in self.active.values()]
        rows = model_rows_batch(shapes, self.precision)
        for _, c, _, _, _ in self.active.values():
            if c.tile_rows() != rows:
                c.set_tile_rows(rows)

    def serve(self, requests, wall_limit=None, keep_fields=False, log=None):
        """Run the stream: requests arrive at their `arrival` offsets from
        now; returns the finished requests (with .result)."""
        pending = sorted(requests, key=lambda r: r.arrival)
        self.prepared = {}
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        nxt = 0
        queue = []
        last_log = 0.0
        while nxt < len(pending) or queue or self.active:
            t = now()
            if wall_limit is not None and t > wall_limit:
                break
            while nxt < len(pending) and pending[nxt].arrival <= t:
                queue.append(pending[nxt])
                self.prepared[pending[nxt].rid] = self.pool.submit(pending[nxt].scenario)
                nxt += 1
            while queue and self._admit(queue[0], t):
                queue.pop(0)
            self._retile()
            if self.active:
                tb = time.perf_counter()
                self.batch.run(self.chunk)
                self.gpu_s += time.perf_counter() - tb
                self._retire(now)
            elif queue:
                time.sleep(0.001)  # waiting for the head request's inputs
            elif nxt < len(pending):
                time.sleep(max(0.0, min(0.05, pending[nxt].arrival - now())))
            if log and now() - last_log > 5.0:
                last_log = now()
                log(f"t={last_log:.1f}s done={len(self.done)} active={len(self.active)} "
                    f"queued={len(queue)} HBM={self.used / 1e9:.2f} GB")
        if not keep_fields:
            for r in self.done:
                r.result.pop("h", None)
        self.backlog = len(queue) + (len(pending) - nxt) + len(self.active)
        self.elapsed = now()
        return self.done

    def close(self):
        self.pool.shutdown(wait=True, cancel_futures=True)
        for rid in list(self.active):
            _, cat, _, _, _ = self.active.pop(rid)
            self.batch.remove(cat)
            cat.close()
        self.batch.close()


def summarize(done, elapsed):
    ttr = np.array([r.result["time_to_result"] for r in done]) if done else np.zeros(0)
    cu = sum(r.result["cell_updates"] for r in done)
    return {
        "completed": len(done),
        "p50_s": float(np.percentile(ttr, 50)) if len(ttr) else None,
        "p99_s": float(np.percentile(ttr, 99)) if len(ttr) else None,
        "max_s": float(ttr.max()) if len(ttr) else None,
        "cell_updates": int(cu),
        "cell_updates_per_s": cu / elapsed if elapsed > 0 else 0.0,
        "elapsed_s": elapsed,
    }


def offered_load(requests, dt_model=0.05):
    """Rough cell-updates the stream asks for (horizon / dt per cell)."""
    tot = 0.0
    for r in requests:
        side = math.sqrt(r.side_lo * r.side_hi)
        tot += side * side * r.horizon / dt_model
    return tot
