"""A serving loop for world requests: continuous batching on one GPU.

Requests (a level, a number of model days, planet settings, a seed) arrive over time. The host-side part of building
each arrived request's world (its planet's mesh, patch tables and seeded initial state: World.prepare) runs at once on
worker threads, off the loop's critical path; the loop admits prepared requests in arrival order as soon as the HBM
budget and the slot count allow, and then only uploads the world and sets it up. An admitted world keeps its state
resident on the device and runs on its own stream with its captured tick graphs, so the work of every admitted world
overlaps on the GPU while new requests join and finished ones leave. Each world is fed a bounded number of model hours
ahead of the GPU (`lead_hours`), so the loop keeps reacting to arrivals instead of queueing all the work at once. A
world's result is bit-identical to the same world run alone (its ticks are exactly its own, whatever runs beside it).

Reported per request: arrival, admission and completion times, the latency per model day (completion - arrival) /
days, and where the time went: queue (arrival to admission, the preparation included), prep (the worker's host
build), build (the loop's device set-up) and compute (admission to completion); in aggregate, model-hours/s over the
run and the percentiles of each. No counterpart in the reference (its server steps one world,
src/insilica/server/app.py); this is the request loop of the batched-serving workload.
"""
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np
import torch

from .world import World


@dataclass
class Request:
    rid: int
    arrival: float                      # s after the start of the run
    level: int
    days: int
    params: dict = field(default_factory=dict)
    admitted: float = None
    finished: float = None
    hash: str = None
    prep_s: float = None                # the worker's host build (World.prepare)
    build_s: float = None               # the loop's device set-up of the world


def draw_requests(n, seed=0, rate=2.0, levels=(5, 6), days=(1, 30)):
    """Seeded Poisson arrivals (mean `rate` per second) of requests at `levels`, each asking for a uniform whole number
    of model days in `days`, with seeded planet settings (as the batched-ensemble workload draws them)."""
    from .ensemble import draw_planets
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    planets = draw_planets(n, seed=seed + 1, levels=levels)
    for k in range(n):
        t += rng.exponential(1.0 / rate)
        lev, p = int(rng.choice(levels)), planets[k][1]
        out.append(Request(k, t, lev, int(rng.integers(days[0], days[1] + 1)), p))
    return out


class Server:
    def __init__(self, hbm_budget=64 << 30, max_slots=64, precision='mixed', lead_hours=2, keep_hashes=False,
                 prep_workers=2):
        self.budget, self.max_slots, self.precision = hbm_budget, max_slots, precision
        self.lead_hours, self.keep_hashes, self.prep_workers = lead_hours, keep_hashes, prep_workers
        self._bytes_per_cell = None

    def _need(self, level):
        """HBM a world of this level holds (tables, state, scratch), calibrated from the first world built."""
        N = 10 * 4 ** level + 2
        return int((self._bytes_per_cell or 4000.0) * N)

    @staticmethod
    def _prepare(r):
        t = time.perf_counter()
        prep = World.prepare(r.level, 'coupled_stack', dict(r.params))
        return prep, time.perf_counter() - t

    def _build(self, r, prepared=None):
        before = torch.cuda.memory_allocated()
        w = World(r.level, 'coupled_stack', precision=self.precision, params=dict(r.params), prepared=prepared)
        if self._bytes_per_cell is None:
            self._bytes_per_cell = 1.1 * (torch.cuda.memory_allocated() - before) / w.grid.N
        return w

    def run(self, requests, clock=time.perf_counter, log=None):
        """Serve `requests` (arrival times relative to now); returns them with their times filled in, and stats.
        log(text), when given, is called as each request is admitted and as it finishes."""
        pending = sorted(requests, key=lambda r: r.arrival)
        active, used, hours_done, preps = [], 0, 0.0, {}
        t0 = clock()
        now = lambda: clock() - t0
        with ThreadPoolExecutor(max(1, self.prep_workers)) as pool:
            while pending or active:
                # every arrived request's host build starts at once, on the workers
                for r in pending:
                    if r.arrival > now():
                        break
                    if r.rid not in preps:
                        preps[r.rid] = pool.submit(self._prepare, r)
                # admit prepared requests in arrival order, while the budget and the slots allow
                while pending and pending[0].rid in preps and preps[pending[0].rid].done() and \
                        len(active) < self.max_slots:
                    r = pending[0]
                    need = self._need(r.level)
                    if active and used + need > self.budget:
                        break
                    pending.pop(0)
                    prep, r.prep_s = preps.pop(r.rid).result()
                    tb = now()
                    w = self._build(r, prep)
                    r.admitted = now()
                    r.build_s = r.admitted - tb
                    active.append(dict(r=r, w=w, need=need, left=int(round(r.days * 24 * w.ticks_per_hour)),
                                       per_hour=int(round(w.ticks_per_hour)), marks=[]))
                    used += need
                    if log:
                        log(f'{r.admitted:8.2f} s  admitted request {r.rid} (level {r.level}, {r.days} days; prep '
                            f'{r.prep_s:.2f} s, build {r.build_s:.2f} s), {len(active)} active')
                if not active:
                    wait = 0.001 if pending and pending[0].arrival <= now() else (
                        pending[0].arrival - now() if pending else 0.0)
                    time.sleep(max(0.0, min(0.01, wait)))
                    continue
                # feed every world up to lead_hours model hours ahead of the GPU
                for a in active:
                    a['marks'] = [e for e in a['marks'] if not e.query()]
                    while a['left'] > 0 and len(a['marks']) < self.lead_hours:
                        k = min(a['per_hour'], a['left'])
                        for _ in range(k):
                            a['w'].launch_tick()
                        a['left'] -= k
                        ev = torch.cuda.Event()
                        ev.record(a['w']._stream)
                        a['marks'].append(ev)
                # retire the worlds whose work has all completed
                still = []
                for a in active:
                    if a['left'] == 0 and all(e.query() for e in a['marks']):
                        r = a['r']
                        r.finished = now()
                        hours_done += 24 * r.days
                        if self.keep_hashes:
                            a['w'].sync()
                            r.hash = a['w'].state_hash()
                        used -= a['need']
                        if log:
                            log(f'{r.finished:8.2f} s  finished request {r.rid} '
                                f'({(r.finished - r.arrival) / r.days:.3f} s per model day)')
                    else:
                        still.append(a)
                active = still
                if active:
                    time.sleep(0.0005)
        makespan = now()
        lat = np.array([(r.finished - r.arrival) / r.days for r in requests])
        pct = lambda x: dict(p50=float(np.percentile(x, 50)), p99=float(np.percentile(x, 99)))
        return requests, dict(model_hours=hours_done, seconds=makespan, model_hours_per_s=hours_done / makespan,
                              day_p50_s=float(np.percentile(lat, 50)), day_p99_s=float(np.percentile(lat, 99)),
                              queue_s=pct([r.admitted - r.arrival for r in requests]),
                              prep_s=pct([r.prep_s for r in requests]), build_s=pct([r.build_s for r in requests]),
                              compute_s=pct([r.finished - r.admitted for r in requests]))
