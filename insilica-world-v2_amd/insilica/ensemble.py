"""Many worlds at once: the reference's `Ensemble` (src/insilica/ensemble.py:1-29).

Each world keeps its own state, stream and captured tick graphs; the ensemble queues every member's ticks before
waiting, so the GPU runs the members' kernels side by side (at level 6 a single world's kernels are latency-bound and
leave most of the 256 CUs idle). A member does exactly what it does alone: results are bit-identical to solo runs.

`advance(hours)` steps every member by the same model time whatever its level (members of different levels take
different ticks per hour) and records, per member, HIP events at the end of each model day on its stream, from which
`day_latencies()` returns each member's wall time per model day.
"""
import numpy as np
import torch

from .world import World


class Ensemble:
    def __init__(self, worlds, batch=False):
        """batch: step the coupled worlds of each level together on their union mesh (insilica.batch.BatchWorld),
        one launch per phase for the whole level, instead of one graph per world. The members end the same either way;
        with batch=True their state lives in the batches (`member_state(i)`)."""
        self.worlds = list(worlds)
        self.units, self.member = self.worlds, [(i, None) for i in range(len(self.worlds))]
        if batch:
            from .batch import BatchWorld
            groups = {}
            for i, w in enumerate(self.worlds):
                groups.setdefault((w.n, w.dt, w.dg.precision, w.tick), []).append(i)
            self.units, self.member = [], [None] * len(self.worlds)
            for idx in groups.values():
                bw = BatchWorld([self.worlds[i] for i in idx])
                for k, i in enumerate(idx):
                    self.member[i] = (len(self.units), k)
                self.units.append(bw)
        self._marks = {i: [] for i in range(len(self.units))}

    def member_state(self, i):
        """(state_arrays, state_hash) of member i, wherever it is stepped."""
        u, k = self.member[i]
        if k is None:
            w = self.units[u]
            return w.state_arrays(), w.state_hash()
        return self.units[u].state_arrays(k), self.units[u].state_hash(k)

    @classmethod
    def build(cls, n, preset, params_list, **kw):
        """One world per entry of params_list, same level and preset (the reference's calibration sweep)."""
        return cls(World(n, preset, params=p, **kw) for p in params_list)

    def step(self, k=1, graph=True):
        """k ticks of every member, interleaved, then wait for all."""
        for _ in range(k):
            for w in self.units:
                w.launch_tick(graph)
        self.sync()
        return self

    def sync(self):
        for w in self.units:
            w.sync()

    def advance(self, hours, graph=True, day_marks=False):
        """Every member by `hours` model hours (a whole number of its ticks), interleaved one tick per member in turn.
        With day_marks, an event is recorded on a member's stream each time it completes a model day."""
        todo = []
        for w in self.units:
            t = hours * w.ticks_per_hour
            if abs(t - round(t)) > 1e-9:
                raise ValueError(f'{hours} h is not a whole number of ticks of a world with dt {w.dt} s')
            todo.append(int(round(t)))
        per_day = [int(round(24 * w.ticks_per_hour)) for w in self.units]
        start = torch.cuda.Event(enable_timing=True)
        if day_marks:
            start.record()
            for w in self.units:
                w._stream.wait_event(start)
        done = [0] * len(self.units)
        while any(d < t for d, t in zip(done, todo)):
            for i, w in enumerate(self.units):
                if done[i] < todo[i]:
                    w.launch_tick(graph)
                    done[i] += 1
                    if day_marks and done[i] % per_day[i] == 0:
                        ev = torch.cuda.Event(enable_timing=True)
                        ev.record(w._stream)
                        self._marks[i].append(ev)
        self.sync()
        self._start = start if day_marks else None
        return self

    def day_latencies(self):
        """Per member, the wall seconds of each model day of the last advance(day_marks=True) (the members of a batch
        share their batch's)."""
        per_unit = []
        for i in range(len(self.units)):
            prev, lat = self._start, []
            for ev in self._marks[i]:
                lat.append(prev.elapsed_time(ev) / 1e3)
                prev = ev
            per_unit.append(lat)
            self._marks[i] = []
        return [per_unit[u] for u, _ in self.member]

    def __iter__(self):
        return iter(self.worlds)

    def __len__(self):
        return len(self.worlds)


def draw_planets(n_worlds, seed=0, levels=(4, 5, 6)):
    """Seeded per-world settings of the batched-ensemble workload: a level from `levels`, planet radius 0.5-1.5 Earth,
    rotation 0.5-2x and gravity 3.7-11 m/s^2 (log-uniform in rotation). The del^4 velocity scale grows with the
    radius (visc4_U = 15 a / a_Earth m/s) so that the grid-scale damping rate, and its stability margin at the level's
    step, stay Earth's."""
    from .constants import OMEGA, R_PLANET
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_worlds):
        lev = int(levels[k % len(levels)])
        r = float(rng.uniform(0.5, 1.5))
        planet = dict(radius=R_PLANET * r, omega=float(OMEGA * np.exp(rng.uniform(np.log(0.5), np.log(2.0)))),
                      grav=float(rng.uniform(3.7, 11.0)))
        out.append((lev, dict(planet=planet, seed=int(rng.integers(1 << 30)), visc4_U=15.0 * r)))
    return out
