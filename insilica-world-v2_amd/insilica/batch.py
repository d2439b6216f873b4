"""Batched worlds: many coupled worlds of one level stepped together, each launch covering all of them.

At level 6 one world's kernels are latency-bound (a kernel boundary costs ~6 us on MI355X) and many small worlds on
separate streams are bound by the host's graph submission (~3 us per kernel node). A batch puts the worlds of one
level on one mesh -- the disjoint union of their (planet-scaled) meshes, world k's cells, edges and vertices at an
offset -- with their states side by side in the same layer-major tensors. The per-world values that differ between
planets (gravity, the del^4 coefficient, the fast sub-step counts and lengths) go to a small device table read by the
kernels (isw_batch); everything else is shared. A batch tick is one launch per phase for all the worlds; every world
steps exactly as it does alone, bit for bit (the kernels gather only within a world, in each world's own order).

No counterpart in the reference, whose Ensemble runs one CUDA graph per world on its own stream
(src/insilica/ensemble.py:17-23).
"""
import ctypes
import hashlib
import math

import numpy as np
import torch

from . import _hip
from .constants import CP_DRY, KAPPA, R_DRY
from .gpu import DeviceGrid
from .grid import d2_stencil
from .layers import Layers
from .transport import Transport

CELL_TABLES = ('eI', 'eJ', 'nbr', 'vCell')          # int tables whose entries are cells
EDGE_TABLES = ('cellEdge', 'cellEdgeAsc', 'vEdge', 'eoe')
VERT_TABLES = ('ev1', 'ev2')
PLAIN = ('cellSign', 'vEdgeSign', 'vSlot', 'le', 'de', 'areaCell', 'areaV', 'fV', 'kite', 'w')


class UnionGrid:
    """The disjoint union of grids (as DeviceGrid reads them): tables offset per grid, pads (-1) kept."""

    def __init__(self, grids):
        self.parts = grids
        self.n = grids[0].n
        self.N = sum(g.N for g in grids)
        self.E = sum(g.E for g in grids)
        self.NV = sum(g.NV for g in grids)
        self.cell0 = np.cumsum([0] + [g.N for g in grids])
        self.edge0 = np.cumsum([0] + [g.E for g in grids])
        self.vert0 = np.cumsum([0] + [g.NV for g in grids])
        eoe0 = np.cumsum([0] + [len(g.eoe) for g in grids])

        def cat(name, off=None):
            out = []
            for k, g in enumerate(grids):
                a = np.asarray(getattr(g, name))
                if off is not None:
                    a = np.where(a >= 0, a.astype(np.int64) + off[k], -1)
                out.append(a)
            return np.concatenate(out, axis=0)

        for k in CELL_TABLES:
            setattr(self, k, cat(k, self.cell0).astype(np.int32))
        for k in EDGE_TABLES:
            setattr(self, k, cat(k, self.edge0).astype(np.int32))
        for k in VERT_TABLES:
            setattr(self, k, cat(k, self.vert0).astype(np.int32))
        for k in PLAIN:
            setattr(self, k, cat(k))
        self.eoeStart = np.concatenate([np.asarray(g.eoeStart[:-1], dtype=np.int64) + eoe0[k]
                                        for k, g in enumerate(grids)] + [[eoe0[-1]]]).astype(np.int32)
        idx, wt = [], []
        for k, g in enumerate(grids):
            i, w = d2_stencil(g)
            idx.append(i.astype(np.int64) + self.cell0[k])
            wt.append(w)
        self.d2 = (np.concatenate(idx).astype(np.int32), np.concatenate(wt))
        self.world_cell = np.repeat(np.arange(len(grids)), [g.N for g in grids]).astype(np.int32)
        self.world_edge = np.repeat(np.arange(len(grids)), [g.E for g in grids]).astype(np.int32)


def _substeps(nfast, dt):
    """Per stage c = 1/3, 1/2, 1: the fast sub-steps ceil(c nfast), their length c dt / n and 1 / n, computed as the
    reference's Layers._stage does (layers.py:276-277)."""
    n, tau, inv = [], [], []
    for c in (1 / 3, 1 / 2, 1.0):
        k = max(1, math.ceil(c * nfast))
        n.append(k)
        tau.append(c * dt / k)
        inv.append(1.0 / k)
    return n, tau, inv


class BatchWorld:
    """Coupled worlds (`World(level, 'coupled_stack', ...)` objects of one level and one precision) stepped together
    on their union mesh, from their current states. `tick`, `step`, `launch_tick` and `ticks_per_hour` as World's;
    `state_hash(k)` and `state_arrays(k)` give world k's state as World's own methods would."""

    def __init__(self, worlds):
        worlds = list(worlds)
        assert worlds and all(w.air and w.sea is not None for w in worlds), 'batches hold coupled air-ocean worlds'
        w0 = worlds[0]
        for w in worlds:
            assert w.n == w0.n and w.dt == w0.dt and w.dg.precision == w0.dg.precision, 'one level, step, precision'
            assert w.sea.cfg['dt'] == w0.sea.cfg['dt'] and w.land is not None and w.tick == w0.tick
        self.worlds, self.n, self.dt, self.L = worlds, w0.n, w0.dt, w0.L
        self.ocean_every = w0.ocean_every
        self.grid = ug = UnionGrid([w.grid for w in worlds])
        self.dg = dg = DeviceGrid(ug, w0.dg.precision)
        dev = dg.dev
        i32 = lambda a: torch.as_tensor(np.asarray(a, dtype=np.int32), device=dev)
        f64 = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev)
        cat1 = lambda ts: torch.cat([t.reshape(-1) for t in ts]).contiguous()
        cat2 = lambda ts: torch.cat(list(ts), dim=1).contiguous()
        wc, we = i32(ug.world_cell), i32(ug.world_edge)

        def table(nfasts, dt, nu4s):
            rows = [_substeps(nf, dt) for nf in nfasts]
            return dict(nworlds=len(worlds), wc=wc, we=we, grav=f64([w.grav for w in worlds]), nu4=f64(nu4s),
                        nsub=i32([r[0] for r in rows]), tau=f64([r[1] for r in rows]), inv_n=f64([r[2] for r in rows]),
                        nmax=[max(r[0][k] for r in rows) for k in range(3)])

        c, s0 = w0.core, w0.sea
        self.core = Layers(dg, 3, b=cat1(w.core.b for w in worlds), grav=w0.grav, gprime=[0.0, 0.0], drag=c.drag,
                           stress=c.stress, nu4=c.nu4, floor=c.floor, div_damp=c.div_damp,
                           air=dict(cp=CP_DRY, kappa=KAPPA, rd=R_DRY), vmix=c.vmix, thin_wind=c.thin_wind,
                           batch=table([w.nfast for w in worlds], self.dt, [w.core.nu4 for w in worlds]))
        self.core.Gc.copy_(cat1(w.core.Gc for w in worlds))
        sc = s0.core
        self.ocean = Layers(dg, s0.L, b=cat1(w.sea.core.b for w in worlds), grav=w0.grav, gprime=sc.gprime,
                            stress=sc.stress, nu4=sc.nu4, floor=sc.floor, vmix=sc.vmix, vmix_cd=sc.vmix_cd,
                            thin=sc.thin, wall=np.concatenate([w.sea.wall_host for w in worlds]),
                            batch=table([w.sea.nfast for w in worlds], s0.cfg['dt'], [w.sea.core.nu4 for w in worlds]))
        self.ocean.Gc.copy_(cat1(w.sea.core.Gc for w in worlds))
        self.tr = Transport(dg, layers=3, beta=w0.tr.beta, tracers=5)
        self.otr = Transport(dg, layers=s0.L, tracers=2)
        # state, side by side
        self.h, self.u, self.M = (cat2(getattr(w, k) for w in worlds) for k in ('h', 'u', 'M'))
        self.rows = cat2(w.rows for w in worlds)
        self.theta = self.rows[0:3]
        self.oh, self.ou, self.oM = (cat2(getattr(w.sea, k) for w in worlds) for k in ('h', 'u', 'M'))
        self.orows = cat2(w.sea.rows for w in worlds)
        self.oTg = cat2(w.sea.Tg for w in worlds)
        self.cap = cat1(w.sea.cap for w in worlds)
        self.tau_acc = cat1(w.sea.tau_acc for w in worlds)
        self.diss = cat1(w.sea.diss for w in worlds)
        self.land = cat1(w.sea.land for w in worlds)
        self.soil, self.soil_dz = s0.soil, s0.soil_dz
        self.ocfg, self.c_land = s0.cfg, s0.c_land
        from .ocean import ocean_struct
        self.ostruct = ocean_struct(s0.cfg, w0.grav)
        self.surface_drag = w0.surface_drag
        z = dg.zeros
        self._Fm, self._M1 = z(3, dg.E), z(3, dg.N)
        self._oFm, self._oM1 = z(s0.L, dg.E), z(s0.L, dg.N)
        self.tick = w0.tick
        self._tick0 = self.tick
        self._graphs = {}
        self._stream = torch.cuda.Stream()
        self._stream.wait_stream(torch.cuda.current_stream())

    @property
    def ticks_per_hour(self):
        return 3600.0 / self.dt

    def _kind(self, tick):
        return 0 if (tick + 1) % self.ocean_every else 1

    def _tick(self, kind):
        dg, st = self.dg, _hip.stream
        self.core.step(self.h, self.u, self.dt, 1, self._Fm, theta=self.theta)
        self.tr.mass_after(self.M, self._Fm, self._M1)
        self.tr.step_rows(self.M, self._M1, self._Fm, self.rows, order=3)
        self.M.copy_(self._M1)
        c = self.ocfg
        _hip.call('isw_couple_stress', dg.ref, 3, _hip.ptr(self.h), _hip.ptr(self.u), _hip.ptr(self.oh),
                  _hip.ptr(self.ou), float(self.surface_drag), float(c['rho']), float(self.dt), _hip.ptr(self.diss),
                  _hip.ptr(self.tau_acc), _hip.ptr(self.ocean.wall), float(c['land_drag']), None, st())
        self.core.vmix_heat(self.h, self.diss, self.theta)
        _hip.call('isw_land_column', dg.ref, len(self.soil['dz']), _hip.ptr(self.land), _hip.ptr(self.orows),
                  _hip.ptr(self.cap), _hip.ptr(self.oTg), _hip.ptr(self.soil_dz), float(self.soil['c']),
                  float(self.soil['k']), float(self.dt), st())
        if kind:
            # the ocean step's convective adjustment first (Ocean.step); its buoyancy comparisons take the first
            # world's gravity, which orders every world's layers as its own (g > 0 scales both sides alike)
            L = self.ocean.L
            _hip.call('isw_ocn_overturn', dg.N, L, ctypes.byref(self.ostruct), _hip.ptr(self.oh),
                      _hip.ptr(self.orows[:L]), _hip.ptr(self.orows[L:]), _hip.ptr(self.land), st())
            self.ocean.step(self.oh, self.ou, c['dt'], 1, self._oFm)
            self.otr.mass_after(self.oM, self._oFm, self._oM1)
            self.otr.step_rows(self.oM, self._oM1, self._oFm, self.orows, order=3)
            self.oM.copy_(self._oM1)
            _hip.call('isw_ocn_cap', dg.ref, _hip.ptr(self.oh), c['rho'] * c['cw'], _hip.ptr(self.cap),
                      _hip.ptr(self.land), float(self.c_land), st())

    def launch_tick(self, graph=True):
        kind = self._kind(self.tick)
        gr = self._graphs.get(kind)
        with torch.cuda.stream(self._stream):
            if graph and gr is None and self.tick > self._tick0:
                gr = torch.cuda.CUDAGraph()
                with torch.cuda.graph(gr):
                    self._tick(kind)
                self._graphs[kind] = gr
            if graph and gr is not None:
                gr.replay()
            else:
                self._tick(kind)
        self.tick += 1

    def step(self, k=1, graph=True):
        for _ in range(k):
            self.launch_tick(graph)
        self.sync()
        return self

    def sync(self):
        self._stream.synchronize()
        torch.cuda.current_stream().wait_stream(self._stream)

    def state_arrays(self, k):
        """World k's state as World.state_arrays() gives it (copies of its slices of the union)."""
        ug = self.grid
        c0, c1, e0, e1 = ug.cell0[k], ug.cell0[k + 1], ug.edge0[k], ug.edge0[k + 1]
        cs = lambda t: t[:, c0:c1].contiguous()
        es = lambda t: t[:, e0:e1].contiguous()
        return {'h': cs(self.h), 'u': es(self.u), 'M': cs(self.M), 'rows': cs(self.rows), 'sea.h': cs(self.oh),
                'sea.u': es(self.ou), 'sea.M': cs(self.oM), 'sea.rows': cs(self.orows),
                'sea.cap': self.cap[c0:c1].contiguous(), 'sea.Tg': cs(self.oTg)}

    def state_hash(self, k):
        h = hashlib.sha1()
        for v in self.state_arrays(k).values():
            h.update(v.detach().cpu().numpy().tobytes())
        h.update(str(self.tick).encode())
        return h.hexdigest()

    def __len__(self):
        return len(self.worlds)
