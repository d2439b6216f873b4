"""Device plumbing: a Grid's tables as torch tensors on the HIP device plus the C-ABI struct the kernels read, and the
deterministic sums. Same surface as the reference's gpu.py (src/insilica/gpu.py:16-126): `PRECISION`,
`set_precision`, `DeviceGrid`, `d2_stencil`, `dsum`; `csum` adds the compensated budget sum.

Precision (INSILICA_PRECISION or set_precision, read when a DeviceGrid is built): 'double' is the plain FP64 form,
'mixed' the mixed form (see include/insilica_hip.h). The reference defaults to 'mixed'; so does this package.
"""
import ctypes
import os

import numpy as np
import torch

from . import _hip
from .grid import d2_stencil  # noqa: F401  (re-exported, as the reference's gpu.d2_stencil)

PRECISION = os.environ.get('INSILICA_PRECISION', 'mixed')


def set_precision(p):
    global PRECISION
    assert p in ('mixed', 'double')
    PRECISION = p


def device():
    """The HIP device. Raises if there is none: this package has no CPU path."""
    if not torch.cuda.is_available():
        raise _hip.HipError('no HIP device: the insilica kernels run on the GPU only')
    return torch.device('cuda', torch.cuda.current_device())


PATCH_CELLS = 64
PATCH_PAIRS = True          # the air's fast sub-steps two per launch where the patches' LDS allows


def fast_patches(grid, pcells=64):
    """The air fast sub-step's patches (isw_grid's ph_* and pe_* tables): patch b holds cells [b pcells, (b + 1)
    pcells), its edges are those whose first cell is in it, its halo the second cells of those edges outside it. Each
    edge's two cells get LDS slots (in the patch: their offset; in the halo: pcells + their index in the sorted halo).
    Returns None for a mesh with an edge lacking a cell. Every invariant the kernel relies on is checked here."""
    N, E = int(grid.N), int(grid.E)
    eI, eJ = np.asarray(grid.eI, np.int64), np.asarray(grid.eJ, np.int64)
    if E == 0 or eI.min() < 0 or eJ.min() < 0 or eI.max() >= N or eJ.max() >= N:
        return None
    nb = (N + pcells - 1) // pcells
    bI, bJ = eI // pcells, eJ // pcells
    order = np.lexsort((np.arange(E), bI))
    pe_off = np.searchsorted(bI[order], np.arange(nb + 1)).astype(np.int64)
    ob, oJ = bI[order], eJ[order]
    out = bJ[order] != ob
    keys = np.unique(ob[out] * N + oJ[out])                          # (patch, halo cell), sorted by patch then cell
    kb = keys // N
    ph_off = np.searchsorted(kb, np.arange(nb + 1)).astype(np.int64)
    ph_cell = keys % N
    sI = eI[order] - ob * pcells
    sJ = np.where(out, 0, oJ - ob * pcells)
    hidx = np.searchsorted(keys, ob * N + oJ)
    sJ = np.where(out, pcells + hidx - ph_off[ob], sJ)
    pslots = int(pcells + np.max(np.diff(ph_off)))
    # the kernel's assumptions: every edge once; its first cell in its patch; its slots in range and naming its cells
    assert np.array_equal(np.sort(order), np.arange(E))
    nown = np.minimum(pcells, N - ob * pcells)
    assert np.all((sI >= 0) & (sI < nown))
    assert np.all(np.where(out, (sJ >= pcells) & (sJ < pslots), (sJ >= 0) & (sJ < nown)))
    assert np.all(np.where(out, ph_cell[np.clip(hidx, 0, len(ph_cell) - 1)] == oJ, ob * pcells + sJ == oJ))
    assert pslots < 1 << 15 and np.all(np.diff(ph_off) >= 0) and ph_off[-1] == len(ph_cell)
    return dict(npatch=nb, pcells=pcells, pslots=pslots, ph_off=ph_off, ph_cell=ph_cell, pe_off=pe_off,
                pe_edge=order, pe_slot=sI | (sJ << 16))


def fast_patches2(grid, pt, max_owned=_hip.ISW_PAIR_THREADS):
    """Two air fast sub-steps per launch on the same patches (isw_grid's ph2_*, px_*, pc_* tables): the cells of the
    first sub-step are the patch, its halo (pt's) and the halo's own halo, the cells of the edges of patch and halo;
    the winds of the first sub-step are needed on every edge of patch and halo (the patch's own edges first, in pt's
    order, then the others); the second sub-step's cells are patch and halo again. Cell slots: the patch at its
    offsets, the halo from pcells, the second halo after it; edge slots: the patch's edges 0 .. ne - 1, the others
    from ne. pc_eslot: each patch-and-halo cell's six edge slots, two per int32 (0xffff for a missing edge). None
    when a patch owns more than max_owned edges or its patch and halo hold more than max_owned cells (one thread per
    owned edge and per patch-and-halo cell in the kernel)."""
    N, PC, nb = int(grid.N), pt['pcells'], pt['npatch']
    eI, eJ = np.asarray(grid.eI, np.int64), np.asarray(grid.eJ, np.int64)
    ce = np.asarray(grid.cellEdge, np.int64)
    ph_off, ph_cell, pe_off, pe_edge = pt['ph_off'], pt['ph_cell'], pt['pe_off'], pt['pe_edge']
    if np.max(np.diff(pe_off)) > max_owned or pt['pslots'] > max_owned:
        return None
    h2_off, h2_cell, x_off, x_edge, x_slot, c_off, c_slot = [0], [], [0], [], [], [0], []
    ps2 = pe2 = 0
    for b in range(nb):
        c0 = b * PC
        own = np.arange(c0, min(c0 + PC, N))
        h1 = ph_cell[ph_off[b]:ph_off[b + 1]]
        c1 = np.concatenate([own, h1])
        ep = pe_edge[pe_off[b]:pe_off[b + 1]]
        ec = np.unique(ce[c1][ce[c1] >= 0])
        ex = np.setdiff1d(ec, ep, assume_unique=True)
        assert np.all(np.isin(ep, ec))
        ca = np.unique(np.concatenate([eI[ec], eJ[ec]]))
        h2 = np.setdiff1d(ca, c1)
        # cell slots in the first sub-step's space
        cells = np.concatenate([own, h1, h2])
        slots = np.concatenate([own - c0, PC + np.arange(len(h1)), PC + len(h1) + np.arange(len(h2))])
        srt = np.argsort(cells)
        cs, ss = cells[srt], slots[srt]
        slot_of = lambda c: ss[np.searchsorted(cs, c)]
        assert np.all(cs[np.searchsorted(cs, eI[ex])] == eI[ex]) and np.all(cs[np.searchsorted(cs, eJ[ex])] == eJ[ex])
        xs = slot_of(eI[ex]) | (slot_of(eJ[ex]) << 16)
        # edge slots: the patch's edges in order, then the others
        edges = np.concatenate([ep, ex])
        esrt = np.argsort(edges)
        es_e, es_s = edges[esrt], np.arange(len(edges))[esrt]
        cc = ce[c1]
        k = np.clip(np.searchsorted(es_e, np.where(cc >= 0, cc, 0)), 0, len(es_e) - 1)
        assert np.all((cc < 0) | (es_e[k] == cc))
        sl6 = np.where(cc >= 0, es_s[k], 0xffff)
        c_slot.append((sl6[:, 0::2] | (sl6[:, 1::2] << 16)).reshape(-1))
        h2_cell.append(h2); x_edge.append(ex); x_slot.append(xs)
        h2_off.append(h2_off[-1] + len(h2)); x_off.append(x_off[-1] + len(ex)); c_off.append(c_off[-1] + len(c1))
        ps2 = max(ps2, PC + len(h1) + len(h2))
        pe2 = max(pe2, len(edges))
    cat = lambda v: np.concatenate(v) if v else np.zeros(0, np.int64)
    out = dict(pslots2=ps2, pedges2=pe2, ph2_off=np.array(h2_off), ph2_cell=cat(h2_cell), px_off=np.array(x_off),
               px_edge=cat(x_edge), px_slot=cat(x_slot), pc_off=np.array(c_off), pc_eslot=cat(c_slot))
    assert ps2 < 1 << 15 and pe2 < 0xffff and len(out['pc_eslot']) == 3 * out['pc_off'][-1]
    return out


def transport_patches(grid, pt):
    """The transport's fused flux-and-stage tables on fast_patches' patches (isw_grid's pi_*, tc_slot): a patch's
    edges are its own (pt's pe_edge, slots 0 .. ne - 1) then its incoming ones -- those whose second cell only is in
    it, ascending (slots from ne); tc_slot (N, 6): each cell's edges in cellEdgeAsc order as slot * 2 + 1 where the
    cell is the edge's first, slot * 2 where it is the second, -1 for none. Every invariant the kernel relies on is
    checked here."""
    N, PC, nb = int(grid.N), pt['pcells'], pt['npatch']
    eI, eJ = np.asarray(grid.eI, np.int64), np.asarray(grid.eJ, np.int64)
    asc = np.asarray(grid.cellEdgeAsc, np.int64)
    bI, bJ = eI // PC, eJ // PC
    inc = np.flatnonzero(bJ != bI)
    order = inc[np.lexsort((inc, bJ[inc]))]                           # incoming edges by patch, then ascending
    pi_off = np.searchsorted(bJ[order], np.arange(nb + 1)).astype(np.int64)
    ne = np.diff(pt['pe_off'])
    slot = np.full(len(eI), -1, np.int64)                             # an edge's slot in its first cell's patch
    slot[pt['pe_edge']] = np.arange(len(eI)) - np.repeat(pt['pe_off'][:-1], ne)
    islot = np.full(len(eI), -1, np.int64)                            # its slot in its second cell's patch
    islot[order] = ne[bJ[order]] + np.arange(len(order)) - pi_off[bJ[order]]
    cells = np.repeat(np.arange(N), 6)
    e = asc.reshape(-1)
    ok = e >= 0
    ec = np.where(ok, e, 0)
    first = eI[ec] == cells
    s_ = np.where(first, slot[ec], np.where(bJ[ec] == bI[ec], slot[ec], islot[ec]))
    tc = np.where(ok, 2 * s_ + first, -1).reshape(N, 6)
    ptedges = int(np.max(ne + np.diff(pi_off)))
    # the kernel's assumptions: every cell's edge has a slot in the cell's own patch naming that edge
    own_edges = [np.concatenate([pt['pe_edge'][pt['pe_off'][b]:pt['pe_off'][b + 1]], order[pi_off[b]:pi_off[b + 1]]])
                 for b in range(nb)]
    for b in range(nb):
        rows = np.arange(b * PC, min((b + 1) * PC, N))
        t = tc[rows]
        m = t >= 0
        assert np.array_equal(own_edges[b][t[m] >> 1], asc[rows][m])
        assert np.array_equal((t[m] & 1) == 1, eI[asc[rows][m]] == np.repeat(rows, 6).reshape(-1, 6)[m])
    assert np.all((asc >= 0) == (tc >= 0)) and ptedges < 1 << 14
    return dict(pi_off=pi_off, pi_edge=order, tc_slot=tc, ptedges=ptedges)


def edge_slots(grid):
    """(E, 2): each edge's slot in its first cell's and its second cell's cellEdge ring."""
    ce = np.asarray(grid.cellEdge, np.int64)
    eI, eJ = np.asarray(grid.eI, np.int64), np.asarray(grid.eJ, np.int64)
    out = np.full((len(eI), 2), -1, np.int64)
    for col, cell in ((0, eI), (1, eJ)):
        hit = ce[cell] == np.arange(len(eI))[:, None]
        assert np.all(hit.sum(1) == 1), 'every edge once in each of its cells\' rings'
        out[:, col] = np.argmax(hit, axis=1)
    return out


_PATCHES = {}


def patch_tables(grid):
    """fast_patches and fast_patches2 of a mesh, built once per mesh topology: meshes scaled to other planets share
    their index arrays with the level's mesh, so they share these tables too. Host work only (safe to call from a
    worker thread)."""
    key = (id(grid.eI), int(grid.N), PATCH_CELLS, PATCH_PAIRS)
    hit = _PATCHES.get(key)
    if hit is None or hit[0] is not grid.eI:
        pt = fast_patches(grid, PATCH_CELLS)
        p2 = fast_patches2(grid, pt) if pt is not None and PATCH_PAIRS else None
        if pt is not None and pt['pcells'] == 64:
            pt.update(transport_patches(grid, pt))
        hit = (grid.eI, pt, p2)
        if len(_PATCHES) > 16:
            _PATCHES.clear()
        _PATCHES[key] = hit
    return hit[1], hit[2]


class DeviceGrid:
    """A Grid's arrays on the GPU: indices int32/int8, geometry float64, reciprocals and FP32 weight copies."""

    INT32 = ('eI', 'eJ', 'nbr', 'cellEdge', 'cellEdgeAsc', 'ev1', 'ev2', 'vEdge', 'vCell', 'eoeStart', 'eoe')
    INT8 = ('cellSign', 'vEdgeSign', 'vSlot')
    REALS = ('le', 'de', 'areaCell', 'areaV', 'fV', 'kite', 'w')

    def __init__(self, grid, precision=None):
        _hip.lib()
        self.host = grid
        self.n, self.N, self.E, self.NV = grid.n, grid.N, grid.E, grid.NV
        self.precision = precision or PRECISION
        assert self.precision in ('mixed', 'double')
        self.mixed = self.precision == 'mixed'
        self.prec = _hip.ISW_MIXED if self.mixed else _hip.ISW_FP64
        self.dev = dev = device()
        self.dtype = torch.float64
        put = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)
        for k in self.INT32:
            setattr(self, k, put(getattr(grid, k), np.int32))
        for k in self.INT8:
            setattr(self, k, put(getattr(grid, k), np.int8))
        for k in self.REALS:
            setattr(self, k, put(getattr(grid, k), np.float64))
        idx, wt = grid.d2 if hasattr(grid, 'd2') else d2_stencil(grid)
        self.d2idx, self.d2wt = put(idx, np.int32), put(wt, np.float64)
        self.rA, self.rde, self.rAV = 1.0 / self.areaCell, 1.0 / self.de, 1.0 / self.areaV
        self.wf, self.d2wtf = self.w.float(), self.d2wt.float()
        s = _hip.IswGrid()
        s.N, s.E, s.NV, s.n_eoe = self.N, self.E, self.NV, int(self.eoe.numel())
        for k in self.INT32 + ('d2idx',):
            setattr(s, k, ctypes.cast(getattr(self, k).data_ptr(), _hip._i32p))
        for k in self.INT8:
            setattr(s, k, ctypes.cast(getattr(self, k).data_ptr(), _hip._i8p))
        for k in self.REALS + ('d2wt', 'rA', 'rde', 'rAV'):
            setattr(s, k, ctypes.cast(getattr(self, k).data_ptr(), _hip._f64p))
        for k in ('wf', 'd2wtf'):
            setattr(s, k, ctypes.cast(getattr(self, k).data_ptr(), _hip._f32p))
        pt, p2 = patch_tables(grid)
        if pt is not None:
            # each edge's slot in its two cells' rings (where the stage's coefficients of its cells are written)
            self.eslot = put(edge_slots(grid), np.int8)
            s.eslot = ctypes.cast(self.eslot.data_ptr(), _hip._i8p)
            s.npatch, s.pcells, s.pslots = pt['npatch'], pt['pcells'], pt['pslots']
            s.pown = int(np.max(np.diff(pt['pe_off'])))
            for k in ('ph_off', 'ph_cell', 'pe_off', 'pe_edge', 'pe_slot'):
                setattr(self, k, put(pt[k], np.int32))
                setattr(s, k, ctypes.cast(getattr(self, k).data_ptr(), _hip._i32p))
            if 'pi_off' in pt:
                s.ptedges = pt['ptedges']
                for k in ('pi_off', 'pi_edge', 'tc_slot'):
                    setattr(self, k, put(pt[k], np.int32))
                    setattr(s, k, ctypes.cast(getattr(self, k).data_ptr(), _hip._i32p))
            if p2 is not None:
                s.pslots2, s.pedges2 = p2['pslots2'], p2['pedges2']
                for k in ('ph2_off', 'ph2_cell', 'px_off', 'px_edge', 'px_slot', 'pc_off', 'pc_eslot'):
                    setattr(self, k, put(p2[k], np.int32))
                    setattr(s, k, ctypes.cast(getattr(self, k).data_ptr(), _hip._i32p))
        self.s = s
        self.ref = ctypes.byref(s)
        self._sum_scratch = None

    def zeros(self, *shape, dtype=torch.float64):
        return torch.zeros(*shape, dtype=dtype, device=self.dev)

    def asarray(self, a):
        return torch.as_tensor(np.asarray(a, dtype=np.float64), device=self.dev).clone()

    def _scratch(self, n):
        need = 2 * (n + 2047) // 2048 + (n + 2048 * 2048 - 1) // (2048 * 2048) + 8
        if self._sum_scratch is None or self._sum_scratch.numel() < need:
            self._sum_scratch = torch.zeros(need, dtype=torch.float64, device=self.dev)
        return self._sum_scratch


def _sum(name, dg, x, y=None):
    x = x.contiguous().reshape(-1)
    y = None if y is None else y.contiguous().reshape(-1)
    out = torch.zeros(2, dtype=torch.float64, device=x.device)
    _hip.call(name, _hip.ptr(x), _hip.ptr(y), x.numel(), _hip.ptr(dg._scratch(x.numel())), _hip.ptr(out), _hip.stream())
    return out


def dsum(dg, x, y=None):
    """The reference's deterministic tree sum of x (or x*y) (src/insilica/gpu.py:111-126). Returns a float."""
    return float(_sum('isw_dsum', dg, x, y)[0])


def csum(dg, x, y=None):
    """Compensated deterministic sum of x (or x*y): the budgets' sum. Returns a float."""
    return float(_sum('isw_csum', dg, x, y)[0])
