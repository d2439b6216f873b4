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
        pt = fast_patches(grid, PATCH_CELLS)
        if pt is not None:
            s.npatch, s.pcells, s.pslots = pt['npatch'], pt['pcells'], pt['pslots']
            for k in ('ph_off', 'ph_cell', 'pe_off', 'pe_edge', 'pe_slot'):
                setattr(self, k, put(pt[k], np.int32))
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
