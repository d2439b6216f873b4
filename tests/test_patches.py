"""The fast sub-steps' patch tables (insilica.gpu.fast_patches, fast_patches2), checked on the host against the data
each phase of the patch kernels reads: one fused sub-step needs Phi* on both cells of every edge it steps; two fused
sub-steps need the first sub-step's winds on every edge of the second's cells, and its Phi* on both cells of each of
those edges. A wrong slot would read another cell's value; these checks name every slot's cell and edge."""
import numpy as np
import pytest

from insilica import grid as gridmod
from insilica.gpu import fast_patches, fast_patches2, transport_patches


def unpack(v):
    return v & 0xffff, v >> 16


@pytest.mark.parametrize('n,pc', [(3, 64), (4, 64), (4, 48), (5, 96)])
def test_single_step_tables(n, pc):
    g = gridmod.get(n)
    t = fast_patches(g, pc)
    N, E = g.N, g.E
    eI, eJ = np.asarray(g.eI), np.asarray(g.eJ)
    assert t['npatch'] == (N + pc - 1) // pc
    assert np.array_equal(np.sort(t['pe_edge']), np.arange(E))           # every edge stepped by exactly one patch
    for b in range(t['npatch']):
        c0, nown = b * pc, min(pc, N - b * pc)
        halo = t['ph_cell'][t['ph_off'][b]:t['ph_off'][b + 1]]
        edges = t['pe_edge'][t['pe_off'][b]:t['pe_off'][b + 1]]
        sI, sJ = unpack(t['pe_slot'][t['pe_off'][b]:t['pe_off'][b + 1]])
        hx = np.concatenate([halo, [-1]])                                   # (a patch may have no halo)
        cell_of = lambda s: np.where(s < pc, c0 + s, hx[np.clip(s - pc, 0, len(hx) - 1)])
        assert np.all(eI[edges] // pc == b)                                 # its edges' first cells are its own
        assert np.all(sI < nown) and np.all(sJ < pc + len(halo)) and np.all((sJ < nown) | (sJ >= pc))
        assert np.array_equal(cell_of(sI), eI[edges]) and np.array_equal(cell_of(sJ), eJ[edges])
        # the halo is exactly the outside cells those edges reach, sorted
        want = np.unique(eJ[edges][eJ[edges] // pc != b])
        assert np.array_equal(halo, want)
        assert pc + len(halo) <= t['pslots']


@pytest.mark.parametrize('n', [3, 4])
def test_two_step_tables(n):
    g = gridmod.get(n)
    t = fast_patches(g, 64)
    t2 = fast_patches2(g, t)
    N, pc = g.N, 64
    eI, eJ, ce = np.asarray(g.eI), np.asarray(g.eJ), np.asarray(g.cellEdge)
    for b in range(t['npatch']):
        c0, nown = b * pc, min(pc, N - b * pc)
        h1 = t['ph_cell'][t['ph_off'][b]:t['ph_off'][b + 1]]
        h2 = t2['ph2_cell'][t2['ph2_off'][b]:t2['ph2_off'][b + 1]]
        own_e = t['pe_edge'][t['pe_off'][b]:t['pe_off'][b + 1]]
        x_e = t2['px_edge'][t2['px_off'][b]:t2['px_off'][b + 1]]
        edges = np.concatenate([own_e, x_e])                               # edge slot -> edge
        slots_cells = np.concatenate([np.arange(c0, c0 + nown), np.full(pc - nown, -1), h1, h2])
        # second sub-step's cells (patch, then halo): each of its edges has a slot holding that edge
        c1 = np.concatenate([np.arange(c0, c0 + nown), h1])
        es = t2['pc_eslot'][3 * t2['pc_off'][b]:3 * t2['pc_off'][b + 1]].reshape(-1, 3)
        es6 = np.stack([es & 0xffff, es >> 16], axis=2).reshape(-1, 6)
        assert len(es6) == len(c1)
        for ci, c in enumerate(c1):
            for m in range(6):
                if ce[c, m] < 0:
                    assert es6[ci, m] == 0xffff
                else:
                    assert edges[es6[ci, m]] == ce[c, m]
        # the first sub-step's winds on the other edges read Phi* at slots naming their cells
        sI, sJ = unpack(t2['px_slot'][t2['px_off'][b]:t2['px_off'][b + 1]])
        assert np.array_equal(slots_cells[sI], eI[x_e]) and np.array_equal(slots_cells[sJ], eJ[x_e])
        assert pc + len(h1) + len(h2) <= t2['pslots2'] and len(edges) <= t2['pedges2']
        assert len(np.intersect1d(h2, c1)) == 0 and len(np.unique(h2)) == len(h2)


@pytest.mark.parametrize('n', [3, 4])
def test_transport_tables(n):
    """The fused transport stage's tables: a patch's edge list is every edge touching its cells, once, its own edges
    first in pe order; each cell's slots name its cellEdgeAsc edges in its own patch's list, odd where it is the
    edge's first cell; every incoming edge's first cell lies outside the patch and its second inside."""
    g = gridmod.get(n)
    t = fast_patches(g, 64)
    r = transport_patches(g, t)
    N, pc = g.N, 64
    eI, eJ, asc = np.asarray(g.eI), np.asarray(g.eJ), np.asarray(g.cellEdgeAsc)
    for b in range(t['npatch']):
        cells = np.arange(b * pc, min((b + 1) * pc, N))
        own = t['pe_edge'][t['pe_off'][b]:t['pe_off'][b + 1]]
        inc = r['pi_edge'][r['pi_off'][b]:r['pi_off'][b + 1]]
        assert np.all(eI[inc] // pc != b) and np.all(eJ[inc] // pc == b) and np.all(np.diff(inc) > 0)
        edges = np.concatenate([own, inc])
        touching = np.unique(asc[cells][asc[cells] >= 0])
        assert np.array_equal(np.sort(edges), touching)
        tc = r['tc_slot'][cells]
        m = asc[cells] >= 0
        assert np.array_equal(edges[tc[m] >> 1], asc[cells][m])
        assert np.array_equal((tc[m] & 1) == 1, eI[asc[cells][m]] == np.repeat(cells, 6).reshape(-1, 6)[m])
        assert len(edges) <= r['ptedges']
    assert np.all((r['tc_slot'] < 0) == (asc < 0))
