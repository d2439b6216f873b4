"""TRiSK C-grid geometry on the Voronoi mesh (Thuburn et al. 2009; Ringler et al. 2010) and the gather tables the
HIP kernels read.

Cells hold mass and tracers; edges the normal wind (positive from eI to eJ, tangent from vertex ev1 to ev2);
vertices (the Delaunay triangles' circumcentres) the vorticity. The tables and their orders are the reference's
(src/insilica/grid/trisk.py:141-269): vertices numbered by first appearance over (cell, slot), kites per
(cell, slot), the tangential weights W in CSR with cell eI's edges first then eJ's, each walked counter-clockwise
from the edge, and the ascending-order gather tables that make every sum run in the reference's order.
"""
import numpy as np

from ..constants import OMEGA, R_PLANET
from .mesh import tri_area, vcross, vdot, vnorm


@np.errstate(invalid='ignore', divide='ignore')
def geometry(g):
    N, E = g['N'], g['E']
    pos = g['pos']
    nbr, deg = g['nbr'].astype(np.int64), g['deg'].astype(np.int64)
    eI, eJ = g['eI'].astype(np.int64), g['eJ'].astype(np.int64)
    eKi, eKj = g['eKi'].astype(np.int64), g['eKj'].astype(np.int64)
    slot = np.arange(6)[None, :]
    valid = slot < deg[:, None]
    nxt = np.where(valid, (slot + 1) % deg[:, None], slot)

    # vertices: one per Delaunay triangle (sorted cell triple), numbered by first appearance over (cell, slot)
    nb = np.where(valid, nbr, 0)
    tri = np.sort(np.stack([np.broadcast_to(np.arange(N)[:, None], (N, 6)), nb, np.take_along_axis(nb, nxt, axis=1)], -1), -1)
    key = np.where(valid, (tri[..., 0] * N + tri[..., 1]) * N + tri[..., 2], -1).ravel()
    live = np.flatnonzero(key >= 0)
    uniq, first, inv = np.unique(key[live], return_index=True, return_inverse=True)
    NV = len(uniq)
    rank = np.empty(NV, dtype=np.int64)
    rank[np.argsort(first, kind='stable')] = np.arange(NV)
    corner_v = np.full(N * 6, -1, dtype=np.int64)
    corner_v[live] = rank[inv]
    corner_v = corner_v.reshape(N, 6)
    v_pos = np.empty((NV, 3))
    v_pos[rank] = g['corners'].reshape(N * 6, 3)[live[first]]

    cell_edge = np.full((N, 6), -1, dtype=np.int64)
    cell_sign = np.zeros((N, 6), dtype=np.int8)
    eid = np.arange(E)
    cell_edge[eI, eKi], cell_sign[eI, eKi] = eid, 1
    cell_edge[eJ, eKj], cell_sign[eJ, eKj] = eid, -1
    mi = deg[eI]
    ev1 = corner_v[eI, (eKi + mi - 1) % mi]
    ev2 = corner_v[eI, eKi]
    de = g['dist'][eI, eKi]
    le = g['edgeLen'][eI, eKi]
    pi, pj = pos[eI], pos[eJ]
    s = pi + pj
    xe = s / vnorm(s)[:, None]
    d = pj - pi
    d = d - vdot(d, xe)[:, None] * xe
    ne = d / vnorm(d)[:, None]

    # kite (cell i, slot k): the quadrilateral (x_i, edge k's crossing, vertex k, edge k+1's crossing)
    ce = np.where(valid, cell_edge, 0)
    ce1 = np.take_along_axis(ce, nxt, axis=1)
    vv = np.where(valid, corner_v, 0)
    Pc = np.broadcast_to(pos[:, None, :], (N, 6, 3))
    kite = (tri_area(Pc, xe[ce], v_pos[vv]) + tri_area(Pc, v_pos[vv], xe[ce1])) * R_PLANET * R_PLANET
    kite = np.where(valid, kite, 0.0)
    area_cell = np.zeros(N)
    for k in range(6):
        area_cell = area_cell + kite[:, k]
    area_v = np.zeros(NV)
    np.add.at(area_v, corner_v[valid], kite[valid])          # sequential in (cell, slot) order

    def circulation(vi):                                      # +1 where n_e runs counter-clockwise about vertex vi
        return np.sign(vdot(ne, vcross(xe, xe - v_pos[vi]))).astype(np.int8)

    circ1, circ2 = circulation(ev1), circulation(ev2)

    # tangential weights: around each of the edge's two cells, counter-clockwise from the edge, the running kite
    # fraction of the vertices passed so far; weight (1/2 - R) * sign * l_e' / d_e, negated on the eJ side
    cols_e, cols_w, cols_ok = [], [], []
    for c, kc, side in ((eI, eKi, 1.0), (eJ, eKj, -1.0)):
        m = deg[c]
        run = np.zeros(E)
        for step in range(1, 6):
            ok = step < m
            k = (kc + step) % m
            run = run + np.where(ok, kite[c, (k + m - 1) % m] / area_cell[c], 0.0)
            e2 = cell_edge[c, k]
            cols_e.append(e2)
            cols_w.append(side * (0.5 - run) * cell_sign[c, k] * le[e2] / de)
            cols_ok.append(ok)
    ok = np.stack(cols_ok, axis=1)
    eoe = np.stack(cols_e, axis=1)[ok]
    w = np.stack(cols_w, axis=1)[ok]
    eoe_start = np.concatenate([[0], np.cumsum(ok.sum(axis=1))])
    fV = 2 * OMEGA * v_pos[:, 2]
    xm = v_pos[ev1] + v_pos[ev2]
    xm = xm / vnorm(xm)[:, None]

    t = dict(NV=NV, vPos=v_pos, xm=xm, cornerV=corner_v.astype(np.int32), cellEdge=cell_edge.astype(np.int32),
             cellSign=cell_sign, ev1=ev1.astype(np.int32), ev2=ev2.astype(np.int32), de=de, le=le, xe=xe, ne=ne,
             kite=kite, areaCell=area_cell, areaV=area_v, circ1=circ1, circ2=circ2, fV=fV,
             eoeStart=eoe_start.astype(np.int32), eoe=eoe.astype(np.int32), w=w)
    t.update(gather_tables(g, t))
    return t


def gather_tables(g, t):
    """cellEdgeAsc (N, 6): a cell's edges ascending (pads -1 last); vEdge/vEdgeSign (NV, 3): a vertex's edges ascending
    with their circulation signs; vCell/vSlot (NV, 3): a vertex's kites by ascending cell."""
    N, E, NV = g['N'], g['E'], t['NV']
    ce = t['cellEdge'].astype(np.int64)
    order = np.argsort(np.where(ce < 0, np.iinfo(np.int64).max, ce), axis=1, kind='stable')
    asc = np.take_along_axis(ce, order, axis=1)
    ev = np.concatenate([t['ev1'], t['ev2']]).astype(np.int64)
    ee = np.concatenate([np.arange(E), np.arange(E)])
    sg = np.concatenate([t['circ1'], t['circ2']])
    o = np.lexsort((ee, ev))
    deg = g['deg'].astype(np.int64)
    valid = np.arange(6)[None, :] < deg[:, None]
    cells = np.broadcast_to(np.arange(N)[:, None], (N, 6))[valid]
    slots = np.broadcast_to(np.arange(6)[None, :], (N, 6))[valid]
    vs = t['cornerV'].astype(np.int64)[valid]
    ov = np.lexsort((cells, vs))
    return dict(cellEdgeAsc=asc.astype(np.int32), vEdge=ee[o].reshape(NV, 3).astype(np.int32),
                vEdgeSign=sg[o].reshape(NV, 3).astype(np.int8), vCell=cells[ov].reshape(NV, 3).astype(np.int32),
                vSlot=slots[ov].reshape(NV, 3).astype(np.int8))


def cell_winds(g, u):
    """Perot wind vectors at cells (m/s, (N, 3)) from edge normal winds u (E,):
    U_i = (1/A_i) sum_e s l_e (x_e - x_i) u_e with x_e the Voronoi edge midpoint, projected on the tangent plane."""
    ce = np.where(g.cellEdge >= 0, g.cellEdge, 0)
    c = np.where(g.cellEdge >= 0, g.cellSign * u[ce] * g.le[ce] * getattr(g, 'radius', R_PLANET), 0.0)
    v = np.sum(c[..., None] * (g.xm[ce] - g.pos[:, None, :]), axis=1) / g.areaCell[:, None]
    return v - np.sum(v * g.pos, axis=1)[:, None] * g.pos
