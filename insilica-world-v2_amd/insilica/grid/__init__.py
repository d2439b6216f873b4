"""The planet's mesh: the icosahedral Voronoi grid smoothed toward a spherical centroidal Voronoi tessellation, with
the TRiSK C-grid tables, built by the native mesh library (csrc/mesh.cpp, libinsilica_mesh.so) and cached on disk as
.npz (`$INSILICA_CACHE/grids`, default ~/.cache/insilica_amd/grids).

The surface is the reference's grid package's (src/insilica/grid/__init__.py:39-127): `get(n)`, `make(n, order)`,
`for_planet(g, radius, omega, grav)`, `planet_of(g)`, `Grid` carrying the reference's array names (cells `pos`, `lat`,
`lon`, `nbr`, `deg`, `area`, `dist`, `edgeLen`, `corners`, ...; edges `eI`, `eJ`, `le`, `de`, ...; vertices `vPos`,
`areaV`, ...), numbered and valued as the reference's generator numbers and values them (pinned by a level-3
fixture of its arrays). Cells follow a Hilbert curve by default (order='hilbert'); 'v1' keeps the icosphere's order.

The native stages do the topology, the Lloyd iteration and the algebra of every small triangle; the element-wise
arcsin and arctan2 that turn chords and triple products into angles run here in NumPy (see csrc/mesh.cpp).
"""
import ctypes
import math
import os
import time
from pathlib import Path

import numpy as np

GRID_VERSION = 2
DEFAULT_ORDER = 'hilbert'
CACHE_DIR = Path(os.environ.get('INSILICA_CACHE', Path.home() / '.cache' / 'insilica_amd'))
LLOYD_TOL, LLOYD_CAP = 1e-11, 200000
GPU_LLOYD_FROM_N = 6          # levels from which Lloyd runs on the GPU when one is present (as the reference's)
LLOYD_CHUNK = 512             # device Lloyd iterations queued between reads of their moves
HILBERT_BITS = 21
# what scales with the planet's radius (lengths, then areas)
LENGTHS, AREAS = ('de', 'dist', 'edgeLen', 'le'), ('area', 'areaCell', 'areaV', 'kite', 'totalArea')

_lib = None


def _native():
    global _lib
    if _lib is None:
        path = Path(__file__).resolve().parents[1] / '_lib' / 'libinsilica_mesh.so'
        if not path.exists():
            raise RuntimeError(f'{path} is missing: build the native libraries first (python __graft_entry__.py)')
        L = ctypes.CDLL(str(path))
        vp, i64, i32, d = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
        p64 = ctypes.POINTER(ctypes.c_int64)
        for name, args in {
            'isw_mesh_sizes': [i32, p64, p64, p64, p64],
            'isw_mesh_icosphere': [i32, vp, vp, vp, vp],
            'isw_mesh_ring': [i64, vp, vp, vp, vp],
            'isw_mesh_lloyd': [i64, vp, vp, vp, i32, d, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(d)],
            'isw_mesh_cells': [i64, vp, vp, vp, vp, vp, vp, vp, vp, vp],
            'isw_mesh_edges': [i64, vp, vp, vp, vp, vp, vp],
            'isw_mesh_hilbert': [i64, vp, i32, vp],
            'isw_mesh_trisk': [i64, i64, i64] + [vp] * 25,
            'isw_mesh_weights': [i64, i64] + [vp] * 14,
        }.items():
            f = getattr(L, name)
            f.restype, f.argtypes = ctypes.c_int, args
        _lib = L
    return _lib


def _call(name, *args):
    rc = getattr(_native(), name)(*[a.ctypes.data if isinstance(a, np.ndarray) else a for a in args])
    if rc != 0:
        raise RuntimeError(f'{name} failed ({rc})')


class Grid:
    """A mesh level's host arrays as attributes (the reference's names), plus n, N, E, NV and the cell order."""

    def __init__(self, arrays):
        self._names = sorted(arrays)
        for k, v in arrays.items():
            setattr(self, k, v)
        self.n, self.N, self.E, self.NV = int(self.n), int(self.N), int(self.E), int(self.NV)
        self.order = str(arrays.get('order', 'v1'))

    def arrays(self):
        return {k: getattr(self, k) for k in self._names}

    def __repr__(self):
        return f'Grid(n={self.n}, N={self.N}, E={self.E}, NV={self.NV}, order={self.order})'


def _base_vertices():
    """The unit icosahedron's 12 vertices before normalisation: the poles and two rings of five at latitude
    +-atan(1/2), the lower ring turned by half a step (NumPy's trigonometry, as the reference builds them)."""
    lat = np.arctan(0.5)
    up = [(np.cos(lat) * np.cos(k * 2 * np.pi / 5), np.cos(lat) * np.sin(k * 2 * np.pi / 5), np.sin(lat))
          for k in range(5)]
    lo = [(np.cos(lat) * np.cos((k + 0.5) * 2 * np.pi / 5), np.cos(lat) * np.sin((k + 0.5) * 2 * np.pi / 5),
           -np.sin(lat)) for k in range(5)]
    return np.ascontiguousarray([(0.0, 0.0, 1.0)] + up + lo + [(0.0, 0.0, -1.0)], dtype=np.float64)


def _edges(nbr, deg, E):
    e = {k: np.zeros(E, dt) for k, dt in (('eI', np.int32), ('eJ', np.int32), ('eKi', np.uint8), ('eKj', np.uint8))}
    _call('isw_mesh_edges', len(deg), nbr, deg, e['eI'], e['eJ'], e['eKi'], e['eKj'])
    return e


def _device_lloyd(pos, ring, deg):
    """Lloyd's iteration in the HIP library (isw_mesh_lloyd_device) on the current GPU: (pos, iterations, last move)."""
    import torch
    from .. import _hip
    N = len(pos)
    dev = torch.device('cuda', torch.cuda.current_device())
    P = torch.from_numpy(np.ascontiguousarray(pos)).to(dev)
    R = torch.from_numpy(np.ascontiguousarray(ring)).to(dev)
    D = torch.from_numpy(np.ascontiguousarray(deg)).to(dev)
    W = torch.empty(3 * N + LLOYD_CHUNK, dtype=torch.float64, device=dev)
    it, mv = ctypes.c_int32(0), ctypes.c_double(0.0)
    _hip.check(_hip.lib().isw_mesh_lloyd_device(N, P.data_ptr(), R.data_ptr(), D.data_ptr(), LLOYD_CAP, LLOYD_TOL,
                                         W.data_ptr(), LLOYD_CHUNK, ctypes.byref(it), ctypes.byref(mv),
                                         torch.cuda.current_stream(dev).cuda_stream), 'isw_mesh_lloyd_device')
    return np.ascontiguousarray(P.cpu().numpy()), it.value, mv.value


def _lloyd_on(n, lloyd):
    if lloyd is None:
        if n < GPU_LLOYD_FROM_N:
            return 'host'
        import torch
        return 'device' if torch.cuda.is_available() else 'host'
    if lloyd not in ('host', 'device'):
        raise ValueError(f'unknown Lloyd placement {lloyd!r}')
    return lloyd


def _icosahedral(n, lloyd='host'):
    """The level-n mesh in the icosphere's own numbering: generators after Lloyd (on the host, or with
    lloyd='device' on the GPU), rings, the Voronoi cells' corners, distances, edge lengths and areas (m, m^2 on the
    Earth-sized sphere), the edge list."""
    from ..constants import R_PLANET
    N, E, NV, ne = (ctypes.c_int64() for _ in range(4))
    _call('isw_mesh_sizes', n, ctypes.byref(N), ctypes.byref(E), ctypes.byref(NV), ctypes.byref(ne))
    N, E = N.value, E.value
    pos = np.zeros((N, 3))
    adj, deg = np.zeros((N, 6), np.int32), np.zeros(N, np.uint8)
    _call('isw_mesh_icosphere', n, _base_vertices(), pos, adj, deg)
    ring = np.zeros((N, 6), np.int32)
    _call('isw_mesh_ring', N, pos, adj, deg, ring)
    if lloyd == 'device':
        pos, it, mv = _device_lloyd(pos, ring, deg)
    else:
        it, mv = ctypes.c_int(0), ctypes.c_double(0.0)
        _call('isw_mesh_lloyd', N, pos, ring, deg, LLOYD_CAP, LLOYD_TOL, ctypes.byref(it), ctypes.byref(mv))
        it, mv = it.value, mv.value
    _call('isw_mesh_ring', N, pos, adj, deg, ring)
    east, north = np.zeros((N, 3)), np.zeros((N, 3))
    corners = np.zeros((N, 6, 3))
    chord_d, chord_e, tri = np.zeros((N, 6)), np.zeros((N, 6)), np.zeros((N, 6, 2))
    _call('isw_mesh_cells', N, pos, ring, deg, east, north, corners, chord_d, chord_e, tri)
    live = np.arange(6)[None, :] < deg[:, None]
    arc = lambda chord: R_PLANET * (2 * np.arcsin(np.minimum(1.0, 0.5 * chord)))
    fan = 2 * np.arctan2(np.abs(tri[..., 0]), tri[..., 1])
    area = np.zeros(N)
    for k in range(6):                                   # each cell's fan in slot order
        area = area + np.where(live[:, k], fan[:, k], 0.0)
    area = area * R_PLANET * R_PLANET
    g = dict(n=n, N=N, pos=pos, lat=np.arcsin(np.clip(pos[:, 2], -1, 1)), lon=np.arctan2(pos[:, 1], pos[:, 0]),
             east=east, north=north, nbr=ring, deg=deg, corners=corners, area=area,
             dist=np.where(live, arc(chord_d), 0.0), edgeLen=np.where(live, arc(chord_e), 0.0),
             totalArea=float(np.sum(area)), lloydIters=it, lloydMoved=mv, lloydOn=lloyd, E=E)
    g.update(_edges(ring, deg, E))
    return g


CELL_ARRAYS = ('pos', 'lat', 'lon', 'area', 'east', 'north', 'deg', 'dist', 'edgeLen', 'corners')


def _hilbert(g):
    """The same mesh with its cells along the Hilbert curve (new cell k = old cell perm[k]); edges listed again."""
    perm = np.zeros(g['N'], np.int32)
    _call('isw_mesh_hilbert', g['N'], np.ascontiguousarray(g['pos']), HILBERT_BITS, perm)
    back = np.empty_like(perm)
    back[perm] = np.arange(len(perm), dtype=np.int32)
    out = dict(g)
    for k in CELL_ARRAYS:
        out[k] = np.ascontiguousarray(g[k][perm])
    nb = g['nbr'][perm]
    out['nbr'] = np.ascontiguousarray(np.where(nb >= 0, back[np.maximum(nb, 0)], -1).astype(np.int32))
    out.update(_edges(out['nbr'], out['deg'], g['E']))
    out['perm'] = perm
    return out


def _tables(g):
    """The TRiSK tables of a mesh in its current cell order (vertices, edge geometry, kites, areas, weights, the
    ascending gather tables)."""
    from ..constants import OMEGA, R_PLANET
    N, E = g['N'], g['E']
    NV, ne = ctypes.c_int64(), ctypes.c_int64()
    _call('isw_mesh_sizes', g['n'], ctypes.byref(ctypes.c_int64()), ctypes.byref(ctypes.c_int64()),
          ctypes.byref(NV), ctypes.byref(ne))
    NV = NV.value
    t = dict(NV=NV, vPos=np.zeros((NV, 3)), cornerV=np.zeros((N, 6), np.int32), cellEdge=np.zeros((N, 6), np.int32),
             cellSign=np.zeros((N, 6), np.int8), ev1=np.zeros(E, np.int32), ev2=np.zeros(E, np.int32),
             xe=np.zeros((E, 3)), ne=np.zeros((E, 3)), xm=np.zeros((E, 3)), circ1=np.zeros(E, np.int8),
             circ2=np.zeros(E, np.int8), cellEdgeAsc=np.zeros((N, 6), np.int32), vEdge=np.zeros((NV, 3), np.int32),
             vEdgeSign=np.zeros((NV, 3), np.int8), vCell=np.zeros((NV, 3), np.int32),
             vSlot=np.zeros((NV, 3), np.int8))
    kt = np.zeros((N, 6, 2, 2))
    _call('isw_mesh_trisk', N, E, NV, g['pos'], g['nbr'], g['deg'], g['corners'], g['eI'], g['eJ'], g['eKi'],
          g['eKj'], t['vPos'], t['cornerV'], t['cellEdge'], t['cellSign'], t['ev1'], t['ev2'], t['xe'], t['ne'],
          t['xm'], kt, t['circ1'], t['circ2'], t['cellEdgeAsc'], t['vEdge'], t['vEdgeSign'], t['vCell'], t['vSlot'])
    live = np.arange(6)[None, :] < g['deg'][:, None]
    half = 2 * np.arctan2(np.abs(kt[..., 0]), kt[..., 1])
    kite = np.where(live, (half[..., 0] + half[..., 1]) * R_PLANET * R_PLANET, 0.0)
    area = np.zeros(N)
    for k in range(6):
        area = area + kite[:, k]
    area_v = np.zeros(NV)
    np.add.at(area_v, t['cornerV'][live], kite[live])   # (sequential, in (cell, slot) order)
    eI, eKi = g['eI'].astype(np.int64), g['eKi'].astype(np.int64)
    de, le = np.ascontiguousarray(g['dist'][eI, eKi]), np.ascontiguousarray(g['edgeLen'][eI, eKi])
    start, eoe, w = np.zeros(E + 1, np.int32), np.zeros(ne.value, np.int32), np.zeros(ne.value)
    _call('isw_mesh_weights', N, E, g['deg'], g['eI'], g['eJ'], g['eKi'], g['eKj'], t['cellEdge'], t['cellSign'],
          np.ascontiguousarray(kite), area, le, de, start, eoe, w)
    t.update(kite=kite, areaCell=area, areaV=area_v, de=de, le=le, eoeStart=start, eoe=eoe, w=w,
             fV=2 * OMEGA * t['vPos'][:, 2])
    return t


def make(n, order=None, lloyd=None):
    """Build level n from scratch. Returns (Grid, seconds). lloyd: 'host', 'device' (the GPU's), or None for the
    GPU from level GPU_LLOYD_FROM_N when one is present (the reference's rule; the two agree to rounding)."""
    order = DEFAULT_ORDER if order is None else order
    if order not in ('v1', 'hilbert'):
        raise ValueError(f'unknown cell order {order!r}')
    lloyd = _lloyd_on(n, lloyd)
    t0 = time.perf_counter()
    g = _icosahedral(n, lloyd)
    if order == 'hilbert':
        g = _hilbert(g)
    g.update(_tables(g))
    g['order'] = order
    return Grid(g), time.perf_counter() - t0


_loaded = {}


def get(n, use_cache=True, order=None):
    """Level n from memory, the disk cache, or built (and then cached)."""
    order = DEFAULT_ORDER if order is None else order
    key = (n, order)
    if key in _loaded:
        return _loaded[key]
    path = CACHE_DIR / 'grids' / f'n{n}-{order}-g{GRID_VERSION}.npz'
    g = None
    if use_cache and path.exists():
        try:
            with np.load(path, allow_pickle=False) as z:
                g = Grid({k: (z[k][()] if z[k].ndim == 0 else z[k]) for k in z.files})
        except (OSError, ValueError):
            g = None
    if g is None:
        g, _ = make(n, order)
        if use_cache:
            try:
                path.parent.mkdir(parents=True, exist_ok=True)
                tmp = path.with_suffix(f'.{os.getpid()}.tmp.npz')
                np.savez(tmp, **g.arrays())
                os.replace(tmp, path)
            except OSError:
                pass
    _loaded[key] = g
    return g


def for_planet(g, radius, omega, grav):
    """The mesh scaled to a planet of this radius (m) and rotation (rad/s), carrying radius, omega and grav; Earth's
    own settings return g itself (reference: grid/__init__.py:39-53)."""
    from ..constants import GRAV, OMEGA, R_PLANET
    if (radius, omega, grav) == (R_PLANET, OMEGA, GRAV):
        return g
    a = g.arrays()
    s = radius / R_PLANET
    for k in LENGTHS:
        a[k] = a[k] * s
    for k in AREAS:
        a[k] = a[k] * s * s
    a['fV'] = a['fV'] * (omega / OMEGA)
    a.update(radius=float(radius), omega=float(omega), grav=float(grav))
    return Grid(a)


def planet_of(g):
    """(radius, rotation, gravity) the mesh was scaled for."""
    from ..constants import GRAV, OMEGA, R_PLANET
    return getattr(g, 'radius', R_PLANET), getattr(g, 'omega', OMEGA), getattr(g, 'grav', GRAV)


def d2_stencil(g):
    """The 3rd-order transport flux's curvature stencil per edge end (reference: src/insilica/gpu.py:86-108): for
    the side cell c of an edge and its neighbour j across it, the second derivative along the line c-j from c, j and
    the cell(s) opposite j on c's ring (one on a hexagon, the mean of the two across on a pentagon), times d_e^2 / 12.
    Returns idx int32 (E * 8) and wt (E * 8): per edge, side eI then eJ, cells (c, j, o1, o2)."""
    nbr, deg, dist = g.nbr.astype(np.int64), g.deg.astype(np.int64), g.dist
    idx = np.zeros((g.E, 2, 4), dtype=np.int32)
    wt = np.zeros((g.E, 2, 4))
    sides = ((g.eI.astype(np.int64), g.eKi.astype(np.int64)), (g.eJ.astype(np.int64), g.eKj.astype(np.int64)))
    for side, (c, k) in enumerate(sides):
        m = deg[c]
        hexagon = m == 6
        across = (k + 3) % 6
        ka = np.where(hexagon, across, (k + 2) % m)
        kb = np.where(hexagon, across, (k + 3) % m)
        near = dist[c, k]
        far = np.where(hexagon, dist[c, across % 6], 0.5 * (dist[c, ka] + dist[c, kb]))
        s = 2 / (near + far) * g.de * g.de / 12
        idx[:, side] = np.stack([c, nbr[c, k], nbr[c, ka], nbr[c, kb]], axis=1)
        wt[:, side, 0] = -s * (1 / near + 1 / far)
        wt[:, side, 1] = s / near
        wt[:, side, 2] = 0.5 * s / far
        wt[:, side, 3] = 0.5 * s / far
    return idx.reshape(-1), wt.reshape(-1)
