"""The planet's mesh: icosahedral Voronoi cells smoothed to an SCVT, with TRiSK C-grid geometry, built once per level
n and cached on disk as .npz (`$INSILICA_CACHE/grids`, default ~/.cache/insilica_amd/grids).

Same surface as the reference's grid package (src/insilica/grid/__init__.py:39-127): `get(n)`, `make(n, order)`,
`for_planet(g, radius, omega, grav)`, `planet_of(g)`, `Grid` with the reference's array names. Cells are renumbered
along a Hilbert curve by default (order='hilbert'); 'v1' keeps the icosphere's numbering.
"""
import os
import time
from pathlib import Path

import numpy as np

from .mesh import CELL_KEYS, build, hilbert_perm, renumber
from .trisk import geometry

GRID_VERSION = 1
DEFAULT_ORDER = 'hilbert'
CACHE_DIR = Path(os.environ.get('INSILICA_CACHE', Path.home() / '.cache' / 'insilica_amd'))
LENGTHS, AREAS = ('de', 'dist', 'edgeLen', 'le'), ('area', 'areaCell', 'areaV', 'kite', 'totalArea')
BASE_KEYS = ('n', 'N', 'totalArea', 'lloydIters', 'lloydMoved', 'nbr') + CELL_KEYS


class Grid:
    """Host (NumPy) copy of one mesh level; attributes carry the reference's array names."""

    def __init__(self, arrays):
        self._keys = sorted(arrays)
        for k, v in arrays.items():
            setattr(self, k, v)
        self.n, self.N, self.E, self.NV = int(self.n), int(self.N), int(self.E), int(self.NV)
        self.order = str(arrays.get('order', 'v1'))

    def arrays(self):
        return {k: getattr(self, k) for k in self._keys}

    def __repr__(self):
        return f'Grid(n={self.n}, N={self.N}, E={self.E}, NV={self.NV}, order={self.order})'


def for_planet(g, radius, omega, grav):
    """The grid scaled to a planet of this radius (m) and rotation (rad/s); carries radius, omega, grav. Earth's
    returns g itself (reference: grid/__init__.py:39-53)."""
    from ..constants import GRAV, OMEGA, R_PLANET
    if radius == R_PLANET and omega == OMEGA and grav == GRAV:
        return g
    a, s = g.arrays(), radius / R_PLANET
    for k in LENGTHS:
        a[k] = a[k] * s
    for k in AREAS:
        a[k] = a[k] * s * s
    a['fV'] = a['fV'] * (omega / OMEGA)
    a.update(radius=float(radius), omega=float(omega), grav=float(grav))
    return Grid(a)


def planet_of(g):
    from ..constants import GRAV, OMEGA, R_PLANET
    return getattr(g, 'radius', R_PLANET), getattr(g, 'omega', OMEGA), getattr(g, 'grav', GRAV)


def make(n, order=None):
    """Build level n from scratch. Returns (Grid, seconds)."""
    order = DEFAULT_ORDER if order is None else order
    t0 = time.perf_counter()
    g = build(n)
    if order == 'hilbert':
        g = renumber(g, hilbert_perm(g['pos']))
    elif order != 'v1':
        raise ValueError(f'unknown cell order {order!r}')
    g.update(geometry(g))
    g['order'] = order
    return Grid(g), time.perf_counter() - t0


_mem = {}


def get(n, use_cache=True, order=None):
    """Level n from memory, the disk cache, or built (and cached)."""
    order = DEFAULT_ORDER if order is None else order
    key = (n, order)
    if key in _mem:
        return _mem[key]
    f = CACHE_DIR / 'grids' / f'n{n}-{order}-g{GRID_VERSION}.npz'
    g = None
    if use_cache and f.exists():
        try:
            with np.load(f, allow_pickle=False) as z:
                g = Grid({k: (z[k][()] if z[k].ndim == 0 else z[k]) for k in z.files})
        except (OSError, ValueError):
            g = None
    if g is None:
        if order != 'v1':
            base = get(n, use_cache, 'v1')
            b = {k: getattr(base, k) for k in BASE_KEYS}
            b.update({k: getattr(base, k) for k in ('E', 'eI', 'eJ', 'eKi', 'eKj')})
            a = renumber(b, hilbert_perm(base.pos))
            a.update(geometry(a))
            a['order'] = order
            g = Grid(a)
        else:
            g, _ = make(n, order='v1')
        if use_cache:
            try:
                f.parent.mkdir(parents=True, exist_ok=True)
                tmp = f.with_suffix(f'.{os.getpid()}.tmp.npz')
                np.savez(tmp, **g.arrays())
                os.replace(tmp, f)
            except OSError:
                pass
    _mem[key] = g
    return g


def d2_stencil(g):
    """Per edge end, weights on four cells (c, j, o1, o2) for the second derivative along the line to the neighbour,
    scaled by d_e^2/12, for the 3rd-order transport flux (reference: src/insilica/gpu.py:86-108). o1, o2 are the
    cells opposite the neighbour (one on a hexagon, the two across on a pentagon). Returns idx int32 (E*8), wt (E*8)."""
    E = g.E
    nbr, deg, dist = g.nbr.astype(np.int64), g.deg.astype(np.int64), g.dist
    idx = np.zeros((E, 2, 4), dtype=np.int32)
    wt = np.zeros((E, 2, 4))
    for side, (c, k) in enumerate(((g.eI.astype(np.int64), g.eKi.astype(np.int64)),
                                   (g.eJ.astype(np.int64), g.eKj.astype(np.int64)))):
        m = deg[c]
        j = nbr[c, k]
        d1 = dist[c, k]
        hexa = m == 6
        ka = np.where(hexa, (k + 3) % 6, (k + 2) % m)
        kb = np.where(hexa, (k + 3) % 6, (k + 3) % m)
        d2 = np.where(hexa, dist[c, (k + 3) % 6], 0.5 * (dist[c, ka] + dist[c, kb]))
        s = 2 / (d1 + d2) * g.de * g.de / 12
        idx[:, side] = np.stack([c, j, nbr[c, ka], nbr[c, kb]], axis=1)
        wt[:, side, 0] = -s * (1 / d1 + 1 / d2)
        wt[:, side, 1] = s / d1
        wt[:, side, 2] = 0.5 * s / d2
        wt[:, side, 3] = 0.5 * s / d2
    return idx.reshape(-1), wt.reshape(-1)
