"""The icosahedral Voronoi mesh: icosphere, Delaunay rings, Lloyd smoothing toward an SCVT, Voronoi geometry, edge
list, and the Hilbert renumbering.

The construction follows the reference's `build_grid` step by step (src/insilica/grid/icosa.py:101-353), because a
drop-in mesh has to number its cells, edges and vertices exactly as the reference does: the same icosahedron and
midpoint numbering, counter-clockwise neighbour rings sorted by angle from local east, Jacobi Lloyd with the poles
pinned to 1e-11 rad, corners shared between the three cells that meet there (first cell in cell order wins), edges
listed by lower cell then slot, and the 3-D Hilbert key of Skilling (2004) on 21-bit quantised positions. The
geometric kernels use the same small-triangle forms (chords in a tangent frame; Van Oosterom-Strackee areas), so the
geometry agrees to rounding. The Lloyd loop runs in C (`libinsilica_mesh.so`, csrc/mesh_lloyd.cpp).
"""
import ctypes
import math
from pathlib import Path

import numpy as np

from ..constants import R_PLANET

LLOYD_TOL = 1e-11
LLOYD_CAP = 200000

_LIB = None


def _mesh_lib():
    global _LIB
    if _LIB is None:
        path = Path(__file__).resolve().parents[1] / '_lib' / 'libinsilica_mesh.so'
        if not path.exists():
            raise RuntimeError(f'{path} is missing: build the native libraries first (python __graft_entry__.py build)')
        lib = ctypes.CDLL(str(path))
        lib.isw_mesh_lloyd.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_double, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]
        lib.isw_mesh_lloyd.restype = ctypes.c_int
        _LIB = lib
    return _LIB


# ---- vector helpers on (..., 3) arrays, written component by component so each product rounds on its own ----

def vnorm(v):
    return np.sqrt(v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1] + v[..., 2] * v[..., 2])


def vdot(a, b):
    return a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]


def vcross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], axis=-1)


def tangent_pair(a):
    """(e1, e2) orthonormal at unit vectors a: e1 = a x (the axis least aligned with a), e2 = a x e1."""
    k = np.argmin(np.abs(a), axis=-1)
    axis = (np.arange(3) == k[..., None]).astype(a.dtype)
    e1 = vcross(a, axis)
    e1 = e1 / vnorm(e1)[..., None]
    return e1, vcross(a, e1)


def arc(a, b):
    """Great-circle angle between unit vectors from their chord, 2 asin(|b - a| / 2)."""
    return 2 * np.arcsin(np.minimum(1.0, 0.5 * vnorm(b - a)))


def tri_area(a, b, c):
    """Spherical triangle area (unit sphere): tan(E/2) = |a.(b x c)| / (1 + a.b + b.c + c.a), the triple product taken
    from the chords in a's tangent frame and the denominator as 4 - (|b-a|^2 + |c-b|^2 + |a-c|^2)/2."""
    u, v = b - a, c - a
    e1, e2 = tangent_pair(a)
    trip = vdot(u, e1) * vdot(v, e2) - vdot(u, e2) * vdot(v, e1)
    w = c - b
    den = 4 - 0.5 * (vdot(u, u) + vdot(w, w) + vdot(v, v))
    return 2 * np.arctan2(np.abs(trip), den)


def circumcentre(a, b, c):
    """Unit circumcentre of spherical triangles: the plane normal of (a, b, c) built from chords in a's frame,
    a + (w1 e1 + w2 e2) / w3, normalised. Degenerate (pad) triangles give NaN; callers mask them."""
    e1, e2 = tangent_pair(a)
    u, v = b - a, c - a
    u1, u2, u3 = vdot(u, e1), vdot(u, e2), -0.5 * vdot(u, u)
    v1, v2, v3 = vdot(v, e1), vdot(v, e2), -0.5 * vdot(v, v)
    w1 = u2 * v3 - u3 * v2
    w2 = u3 * v1 - u1 * v3
    w3 = u1 * v2 - u2 * v1
    p = a + (w1 / w3)[..., None] * e1 + (w2 / w3)[..., None] * e2
    return p / vnorm(p)[..., None]


# ---- topology ----

def icosphere(n):
    """Unit icosahedron (a vertex at each pole, two rings of five at latitude +-atan(1/2)) subdivided n times by edge
    midpoints; new points are numbered as the faces are walked. Returns (positions (N, 3), faces (F, 3))."""
    pts = []

    def put(x, y, z):
        r = (x * x + y * y + z * z) ** 0.5
        pts.append((x / r, y / r, z / r))
        return len(pts) - 1

    put(0.0, 0.0, 1.0)
    lat = np.arctan(0.5)
    for k in range(5):
        lo = k * 2 * np.pi / 5
        put(np.cos(lat) * np.cos(lo), np.cos(lat) * np.sin(lo), np.sin(lat))
    for k in range(5):
        lo = (k + 0.5) * 2 * np.pi / 5
        put(np.cos(lat) * np.cos(lo), np.cos(lat) * np.sin(lo), -np.sin(lat))
    put(0.0, 0.0, -1.0)
    faces = []
    for k in range(5):
        a, b = 1 + k, 1 + (k + 1) % 5
        c, d = 6 + k, 6 + (k + 1) % 5
        faces.extend([(0, a, b), (a, c, b), (b, c, d), (11, d, c)])
    for _ in range(n):
        cache = {}

        def mid(i, j):
            key = (i, j) if i < j else (j, i)
            m = cache.get(key)
            if m is None:
                p, q = pts[i], pts[j]
                m = cache[key] = put(p[0] + q[0], p[1] + q[1], p[2] + q[2])
            return m

        sub = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            sub.extend([(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)])
        faces = sub
    return np.array(pts, dtype=np.float64), np.array(faces, dtype=np.int64)


def adjacency(faces, N):
    """Neighbour lists from the faces, each cell's neighbours ascending (pads -1), and the degrees."""
    a, b, c = faces.T
    pairs = np.unique(np.concatenate([a, a, b, b, c, c]) * N + np.concatenate([b, c, a, c, a, b]))
    src, dst = pairs // N, pairs % N
    deg = np.bincount(src, minlength=N)
    first = np.concatenate([[0], np.cumsum(deg)[:-1]])
    adj = np.full((N, 6), -1, dtype=np.int64)
    adj[src, np.arange(len(src)) - first[src]] = dst
    return adj, deg


def local_frame(P):
    """Local east and north at unit vectors P; east is (0, 1, 0) exactly at the poles."""
    ex, ey = -P[:, 1], P[:, 0].copy()
    er = np.sqrt(ex * ex + ey * ey)
    pole = er < 1e-12
    ex, ey, er = np.where(pole, 0.0, ex), np.where(pole, 1.0, ey), np.where(pole, 1.0, er)
    east = np.stack([ex / er, ey / er, np.zeros_like(ex)], axis=1)
    return east, vcross(P, east)


def ring_sort(P, adj):
    """Each cell's neighbours counter-clockwise by angle from local east (stable; pads last)."""
    east, north = local_frame(P)
    Q = P[np.maximum(adj, 0)]
    ang = np.arctan2(vdot(Q, north[:, None, :]), vdot(Q, east[:, None, :]))
    ang = np.where(adj < 0, np.inf, ang)
    return np.take_along_axis(adj, np.argsort(ang, axis=1, kind='stable'), axis=1)


def slot_next(deg):
    k, m = np.arange(6)[None, :], deg[:, None]
    return np.where(k < m, (k + 1) % m, k)


def slot_prev(deg):
    k, m = np.arange(6)[None, :], deg[:, None]
    return np.where(k < m, (k + m - 1) % m, k)


def lloyd(P, ring, deg, max_iters=LLOYD_CAP, tol=LLOYD_TOL):
    """Jacobi Lloyd toward the SCVT in C. Returns (positions, iterations, last move in radians)."""
    pos = np.ascontiguousarray(P, dtype=np.float64).copy()
    rg = np.ascontiguousarray(np.where(ring < 0, -1, ring), dtype=np.int32)
    dg = np.ascontiguousarray(deg, dtype=np.int32)
    it, mv = ctypes.c_int(0), ctypes.c_double(0.0)
    rc = _mesh_lib().isw_mesh_lloyd(len(pos), pos.ctypes.data, rg.ctypes.data, dg.ctypes.data, int(max_iters),
                                    float(tol), ctypes.byref(it), ctypes.byref(mv))
    if rc != 0:
        raise RuntimeError(f'isw_mesh_lloyd failed ({rc})')
    return pos, it.value, mv.value


def edge_list(nbr, deg):
    """Edges (i < j) listed by their lower cell i, then i's slot; with each side's slot."""
    N = nbr.shape[0]
    valid = np.arange(6)[None, :] < deg.astype(np.int64)[:, None]
    ii, kk = np.nonzero(valid & (nbr > np.arange(N)[:, None]))
    jj = nbr[ii, kk]
    kj = np.argmax(nbr[jj] == ii[:, None], axis=1)
    return dict(E=len(ii), eI=ii.astype(np.int32), eJ=jj.astype(np.int32), eKi=kk.astype(np.uint8), eKj=kj.astype(np.uint8))


@np.errstate(invalid='ignore', divide='ignore')
def build(n, max_iters=LLOYD_CAP):
    """Level n in the icosphere's own (v1) numbering: generators, rings, Voronoi corners, lengths and areas (m, m^2 on
    the Earth-sized sphere), edge list."""
    P0, faces = icosphere(n)
    N = len(P0)
    adj, deg = adjacency(faces, N)
    pos, iters, moved = lloyd(P0, ring_sort(P0, adj), deg, max_iters)
    lat = np.arcsin(np.clip(pos[:, 2], -1, 1))
    lon = np.arctan2(pos[:, 1], pos[:, 0])
    east, north = local_frame(pos)
    valid = np.arange(6)[None, :] < deg[:, None]
    nbr = np.where(valid, ring_sort(pos, adj), -1)
    nb = np.where(valid, nbr, nbr[:, :1])
    nxt, prv = slot_next(deg), slot_prev(deg)
    nb1 = np.take_along_axis(nb, nxt, axis=1)
    Pc = np.broadcast_to(pos[:, None, :], (N, 6, 3))
    corners = np.where(valid[..., None], circumcentre(Pc, pos[nb], pos[nb1]), 0.0)
    # one copy of each corner for its three cells: the first cell (in cell order) that meets it
    tri = np.sort(np.stack([np.broadcast_to(np.arange(N)[:, None], (N, 6)), nb, nb1], -1), -1)
    key = ((tri[..., 0] * N + tri[..., 1]) * N + tri[..., 2]).ravel()
    live = np.flatnonzero(valid.ravel())
    _, first, inv = np.unique(key[live], return_index=True, return_inverse=True)
    flat = corners.reshape(-1, 3)
    flat[live] = flat[live[first[inv]]]
    corners = flat.reshape(N, 6, 3)
    prev_c = np.take_along_axis(corners, prv[..., None], axis=1)
    dist = np.where(valid, R_PLANET * arc(Pc, pos[nb]), 0.0)
    edge_len = np.where(valid, R_PLANET * arc(prev_c, corners), 0.0)
    next_c = np.take_along_axis(corners, nxt[..., None], axis=1)
    area = np.zeros(N)
    for k in range(6):
        area = area + np.where(valid[:, k], tri_area(pos, corners[:, k], next_c[:, k]), 0.0)
    area = area * R_PLANET * R_PLANET
    g = dict(n=n, N=N, pos=pos, lat=lat, lon=lon, area=area, east=east, north=north, nbr=nbr.astype(np.int32),
             deg=deg.astype(np.uint8), dist=dist, edgeLen=edge_len, corners=corners, totalArea=float(np.sum(area)),
             lloydIters=iters, lloydMoved=moved)
    g.update(edge_list(g['nbr'], g['deg']))
    return g


CELL_KEYS = ('pos', 'lat', 'lon', 'area', 'east', 'north', 'deg', 'dist', 'edgeLen', 'corners')


def hilbert_perm(pos, bits=21):
    """Cell order along a 3-D Hilbert curve through the quantised unit positions (Skilling 2004: axes to the
    transposed Hilbert index, then its bits interleaved into one key)."""
    q = np.clip(((pos + 1) * 0.5 * (2 ** bits - 1)).astype(np.uint64), 0, 2 ** bits - 1)
    X = [q[:, 0].copy(), q[:, 1].copy(), q[:, 2].copy()]
    Q = 1 << (bits - 1)
    while Q > 1:
        low, bit = np.uint64(Q - 1), np.uint64(Q)
        for i in range(3):
            on = (X[i] & bit) != 0
            t = (X[0] ^ X[i]) & low
            x0 = np.where(on, X[0] ^ low, X[0] ^ t)
            xi = np.where(on, X[i], X[i] ^ t)
            X[0] = x0
            if i:
                X[i] = xi
        Q >>= 1
    X[1] = X[1] ^ X[0]
    X[2] = X[2] ^ X[1]
    t = np.zeros_like(X[0])
    Q = 1 << (bits - 1)
    while Q > 1:
        t = np.where((X[2] & np.uint64(Q)) != 0, t ^ np.uint64(Q - 1), t)
        Q >>= 1
    X = [x ^ t for x in X]
    key = np.zeros_like(X[0])
    one = np.uint64(1)
    for b in range(bits - 1, -1, -1):
        for i in range(3):
            key = (key << one) | ((X[i] >> np.uint64(b)) & one)
    return np.argsort(key, kind='stable')


def renumber(g, perm):
    """The same mesh with new cell k = old cell perm[k]; edges re-listed by the usual rule."""
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    out = dict(g)
    for k in CELL_KEYS:
        out[k] = g[k][perm]
    nb = g['nbr'][perm]
    out['nbr'] = np.where(nb >= 0, inv[np.maximum(nb, 0)], -1).astype(np.int32)
    out.update(edge_list(out['nbr'], out['deg']))
    out['perm'] = perm.astype(np.int32)
    return out
