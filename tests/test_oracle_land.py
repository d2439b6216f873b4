"""The oracle's land water routing and ice flow (CPU) against the rules the reference states (kernels/land.cu:1-11,
97-103; land.py): every kilogram kept (land's pond and soil plus what the sea cells' mouths received; the land's ice
plus what the sea cells calved), a level water surface and level water tables moving nothing, one wet cell among
dry neighbours on level ground giving each neighbour rho_w min(3/4 U d le, h / 12 A_min) (its depth d = h the head),
scaled so that it gives at most half its pond, seepage between two tables by Darcy's law rho_w K (le / de) dz within
its 1/12 cap, and ice flowing only downhill, faster as the ice thickens (Glen's H^5 |grad s|^2)."""
import numpy as np
import pytest

import oracle
from oracle import land as oland

RHO_W, RHO_ICE = 1000.0, 917.0
GAM = 2.0 * 1.0e-16 / (365.0 * 86400.0) * (RHO_ICE * 9.81) ** 3 / 5.0


def terrain(g, seed):
    rng = np.random.default_rng(seed)
    z = np.asarray(g.pos)[:, 2]
    b = 800.0 * np.maximum(0.0, np.sin(3 * np.arcsin(z)) + 0.3) + rng.uniform(0, 50, g.N)
    land = (b > 300.0).astype(float)
    soilcap = 100.0 + 0.5 * np.maximum(b, 0)
    return b, land, soilcap


@pytest.mark.parametrize('n', [4, 5])
def test_route_keeps_water_and_ice(grids, n):
    g = grids(n)
    og = oracle.Grid(g)
    b, land, soilcap = terrain(g, n)
    rng = np.random.default_rng(10 + n)
    pond = land * rng.uniform(0, 500, g.N)
    soil = land * soilcap * rng.uniform(0, 1, g.N)
    ice = land * np.where(rng.uniform(size=g.N) < 0.3, rng.uniform(0, 3e5, g.N), 0.0)
    mouth, calved = np.zeros(g.N), np.zeros(g.N)
    A = g.areaCell
    w0 = np.sum(A * (pond + soil)) + np.sum(A * mouth)
    i0 = np.sum(A * ice * land) + np.sum(A * calved)
    dt = 3600.0
    moved = 0.0
    for _ in range(24):
        p1, s1, mouth, _ = oland.route(og, b, land, soilcap, pond, soil, ice, RHO_W, RHO_ICE, 1.0 * dt, 2e4 / 3600 * dt,
                                       mouth)
        moved += np.abs(p1 - pond).sum()
        pond, soil = p1, s1
        ice, calved, _, _ = oland.ice_flow(og, b, land, ice, RHO_ICE, GAM * dt, calved)
    w1 = np.sum(A * (pond + soil)) + np.sum(A * mouth)
    i1 = np.sum(A * ice * land) + np.sum(A * calved)
    assert moved > 0 and mouth.max() > 0
    assert abs(w1 - w0) <= 1e-13 * w0 and abs(i1 - i0) <= 1e-13 * i0
    assert pond.min() >= -1e-9 and soil.min() >= -1e-9 and ice.min() >= -1e-6


def test_level_water_stays(grids):
    """All land, level ground, the same pond and soil everywhere: no head, no table difference, nothing moves."""
    g = grids(4)
    og = oracle.Grid(g)
    N = g.N
    b, land, soilcap = np.full(N, 500.0), np.ones(N), np.full(N, 300.0)
    pond, soil, ice = np.full(N, 120.0), np.full(N, 150.0), np.zeros(N)
    p1, s1, m1, _ = oland.route(og, b, land, soilcap, pond, soil, ice, RHO_W, RHO_ICE, 3600.0, 2e4, np.zeros(N))
    assert np.array_equal(p1, pond) and np.array_equal(s1, soil) and not np.any(m1)


def _edges(g, i):
    return [int(e) for e in g.cellEdge[i] if e >= 0]


def test_one_wet_cell_gives_by_the_rule(grids):
    """One wet cell on level dry land: through each edge rho_w min(3/4 U h le, h / 12 min(A_i, A_o)), h its depth;
    with a slow flow the first term (3/4 U h le), with a fast one the 1/12 cap; then at most half its pond in all."""
    g = grids(4)
    og = oracle.Grid(g)
    N = g.N
    b, land, soilcap = np.full(N, 500.0), np.ones(N), np.full(N, 300.0)
    soil, ice = np.zeros(N), np.zeros(N)
    i = 777
    h = 0.3
    pond = np.zeros(N)
    pond[i] = RHO_W * h
    A = g.areaCell
    for u_dt in (10.0, 1e7):                          # slow: 3/4 U h le; fast: the cap, then the half-pond limit
        p1, _, _, _ = oland.route(og, b, land, soilcap, pond, soil, ice, RHO_W, RHO_ICE, u_dt, 0.0, np.zeros(N))
        flows = {}
        for e in _edges(g, i):
            o = int(g.eJ[e] if g.eI[e] == i else g.eI[e])
            flows[o] = RHO_W * min(0.75 * u_dt * h * g.le[e], h / 12 * min(A[i], A[o]))
        out = sum(flows.values())
        s = min(1.0, 0.5 * pond[i] * A[i] / out)
        for o, f in flows.items():
            assert abs(p1[o] - f * s / A[o]) <= 1e-12 * p1[o], (u_dt, o)
        assert abs(p1[i] - (pond[i] - out * s / A[i])) <= 1e-12 * pond[i]
        assert out * s <= 0.5 * pond[i] * A[i] * (1 + 1e-12)
        if u_dt > 1e6:                                # every edge at its cap h / 12 min(A_i, A_o)
            for o, f in flows.items():
                assert f == RHO_W * (h / 12 * min(A[i], A[o]))


def test_seepage_by_darcy(grids):
    """Two water tables on level land: the wetter cell's soil seeps rho_w K (le / de) dz into each drier neighbour,
    dz = (soil / soilcap) b less the neighbour's, within the cap dz / (r_i + r_o) / 12 min(A) (r: table per kg)."""
    g = grids(4)
    og = oracle.Grid(g)
    N = g.N
    b, land, soilcap = np.full(N, 500.0), np.ones(N), np.full(N, 300.0)
    soil = np.full(N, 60.0)
    i = 1234
    soil[i] = 240.0
    ks = 2e4 / 3600 * 60.0
    _, s1, _, _ = oland.route(og, b, land, soilcap, np.zeros(N), soil, np.zeros(N), RHO_W, RHO_ICE, 0.0, ks,
                              np.zeros(N))
    A = g.areaCell
    zt = soil / soilcap * b
    for e in _edges(g, i):
        o = int(g.eJ[e] if g.eI[e] == i else g.eI[e])
        dz = zt[i] - zt[o]
        ra, rb = zt[i] / soil[i], zt[o] / soil[o]
        f = min(RHO_W * ks * g.le[e] / g.de[e] * dz, dz / (ra + rb) / 12 * min(A[i], A[o]))
        assert abs((s1[o] - soil[o]) * A[o] - f) <= 1e-10 * f, o


def test_ice_flows_downhill_and_thick_ice_faster(grids):
    """A dome of ice on level land spreads to its neighbours, none flows up; twice as thick a dome sends 2^8 = 256
    times as much through each edge while the 1/12 cap does not bind (Glen's H^5, times |grad s|^2's 4 and the
    head's own 2)."""
    g = grids(4)
    og = oracle.Grid(g)
    N = g.N
    b, land = np.full(N, 500.0), np.ones(N)
    i = 321
    res = []
    for H in (200.0, 400.0):
        ice = np.zeros(N)
        ice[i] = RHO_ICE * H
        i1, calved, _, _ = oland.ice_flow(og, b, land, ice, RHO_ICE, GAM * 3600.0, np.zeros(N))
        res.append(i1)
        gained = np.flatnonzero(i1 > 0)
        assert set(gained) >= {i} and np.all(i1 >= 0) and not np.any(calved)
        nb = {int(g.eJ[e] if g.eI[e] == i else g.eI[e]) for e in _edges(g, i)}
        assert set(gained) - {i} == nb
    ratio = [res[1][o] / res[0][o] for o in np.flatnonzero(res[0] > 0) if o != i]
    np.testing.assert_allclose(ratio, 256.0, rtol=1e-12)
