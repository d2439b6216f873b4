"""The land's water routing and ice flow on the GPU (isw_land_route, isw_ice_flow through `LandWater`) against the
oracle, on seeded terrain, ponds, soils and ice sheets.

Bar: bit-identical. Both sides compute each edge's flows with the same operations in the same order (contraction off)
and sum a cell's edges in its own edge order, so every pond, soil, ice, mouth and calved value matches to the bit after
every step. At level 7 (163842 cells) the water and ice budgets also close to 1e-13.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GAM_PER_S = 2.0 * 1.0e-16 / (365.0 * 86400.0) * (917.0 * 9.81) ** 3 / 5.0


def _world(g, seed):
    rng = np.random.default_rng(seed)
    z = np.asarray(g.pos)[:, 2]
    b = 800.0 * np.maximum(0.0, np.sin(3 * np.arcsin(z)) + 0.3) + rng.uniform(0, 50, g.N)
    land = (b > 300.0).astype(float)
    soilcap = 100.0 + 0.5 * b
    pond = land * rng.uniform(0, 500, g.N)
    soil = land * soilcap * rng.uniform(0, 1, g.N)
    ice = land * np.where(rng.uniform(size=g.N) < 0.3, rng.uniform(0, 3e5, g.N), 0.0)
    return b, land, soilcap, pond, soil, ice


def _run(grids, n, steps, seed, with_prev):
    import oracle
    from oracle import land as oland
    from insilica.gpu import DeviceGrid
    from insilica.land import LandWater, RIVER_SPEED, K_SEEP, RHO_W
    print(f'case: land water n={n} steps={steps} seed={seed} ice_prev={with_prev}', flush=True)
    g = grids(n)
    dg, og = DeviceGrid(g, 'double'), oracle.Grid(g)
    b, land, soilcap, pond, soil, ice = _world(g, seed)
    T = lambda a: torch.tensor(a, dtype=torch.float64, device=dg.dev)
    lw = LandWater(dg, T(b), T(land), T(soilcap), T(pond), T(soil), T(ice))
    if with_prev:
        lw.ice_prev = T(np.zeros(g.N))
    mouth, calved = np.zeros(g.N), np.zeros(g.N)
    prev = np.zeros(g.N) if with_prev else None
    dt = 1800.0
    for k in range(steps):
        lw.step(dt)
        pond, soil, mouth, _ = oland.route(og, b, land, soilcap, pond, soil, ice, RHO_W, 917.0, RIVER_SPEED * dt,
                                           K_SEEP * dt, mouth)
        ice, calved, prev, _ = oland.ice_flow(og, b, land, ice, 917.0, lw.gam * dt, calved, ice_prev=prev)
        torch.cuda.synchronize()
        got = dict(pond=lw.pond, soil=lw.soil, ice=lw.ice, mouth=lw.mouth, calved=lw.calved)
        want = dict(pond=pond, soil=soil, ice=ice, mouth=mouth, calved=calved)
        if with_prev:
            got['ice_prev'], want['ice_prev'] = lw.ice_prev, prev
        for key in got:
            a = got[key].cpu().numpy()
            bad = np.flatnonzero(a != want[key])
            assert bad.size == 0, (k, key, bad[:5], a[bad[:5]], want[key][bad[:5]])
    return g, lw, (b, land, soilcap, pond, soil, ice, mouth, calved)


@pytest.mark.parametrize('n,steps,seed', [(4, 6, 1), (6, 4, 2)])
def test_land_water_bitwise(grids, n, steps, seed):
    _run(grids, n, steps, seed, with_prev=(n == 6))


def test_land_water_budgets_level7(grids):
    """Level 7: the GPU's pond + soil + mouth and land ice + calved after 8 steps hold what they started with."""
    from insilica.gpu import DeviceGrid
    from insilica.land import LandWater
    print('case: land water budgets n=7', flush=True)
    g = grids(7)
    dg = DeviceGrid(g, 'double')
    b, land, soilcap, pond, soil, ice = _world(g, 7)
    T = lambda a: torch.tensor(a, dtype=torch.float64, device=dg.dev)
    lw = LandWater(dg, T(b), T(land), T(soilcap), T(pond), T(soil), T(ice))
    A = g.areaCell
    w0 = np.sum(A * (pond + soil))
    i0 = np.sum(A * ice)
    for _ in range(8):
        lw.step(3600.0)
    torch.cuda.synchronize()
    p, s, m = lw.pond.cpu().numpy(), lw.soil.cpu().numpy(), lw.mouth.cpu().numpy()
    ic, cv = lw.ice.cpu().numpy(), lw.calved.cpu().numpy()
    w1 = np.sum(A * (p + s)) + np.sum(A * m)
    i1 = np.sum(A * ic * land) + np.sum(A * cv)
    assert m.max() > 0 and np.abs(p - pond).max() > 0
    assert abs(w1 - w0) <= 1e-13 * w0 and abs(i1 - i0) <= 1e-13 * i0
