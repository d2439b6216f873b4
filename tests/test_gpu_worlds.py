"""Whole worlds on the GPU against the oracle world, and their save files (GPU).

Bars, written here:
  * FP64 form against the oracle: <= 1e-12 relative (max |diff| / max |oracle|) per field after the first tick, and
    <= 1e-10 after a model day (device pow/exp/sqrt round once differently from glibc's, and a day of dynamics grows
    that); every budget <= 1e-13 relative;
  * the floor limiter: every thickness >= floor - 1e-9 with the lower layer held at the floor somewhere;
  * save files: the same state saves to the same bytes, load then save gives the same bytes, and a loaded world
    continues bit for bit as the saved one does;
  * the mixed form at level 6 after one model day of the coupled stack: relative L2 error of each air layer's
    thickness <= 1e-5 against the FP64 form.
"""
import numpy as np
import pytest
import torch

from conftest import relmax

pytestmark = pytest.mark.gpu

TOL1, TOLD = 1e-12, 1e-10


def _np(t):
    return t.detach().cpu().numpy()


def _compare(w, o, tol, when):
    st = w.state_arrays()
    worst = 0.0
    for k, r in o.state().items():
        if k in st:
            e = relmax(_np(st[k]), r)
            worst = max(worst, e)
            assert e <= tol, f'{when}: {k} {e:.2e}'
    return worst


def test_jet_level4_day():
    """C1: one 8 km layer in the balanced jet with three bumps, level 4, dt 600 s, 144 steps (24 h), no physics."""
    from insilica.world import World
    from oracle.world import OracleWorld
    print('case: world jet n=4 144 steps', flush=True)
    w = World(4, 'jet', precision='double')
    o = OracleWorld(w)
    m0 = w.totals()
    w.step(1, graph=False)
    o.step(1)
    _compare(w, o, TOL1, 'step 1')
    w.step(143)
    o.step(143)
    e = _compare(w, o, TOLD, 'step 144')
    m1 = w.totals()
    print(f'jet: worst field error after 144 steps {e:.2e}', flush=True)
    assert abs(m1['M0'] - m0['M0']) / m0['M0'] <= 1e-13
    assert abs(m1['tracer0'] - m0['tracer0']) / m0['tracer0'] <= 1e-13


def test_outcrop_floor_limiter():
    """Two layers over a ridge higher than the lower layer is deep (the reference's 'outcrop'): the floor limiter is
    active, and the FP64 world matches the oracle's with every thickness at or above the floor."""
    from insilica.world import World
    from oracle.world import OracleWorld
    print('case: world outcrop n=4 1 day', flush=True)
    w = World(4, 'outcrop', precision='double')
    o = OracleWorld(w)
    t0 = w.totals()
    w.step(1, graph=False)
    o.step(1)
    _compare(w, o, TOL1, 'tick 1')
    low, held = np.inf, 0
    for _ in range(23):
        w.step(1)
        o.step(1)
        h = _np(w.h)
        low = min(low, float(h.min()))
        held = max(held, int(np.sum(h[1] < w.core.floor + 1e-6)))
    e = _compare(w, o, TOLD, 'day 1')
    t1 = w.totals()
    print(f'outcrop: thinnest {low:.6f} m, {held} cells at the floor, worst error {e:.2e}', flush=True)
    assert low >= w.core.floor - 1e-9 and held > 0
    for k in ('M0', 'M1', 'tracer0', 'tracer1'):
        assert abs(t1[k] - t0[k]) / abs(t0[k]) <= 1e-13, k


def test_coupled_with_mixing():
    """The reference's 'coupled' world (moist air over the three-layer ocean): a model day at level 4 with the
    ocean's hourly steps, its forcing from the air column every tick and the mixed layer's daily mixing, against the
    oracle world; the ocean's volume kept, and its salt changed exactly by the virtual flux of the fresh water,
    -S_ref fw / rho_w with fw = (rain - evaporation) dt - the sea ice's growth (reference: tests/test_world.py:
    120-173, salt against fresh water <= 1e-13)."""
    from insilica.world import World
    from oracle.world import OracleWorld
    print('case: world coupled n=4 1 day', flush=True)
    w = World(4, 'coupled', precision='double')
    o = OracleWorld(w)
    day = int(round(86400 / w.dt))
    assert w._kind(day - 1) & 4 and not w._kind(day - 2) & 4
    w.step(1, graph=False)
    o.step(1)
    _compare(w, o, TOL1, 'tick 1')
    A = w.grid.areaCell
    hS = lambda: float(np.sum(A * np.sum(_np(w.sea.h) * _np(w.sea.S), axis=0)))
    s0, v0, ice0 = hS(), float(np.sum(A * _np(w.sea.h))), _np(w.ice).copy()
    fw = 0.0
    for _ in range(day - 1):
        w.step(1)
        fw += float(np.sum(A * (_np(w.physics.rain) - _np(w.physics.evap)))) * w.dt
    o.step(day - 1)
    e = _compare(w, o, TOLD, 'day 1')
    fw -= float(np.sum(A * (_np(w.ice) - ice0)))
    wx = _np(w.sea.wx)
    c = w.sea.cfg
    drift = abs((hS() - s0) + c['S_ref'] * fw / c['rho']) / s0
    print(f'coupled: worst error after a day {e:.2e}; mixing moved up to {np.abs(wx).max():.1f} m; salt against '
          f'fresh water {drift:.2e}', flush=True)
    assert np.abs(wx).max() > 0.1
    assert abs(float(np.sum(A * _np(w.sea.h))) - v0) / v0 <= 1e-13
    assert drift <= 1e-13


SAVE_CASES = {'jet': ('jet', None), 'coupled': ('coupled', None),
              # (the coupled world with both drags over seeded orography and the tides over a seeded rough floor:
              # their settings, arrays included, go into the file and rebuild the same world)
              'coupled+drags+tides': ('coupled', {'physics': {'oro': 1.0, 'gw': 1.0}, 'tides': True,
                                                  'sso': 'synthetic',
                                                  'rough': [float(x) for x in np.random.default_rng(9).uniform(
                                                      0, 4e4, 2562)]})}


@pytest.mark.parametrize('case', sorted(SAVE_CASES))
def test_save_load_continues_identically(tmp_path, case):
    from insilica.world import World
    preset, params = SAVE_CASES[case]
    print(f'case: save/load {case}', flush=True)
    w = World(4, preset, precision='mixed', params=params)
    w.step(7)
    a, b = tmp_path / 'a.iw2', tmp_path / 'b.iw2'
    w.save(a)
    w.save(b)
    assert a.read_bytes() == b.read_bytes()
    r = World.load(a)
    r.save(b)
    assert a.read_bytes() == b.read_bytes()
    assert r.tick == w.tick and r.state_hash() == w.state_hash()
    w.step(50)
    r.step(50)
    assert r.state_hash() == w.state_hash()


def test_mixed_level6_one_day():
    """The mixed form against the FP64 form, coupled stack at level 6, one model day (192 ticks)."""
    from insilica.world import World
    print('case: mixed vs fp64 coupled_stack n=6 1 day', flush=True)
    a = World(6, 'coupled_stack', precision='double', params=dict(nfast=15))
    b = World(6, 'coupled_stack', precision='mixed', params=dict(nfast=15))
    ticks = int(round(86400 / a.dt))
    a.step(ticks)
    b.step(ticks)
    A = a.grid.areaCell
    ha, hb = _np(a.h), _np(b.h)
    errs = [float(np.sqrt(np.sum(A * (hb[k] - ha[k]) ** 2) / np.sum(A * ha[k] ** 2))) for k in range(3)]
    print('level 6 mixed h L2 errors per air layer', errs, flush=True)
    assert max(errs) <= 1e-5
