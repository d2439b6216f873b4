"""The tides on the GPU (isw_tide_pot, isw_tide_itd, isw_tide_mix through `Tides`, and the tide's hooks in the
layers' step) against the oracle.

Bar: FP64 parity <= 1e-12 relative per field for the kernels (the bodies' places go through sin, cos, atan2, asin and
Kepler's Newton steps, which the device math library and glibc may round an ulp apart); a coupled world with the
tides (the potential on every ocean layer, the bottom's quadratic and internal-tide drag, the mixing) against the
oracle's world <= 1e-10 after four ocean steps; the mixing keeps each column's heat and salt to 1e-13.
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def rel(a, b):
    b = np.asarray(b)
    return float(np.max(np.abs(np.asarray(a) - b)) / max(float(np.max(np.abs(b))), 1e-300))


def _dev(dg, a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dg.dev)


@pytest.mark.parametrize('n', [4, 6])
def test_tide_kernels_match_oracle(grids, n):
    import oracle
    from oracle import tides as ot
    from insilica import _hip
    from insilica.atmosphere import DEFAULTS
    from insilica.gpu import DeviceGrid
    from insilica.tides import settings, sky, sky_array
    print(f'case: tide kernels n={n}', flush=True)
    g = grids(n)
    dg, og = DeviceGrid(g, 'double'), oracle.Grid(g)
    p = dict(DEFAULTS, ecc=0.0167, lam_peri=math.radians(283.0))
    s = sky(p, 6.371e6, {'moon': {'L0': 30.0, 'node0': 125.0, 'aps0': 80.0}})
    tp = sky_array(s)
    pos = np.ascontiguousarray(g.pos, dtype=float)
    V = _dev(dg, np.zeros(g.N))
    for t in (0.0, 3.7e6, 2.9e7):
        time = _dev(dg, [t])
        _hip.call('isw_tide_pot', dg.ref, _hip.ctypes.byref(s), _hip.ptr(time), 1800.0, _hip.ptr(_dev(dg, pos)),
                  _hip.ptr(V), _hip.stream())
        want = np.zeros(g.N)
        ot.tide_pot(og, tp, t, 1800.0, pos, want)
        assert rel(V.cpu().numpy(), want) <= 1e-12, t
    # the internal tide's drag and the mixing on a seeded two-layer ocean
    from insilica.ocean import OCEAN, ocean_struct
    rng = np.random.default_rng(3)
    o = ocean_struct(OCEAN, 9.80616)
    eos = (9.80616, OCEAN['alpha'], OCEAN['alpha2'], OCEAN['beta'])
    h = np.stack([rng.uniform(50, 150, g.N), rng.uniform(3000, 4000, g.N)])
    T = np.stack([rng.uniform(285, 300, g.N), rng.uniform(274, 280, g.N)])
    S = np.stack([rng.uniform(34.5, 35, g.N), rng.uniform(34.8, 35.2, g.N)])
    u = rng.uniform(-0.5, 0.5, (2, g.E))
    rough = rng.uniform(0, 2e4, g.N)
    land = (rng.uniform(size=g.N) < 0.2).astype(float)
    cfg = settings(True)
    cb = _dev(dg, np.zeros(g.E))
    Hd, Td, Sd = _dev(dg, h), _dev(dg, T), _dev(dg, S)
    _hip.call('isw_tide_itd', dg.ref, 2, _hip.ctypes.byref(o), _hip.ptr(Hd), _hip.ptr(Td), _hip.ptr(Sd),
              _hip.ptr(_dev(dg, rough)), cfg['kappa_it'], None, _hip.ptr(cb), _hip.stream())
    cbw = np.zeros(g.E)
    ot.tide_itd(og, 2, h, T, S, rough, cfg['kappa_it'], eos, cbw)
    assert rel(cb.cpu().numpy(), cbw) <= 1e-12
    ds = _dev(dg, np.zeros(g.N))
    _hip.call('isw_tide_mix', dg.ref, 2, _hip.ctypes.byref(o), _hip.ptr(Hd), _hip.ptr(_dev(dg, u)), _hip.ptr(Td),
              _hip.ptr(Sd), _hip.ptr(_dev(dg, land)), 2.5e-4, 1025.0, 0.2, 3600.0, _hip.ptr(ds), cfg['cd'],
              _hip.ptr(cb), cfg['zeta'], _hip.stream())
    T1, S1, dsw = T.copy(), S.copy(), np.zeros(g.N)
    ot.tide_mix(og, 2, h, u, T1, S1, land, 2.5e-4, 1025.0, 0.2, 3600.0, eos, dsw, cfg['cd'], cbw, cfg['zeta'])
    assert np.abs(T1 - T).max() > 1e-6                               # (it mixed)
    assert rel(Td.cpu().numpy(), T1) <= 1e-12 and rel(Sd.cpu().numpy(), S1) <= 1e-12
    assert rel(ds.cpu().numpy(), dsw) <= 1e-12
    heat = lambda X: np.sum(h * X, axis=0)
    assert np.max(np.abs(heat(Td.cpu().numpy()) - heat(T)) / heat(T)) <= 1e-13
    assert np.max(np.abs(heat(Sd.cpu().numpy()) - heat(S)) / heat(S)) <= 1e-13


@pytest.mark.parametrize('rough', [False, True])
def test_coupled_world_with_tides_matches_oracle(grids, rough):
    """The reference's 'coupled' world with Earth's tides (and the internal tide's drag over a seeded rough floor):
    the FP64 form against the oracle's world for four ocean steps; the tides move the ocean."""
    from insilica.world import World
    from oracle.world import OracleWorld
    print(f'case: coupled world with tides n=4 rough={rough}', flush=True)
    g = grids(4)
    params = {'tides': True}
    if rough:
        params['rough'] = np.random.default_rng(5).uniform(0, 4e4, g.N)
    w = World(4, 'coupled', precision='double', params=params)
    assert w.tides is not None
    o = OracleWorld(w)
    ticks = 4 * w.ocean_every
    w.step(ticks, graph=False)
    o.step(ticks)
    st = w.state_arrays()
    worst = 0.0
    for k, ref in o.state().items():
        if k in st:
            worst = max(worst, rel(st[k].detach().cpu().numpy(), ref))
    assert worst <= 1e-10, worst
    v = w.tides.V.cpu().numpy()
    assert np.abs(v).max() > 0.1                                      # (the potential is set, ~1 m^2/s^2)
    ref = World(4, 'coupled', precision='double')
    ref.step(ticks, graph=False)
    assert not torch.equal(ref.sea.h, w.sea.h)                        # (and it moves the ocean)
    # a graph replays the same ticks
    w2 = World(4, 'coupled', precision='double', params=params)
    w2.step(ticks, graph=True)
    torch.cuda.synchronize()
    for k, a in w.state_arrays().items():
        assert torch.equal(a, w2.state_arrays()[k]), k
