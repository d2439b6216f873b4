"""The Held-Suarez column and the surface-exchange kernels on the GPU against the CPU oracle, through the C ABI.

FP64 in both forms; bar: <= 1e-12 relative (max |diff| / max |oracle|) on every output (one rounding of the device's
pow/exp/expm1/log/cbrt against glibc's is the only difference left), the crossing's masses to the same bar.
"""
import ctypes

import numpy as np
import pytest
import torch

from conftest import relmax

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device='cuda')


def _np(t):
    return t.detach().cpu().numpy().copy()


def par(L, **kw):
    from insilica.atmosphere import par_array
    stars = list(np.linspace(360.0, 300.0, L - 1)) if L > 2 else [300.0]
    return par_array(L, dict(kw, theta_stars=stars))


@pytest.mark.parametrize('L', [2, 3, 5])
def test_hs_column(L):
    from insilica import _hip
    from oracle import column as ocol
    print(f'case: hs_column L={L} N=40962', flush=True)
    rng = np.random.default_rng(L)
    N = 40962
    mu = rng.uniform(1500.0, 4000.0, (L, N))
    th = np.sort(rng.uniform(270.0, 360.0, (L, N)), axis=0)[::-1].copy()
    lat = rng.uniform(-1.5, 1.5, N)
    p = par(L, bl_mix=1.0 if L >= 3 else 0.0, bl_cap=0.5)
    hs = np.array([1 / 86400.0, 1 / 7200.0])
    dt = 1800.0
    t_th, t_wx, t_q, t_s = _dev(th), torch.zeros(L - 1, N, dtype=torch.float64, device='cuda'), \
        torch.zeros(N, dtype=torch.float64, device='cuda'), torch.zeros(N, dtype=torch.float64, device='cuda')
    t_mu, t_lat = _dev(mu), _dev(lat)
    _hip.call('isw_hs_column', N, L, p.ctypes.data_as(ctypes.c_void_p), hs.ctypes.data_as(ctypes.c_void_p),
              _hip.ptr(t_lat), _hip.ptr(t_mu), _hip.ptr(t_th), dt, _hip.ptr(t_wx), _hip.ptr(t_q), _hip.ptr(t_s),
              _hip.stream())
    torch.cuda.synchronize()
    o_th, o_wx, o_q, o_s = ocol.hs_column(p, hs, lat, mu, th, dt)
    e = [relmax(_np(t_th), o_th), relmax(_np(t_wx), o_wx[:L - 1]), relmax(_np(t_q), o_q), relmax(_np(t_s), o_s)]
    print('hs_column errors', e, 'moved', int(np.count_nonzero(o_wx)), flush=True)
    assert np.count_nonzero(o_wx) > 0
    assert max(e) <= TOL


def test_surface_layer():
    import ctypes as C
    from insilica import _hip
    from insilica.radiation import _air
    from oracle import column as ocol
    print('case: surface_layer N=40962 L=3', flush=True)
    rng = np.random.default_rng(7)
    L, N = 3, 40962
    mu = rng.uniform(2000.0, 4500.0, (L, N))
    ps = 9.80616 * mu.sum(0)
    Ta = rng.uniform(240.0, 305.0, N)
    theta = np.tile(Ta * (1e5 / ps) ** (287.04 / 1004.6), (L, 1)) + rng.normal(0, 0.1, (L, N))
    qv = rng.uniform(0, 0.015, (L, N))
    Ts = Ta + rng.normal(0, 4.0, N)
    ice = np.where(rng.uniform(size=N) < 0.2, rng.uniform(0, 100, N), 0.0)
    land = (rng.uniform(size=N) < 0.3).astype(float)
    lc = rng.uniform(0.01, 1.0, N)
    uabs = rng.uniform(0, 25, N)
    p = par(L, moist=1.0, sfc=1.0, q_sea=0.98)
    cd, ch = torch.zeros(N, dtype=torch.float64, device='cuda'), torch.zeros(N, dtype=torch.float64, device='cuda')
    args = [_dev(x) for x in (uabs, mu, theta, qv, Ts, ice, land, lc)]
    _hip.call('isw_surface_layer', N, L, p.ctypes.data_as(C.c_void_p), C.byref(_air(None)), *(_hip.ptr(a) for a in args[:7]),
              _hip.ptr(cd), _hip.ptr(ch), _hip.ptr(args[7]), _hip.stream())
    torch.cuda.synchronize()
    ocd, och = ocol.surface_layer(p, uabs, mu, theta, Ts, qv=qv, ice=ice, land=land, lc=lc)
    e = (relmax(_np(cd), ocd), relmax(_np(ch), och))
    print('surface_layer errors', e, flush=True)
    assert max(e) <= TOL


def test_cell_speeds_and_clock(grids):
    from insilica import _hip
    from insilica.atmosphere import AirPhysics, perot_operators
    from insilica.gpu import DeviceGrid
    from oracle import column as ocol
    print('case: cell speeds level 6, advance_time, gas_src', flush=True)
    g = grids(6)
    dg = DeviceGrid(g, precision='double')
    rng = np.random.default_rng(3)
    u = rng.normal(0, 10, (2, g.E))
    ap = AirPhysics(dg, 2, params=dict(sfc_wind=1.0), forcing='held_suarez')
    ap._speed(_dev(u))
    torch.cuda.synchronize()
    assert relmax(_np(ap.uabs), ocol.cell_speed_perot(g, u[1], perot_operators(g))) <= TOL
    ap.perot = None
    ap._speed(_dev(u))
    torch.cuda.synchronize()
    assert relmax(_np(ap.uabs), ocol.cell_speed(g, u[1])) <= TOL
    t = _dev([100.0])
    for _ in range(3):
        AirPhysics.advance_time(t, 450.0)
    torch.cuda.synchronize()
    assert float(t[0]) == 1450.0
    from insilica.atmosphere import gas_sources
    L, G = 3, 3
    src, h, x = rng.uniform(0, 1e-6, (G, g.N)), rng.uniform(1e3, 5e3, (L, g.N)), rng.uniform(0, 1e-3, (G * L, g.N))
    tx = _dev(x)
    gas_sources(_dev(src), _dev(h), 450.0, tx, L)
    torch.cuda.synchronize()
    assert np.array_equal(_np(tx), ocol.gas_src(src, h, 450.0, x, L))


def test_held_suarez_world():
    """The 'held_suarez' world (two air layers, the column and the exchange every tick) against the oracle's tick,
    12 ticks at level 4 in the FP64 form (<= 1e-10: the dynamics' pow/exp ulps grow slowly, as the core's own
    multi-step bar); graph replay bit-identical to direct launches; the column moved air across the interface."""
    from insilica.world import World
    from oracle.world import OracleWorld
    print('case: held_suarez world level 4, 12 ticks', flush=True)
    w = World(4, 'held_suarez', precision='double')
    o = OracleWorld(w)
    w.step(12, graph=False)
    o.step(12)
    got = {k: _np(v) for k, v in w.state_arrays().items()}
    ref = o.state()
    errs = {k: relmax(got[k], ref[k]) for k in ref}
    print('held_suarez world errors', errs, flush=True)
    assert max(errs.values()) <= 1e-10
    assert np.count_nonzero(_np(w.physics.wx)) > 0
    w2 = World(4, 'held_suarez', precision='double')
    w2.step(12, graph=True)
    assert w2.state_hash() == w.state_hash()
