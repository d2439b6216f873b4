"""The ocean's column processes on the GPU against the CPU oracle, through the C ABI (isw_ocn_*).

FP64 in both forms; bar: <= 1e-12 relative (max |diff| / max |oracle|) on every output. The arithmetic is the
oracle's in the same order with no contraction; what is left is one rounding of the device's exp/expm1/sqrt against
glibc's (the Kraus-Turner bisection and Paulson & Simpson's light), which the tolerance covers.
"""
import ctypes

import numpy as np
import pytest
import torch

from conftest import relmax

pytestmark = pytest.mark.gpu

TOL = 1e-12


_KEEP = []


def _dev(a):
    """A device copy of a, kept alive until the next test: a pointer taken from a temporary tensor would be freed
    before the call and its memory handed to the next argument's copy."""
    t = torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device='cuda')
    _KEEP.append(t)
    return t


@pytest.fixture(autouse=True)
def _release():
    yield
    torch.cuda.synchronize()
    _KEEP.clear()


def _np(t):
    return t.detach().cpu().numpy().copy()


def settings(**kw):
    from insilica.ocean import ocean_struct, OCEAN
    c = dict(OCEAN, **kw)
    return c, ocean_struct(c, 9.80616)


def oracle_par(c, grav=9.80616):
    from insilica.ocean import ALPHA_REF
    d = {k: c[k] for k in ('S_ref', 'hmin', 'hmax', 'dTmin', 'm_kt', 'n_kt', 'sw_R', 'sw_z1', 'sw_z2', 'kappa_d', 'dwf',
                           'ri_b', 'reach', 'h_ent', 'det_class', 'fw_local', 'alpha', 'alpha2', 'beta')}
    d.update(grav=grav, alpha_ref=ALPHA_REF, rho_w=c['rho'], cw=c['cw'])
    return d


def columns(rng, N, L):
    h = np.vstack([5 + 150 * rng.random((1, N)), 10.5 + 900 * rng.random((L - 1, N))])
    h[1:, rng.random(N) < 0.1] = 10.2                                     # some layers at their floor (reach)
    T = 273.15 + 25 * np.sort(rng.random((L, N)), axis=0)[::-1]
    flip = rng.random(N) < 0.3
    T[:, flip] = T[::-1, flip]
    S = 35 + 0.3 * rng.standard_normal((L, N))
    return np.ascontiguousarray(h), np.ascontiguousarray(T), np.ascontiguousarray(S)


@pytest.mark.parametrize('L', [2, 3, 5])
def test_overturn(L):
    from insilica import _hip
    from oracle import ocean as oo
    print(f'case: ocn_overturn L={L}', flush=True)
    rng = np.random.default_rng(L)
    N = 40962
    h, T, S = columns(rng, N, L)
    land = (rng.random(N) < 0.3).astype(float)
    c, o = settings()
    tT, tS = _dev(T), _dev(S)
    _hip.call('isw_ocn_overturn', N, L, ctypes.byref(o), _hip.ptr(_dev(h)), _hip.ptr(tT), _hip.ptr(tS),
              _hip.ptr(_dev(land)), _hip.stream())
    torch.cuda.synchronize()
    p = oracle_par(c)
    oo.overturn(N, L, h, T, S, p['grav'], p['alpha'], p['alpha2'], p['beta'], land=land)
    assert relmax(_np(tT), T) <= TOL and relmax(_np(tS), S) <= TOL


@pytest.mark.parametrize('L,opts', [(3, {}), (4, dict(kappa_d=1e-4, dwf=1.0, ri_b=0.65, h_ent=30.0)),
                                    (4, dict(reach=1.0, kappa_d=1e-4)), (5, dict(det_class=1.0, reach=1.0))])
def test_column(L, opts):
    from insilica import _hip
    from oracle import ocean as oo
    print(f'case: ocn_column L={L} {opts}', flush=True)
    rng = np.random.default_rng(10 + L)
    N = 40962
    h, T, S = columns(rng, N, L)
    taum = 0.4 * rng.random(N) ** 2
    qns = -250 + 300 * rng.random(N)
    qsun = 250 * rng.random(N)
    qfw = 2e-5 * rng.standard_normal(N)
    shear = 0.6 * rng.random(N)
    land = (rng.random(N) < 0.3).astype(float)
    Tc = np.concatenate([[0.0], 273.15 + np.linspace(20, 0, L - 1)]) if opts.get('det_class') else None
    c, o = settings(**opts)
    dt = 86400.0
    tT, tS = _dev(T), _dev(S)
    twx = torch.full((L - 1, N), np.nan, dtype=torch.float64, device='cuda')
    _hip.call('isw_ocn_column', N, L, ctypes.byref(o), _hip.ptr(_dev(h)), _hip.ptr(tT), _hip.ptr(tS),
              _hip.ptr(_dev(taum)), _hip.ptr(_dev(qns)), _hip.ptr(_dev(qfw)), _hip.ptr(_dev(qsun)), dt, _hip.ptr(twx),
              _hip.ptr(_dev(land)), _hip.ptr(_dev(shear) if opts.get('ri_b') else None),
              _hip.ptr(_dev(Tc) if Tc is not None else None), None, _hip.stream())
    torch.cuda.synchronize()
    wx = np.zeros((L - 1, N))
    oo.column(N, L, oracle_par(c), h, T, S, taum, qns, qfw, qsun, dt, wx, land=land,
              shear=shear if opts.get('ri_b') else None, Tc=Tc)
    e = [relmax(_np(tT), T), relmax(_np(tS), S), relmax(_np(twx), wx)]
    print('ocn_column errors', e, 'deepened', int(np.sum(wx[0] > 0)), 'shoaled', int(np.sum(wx[0] < 0)), flush=True)
    assert np.sum(wx[0] > 0) > N // 10 and np.sum(wx[0] < 0) > N // 10
    assert max(e) <= TOL


@pytest.mark.parametrize('seq', [0, 1])
def test_xch_salt_pen_fresh(seq):
    from insilica import _hip
    from oracle import ocean as oo
    print(f'case: ocn_xch_salt/pen/fresh seq={seq}', flush=True)
    rng = np.random.default_rng(20 + seq)
    N, L = 40962, 4
    h, T, S = columns(rng, N, L)
    wx = (rng.random((L - 1, N)) - 0.5) * np.minimum(h[:-1], h[1:]) * 0.5
    c, o = settings(fw_local=float(seq))
    tS = _dev(S)
    _hip.call('isw_ocn_xch_salt', N, L, _hip.ptr(_dev(wx)), _hip.ptr(_dev(h)), _hip.ptr(tS), seq, _hip.stream())
    qsw, qoc = 300 * rng.random(N), 200 * rng.standard_normal(N)
    up = rng.standard_normal(N)
    land = (rng.random(N) < 0.3).astype(float)
    qa, sa = rng.random(N), rng.random(N)
    tT, tqa, tsa = _dev(T), _dev(qa), _dev(sa)
    dt = 450.0
    _hip.call('isw_ocn_pen', N, L, ctypes.byref(o), _hip.ptr(_dev(h)), _hip.ptr(tT), _hip.ptr(_dev(qoc)),
              _hip.ptr(_dev(qsw)), dt, _hip.ptr(tqa), _hip.ptr(tsa), _hip.ptr(_dev(land)),
              _hip.ptr(_dev(up) if seq else None), _hip.stream())
    evap, rain = 1e-4 * rng.random(N), 1e-4 * rng.random(N)
    ice, ice_prev = 10 * rng.random(N), 10 * rng.random(N)
    mouth = rng.random(N)
    area = 1e11 * (1 + rng.random(N))
    hw, M, fw = h[0].copy(), np.zeros(N), rng.random(N)
    S0 = S[0].copy()                                      # (the mixed layer's salinity before any move)
    tS0 = _dev(S0)
    tip, tfw, tmouth, thw, tM = _dev(ice_prev), _dev(fw), _dev(mouth), _dev(hw), _dev(M)
    _hip.call('isw_ocn_fresh', N, ctypes.byref(o), _hip.ptr(_dev(h[0])), _hip.ptr(tS0), _hip.ptr(_dev(evap)),
              _hip.ptr(_dev(rain)), _hip.ptr(_dev(ice)), _hip.ptr(tip), dt, _hip.ptr(tfw), _hip.ptr(_dev(land)),
              _hip.ptr(tmouth), _hip.ptr(thw), _hip.ptr(tM), _hip.ptr(_dev(area)), _hip.stream())
    torch.cuda.synchronize()
    p = oracle_par(c)
    oo.xch_salt(N, L, wx, h, S, seq=bool(seq))
    oo.pen(N, L, p, h, T, qoc, qsw, dt, qa, sa, land=land, up=up if seq else None)
    oo.fresh(N, p, np.ascontiguousarray(h[0]), S0, evap, rain, ice, ice_prev, dt, fw, land=land, mouth=mouth, hw=hw,
             M=M, area=area)
    e = [relmax(_np(tS), S), relmax(_np(tT), T), relmax(_np(tqa), qa), relmax(_np(tsa), sa), relmax(_np(tS0), S0),
         relmax(_np(tip), ice_prev), relmax(_np(tfw), fw), relmax(_np(tmouth), mouth)]
    if seq:
        e += [relmax(_np(thw), hw), relmax(_np(tM), M)]
    print('xch_salt/pen/fresh errors', e, flush=True)
    assert max(e) <= TOL


def test_diffuse_and_forcing(grids):
    from insilica import _hip
    from insilica.gpu import DeviceGrid
    import oracle
    from oracle import ocean as oo
    print('case: ocn_diffuse + ocn_forcing level 5', flush=True)
    g = grids(5)
    dg = DeviceGrid(g, 'double')
    og = oracle.Grid(g)
    rng = np.random.default_rng(30)
    L, N, E = 3, g.N, g.E
    h = 20 + 500 * rng.random((L, N))
    T = 273.15 + 20 * rng.random((L, N))
    S = 35 + rng.standard_normal((L, N))
    wall = (rng.random(E) > 0.2).astype(float)
    tTn = torch.zeros(L, N, dtype=torch.float64, device='cuda')
    tSn = torch.zeros(L, N, dtype=torch.float64, device='cuda')
    _hip.call('isw_ocn_diffuse', dg.ref, L, _hip.ptr(_dev(h)), _hip.ptr(_dev(T)), _hip.ptr(_dev(S)), _hip.ptr(tTn),
              _hip.ptr(tSn), _hip.ptr(_dev(wall)), 800.0, 86400.0, _hip.stream())
    tau = rng.random(E) * 1e4
    q, s, f = rng.standard_normal(N) * 1e7, rng.random(N) * 1e7, rng.standard_normal(N)
    ttau, ttm, tq, ts, tf = _dev(tau), torch.zeros(N, dtype=torch.float64, device='cuda'), _dev(q), _dev(s), _dev(f)
    _hip.call('isw_ocn_forcing', dg.ref, 86400.0, _hip.ptr(ttau), _hip.ptr(ttm), _hip.ptr(tq), _hip.ptr(ts),
              _hip.ptr(tf), _hip.stream())
    torch.cuda.synchronize()
    Tn, Sn = np.zeros((L, N)), np.zeros((L, N))
    oo.diffuse(og, L, h, T, S, Tn, Sn, wall, 800.0, 86400.0)
    rms = np.sqrt(tau * (1.0 / 86400.0))
    ce = g.cellEdge
    ok = ce >= 0
    cs = np.where(ok, ce, 0)
    ke = np.zeros(N)
    for m in range(6):                                    # in table order, as the kernel sums
        ke = ke + np.where(ok[:, m], g.le[cs[:, m]] * g.de[cs[:, m]] * 0.25 * rms[cs[:, m]] * rms[cs[:, m]], 0.0)
    taum = np.sqrt(2 * ke / g.areaCell)
    e = [relmax(_np(tTn), Tn), relmax(_np(tSn), Sn), relmax(_np(ttau), rms), relmax(_np(ttm), taum),
         relmax(_np(tq), q / 86400.0), relmax(_np(ts), s / 86400.0), relmax(_np(tf), f / 86400.0)]
    print('diffuse/forcing errors', e, flush=True)
    assert max(e) <= TOL
