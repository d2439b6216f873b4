"""Pins of the oracle's ocean column processes (CPU only): oracle/ocean.c against the reference's ocean.cu kernels.

Each check is computed here independently of the oracle's code:

  * the convective adjustment (reference: kernels/ocean.cu:55-72, 231-237) leaves every column stable (buoyancy not
    increasing downward between its mixed blocks), each block uniform with the thickness-weighted mean of the heat
    and salt it started with, and every column's heat and salt unchanged;
  * the mixed layer's Kraus-Turner step (ocean.cu:117-213): with wind alone it deepens at 2 m u*^3 / (h db) dt (u* from
    the stress, db the buoyancy jump at its base, from the equation of state written out here); with strong sunshine
    and little wind it shoals to the depth where the turbulent energy balances, found here by quadrature of Paulson &
    Simpson's light; the shear instability takes it to sqrt(ri_b du^2 h / db); kappa_d moves 2 kappa_d dt / (h_k +
    h_k+1) across each interior interface; every move stays within hmin and hmax;
  * salt moved with the water (ocean.cu:352-381) and the exchange conserve each column's heat, salt and volume;
  * the sunlight passing the mixed layer's base (ocean.cu:387-407) heats each layer below by the light absorbed within
    it (Paulson & Simpson's two exponentials) and takes exactly that from the mixed layer;
  * fresh water (ocean.cu:245-261) changes the mixed layer's salt by -S_ref fw / rho_w (the virtual flux), or with
    fw_local keeps the salt and adds the volume;
  * mixing along the layers (ocean.cu:268-292) keeps every layer's heat and salt, leaves a uniform field unchanged
    and never makes a new extremum (a wrong sign would).
"""
import numpy as np
import pytest

import oracle
from oracle import ocean as oo

EOS = dict(grav=9.80616, alpha=0.52e-4, alpha2=1.0e-5, beta=7.6e-4)


def buoy(T, S, grav=9.80616, alpha=0.52e-4, alpha2=1.0e-5, beta=7.6e-4):
    th = T - 273.15
    return grav * (alpha * th + 0.5 * alpha2 * th * th - beta * S)


def columns(rng, N, L, unstable=True):
    h = np.vstack([50 + 100 * rng.random((1, N)), 100 + 900 * rng.random((L - 1, N))])
    T = 273.15 + 25 * np.sort(rng.random((L, N)), axis=0)[::-1]      # warm over cold
    S = 35 + rng.standard_normal((L, N)) * 0.3
    if unstable:
        flip = rng.random(N) < 0.5
        T[:, flip] = T[::-1, flip]                                   # cold over warm in half the columns
    return np.ascontiguousarray(h), np.ascontiguousarray(T), np.ascontiguousarray(S)


@pytest.mark.parametrize('L', [2, 3, 5])
def test_overturn_stable_blocks_of_means(L):
    rng = np.random.default_rng(L)
    N = 4000
    h, T0, S0 = columns(rng, N, L)
    T, S = T0.copy(), S0.copy()
    oo.overturn(N, L, h, T, S, **EOS)
    np.testing.assert_allclose(np.sum(h * T, axis=0), np.sum(h * T0, axis=0), rtol=1e-14)
    np.testing.assert_allclose(np.sum(h * S, axis=0), np.sum(h * S0, axis=0), rtol=1e-14)
    changed = 0
    for i in range(N):
        # the blocks: runs of equal (T, S) from the top
        ks = [0]
        for k in range(1, L):
            if not (T[k, i] == T[k - 1, i] and S[k, i] == S[k - 1, i]):
                ks.append(k)
        ks.append(L)
        for a, b in zip(ks[:-1], ks[1:]):
            hs = np.sum(h[a:b, i])
            assert abs(T[a, i] - np.sum(h[a:b, i] * T0[a:b, i]) / hs) <= 1e-12 * T[a, i]
            assert abs(S[a, i] - np.sum(h[a:b, i] * S0[a:b, i]) / hs) <= 1e-12 * S[a, i]
        bb = [buoy(T[a, i], S[a, i]) for a in ks[:-1]]
        assert all(bb[j] >= bb[j + 1] for j in range(len(bb) - 1))      # lighter over denser
        changed += len(ks) - 1 < L
    assert changed > N // 4


def mix_defaults(**kw):
    d = dict(S_ref=35.0, alpha_ref=2.0e-4, rho_w=1025.0, cw=3990.0, m_kt=0.5, n_kt=0.2, hmin=10.0, hmax=1000.0,
             dTmin=0.1, sw_R=0.58, sw_z1=0.35, sw_z2=23.0, kappa_d=0.0, dwf=0.0, ri_b=0.0, reach=0.0, h_ent=0.0,
             det_class=0.0, fw_local=0.0)
    d.update(EOS)
    d.update(kw)
    return d


def stable_columns(rng, N, L):
    h, T, S = columns(rng, N, L, unstable=False)
    S[:] = 35.0
    return h, T, S


def test_kraus_turner_wind_deepening():
    rng = np.random.default_rng(1)
    N, L = 2000, 3
    h, T, S = stable_columns(rng, N, L)
    taum = 0.02 + 0.3 * rng.random(N)
    z = np.zeros(N)
    dt = 86400.0
    p = mix_defaults()
    wx = np.zeros((L - 1, N))
    oo.column(N, L, p, h, T.copy(), S.copy(), taum, z, z, z, dt, wx)
    us3 = (taum / p['rho_w']) ** 1.5
    db = buoy(T[0], S[0]) - buoy(T[1], S[1])
    want = 2 * p['m_kt'] * us3 / (h[0] * db) * dt
    want = np.where(db > p['grav'] * p['alpha_ref'] * p['dTmin'], want, 0.0)   # an unstratified base entrains nothing
    want = np.minimum(want, np.minimum(h[1] - p['hmin'], p['hmax'] - h[0]))
    assert np.sum(want == 0) > 0 and np.sum(want > 0) > N // 2
    np.testing.assert_allclose(wx[0], want, rtol=1e-12)
    assert np.all(wx[1] == 0)
    assert np.any(wx[0] < h[1] - p['hmin'])          # some columns below the cap: the rate itself is checked


def light(z, R=0.58, z1=0.35, z2=23.0):
    return R * np.exp(-z / z1) + (1 - R) * np.exp(-z / z2)


def test_kraus_turner_shoals_to_balance():
    from scipy.integrate import quad
    rng = np.random.default_rng(2)
    N, L = 300, 3
    h, T, S = stable_columns(rng, N, L)
    h[0] = 80 + 100 * rng.random(N)
    taum = 0.01 + 0.02 * rng.random(N)
    qsun = 150 + 150 * rng.random(N)
    z = np.zeros(N)
    p = mix_defaults()
    wx = np.zeros((L - 1, N))
    oo.column(N, L, p, h, T.copy(), S.copy(), taum, z, z, qsun, 86400.0, wx)
    hn = h[0] + wx[0]
    assert np.all(wx[0] < 0)
    us3 = (taum / p['rho_w']) ** 1.5
    ga = p['grav'] * (p['alpha'] + p['alpha2'] * (T[0] - 273.15))
    for i in range(N):
        hh = hn[i]
        if hh <= p['hmin'] * (1 + 1e-9):
            continue
        Bsw = ga[i] * qsun[i] / (p['rho_w'] * p['cw'])
        integral = quad(light, 0, hh, epsabs=0, epsrel=1e-13)[0]
        B = Bsw * (1 + light(hh) - 2 * integral / hh)
        P = 2 * p['m_kt'] * us3[i] - 0.5 * hh * ((1 + p['n_kt']) * B + (1 - p['n_kt']) * abs(B))
        assert abs(P) <= 1e-9 * 2 * p['m_kt'] * us3[i], (i, P)


def test_shear_depth_and_interior_mixing():
    rng = np.random.default_rng(3)
    N, L = 1000, 4
    h, T, S = stable_columns(rng, N, L)
    z = np.zeros(N)
    du = 0.2 + 0.5 * rng.random(N)
    p = mix_defaults(ri_b=0.65, kappa_d=1e-4)
    wx = np.zeros((L - 1, N))
    oo.column(N, L, p, h, T.copy(), S.copy(), z + 1e-12, z, z, z, 86400.0, wx, shear=du)
    db = buoy(T[0], S[0]) - buoy(T[1], S[1])
    want = np.sqrt(p['ri_b'] * du * du * h[0] / db) - h[0]
    want = np.clip(want, 0.0, None)
    want = np.minimum(want, np.minimum(h[1] - p['hmin'], p['hmax'] - h[0]))
    np.testing.assert_allclose(wx[0], want, rtol=1e-12, atol=1e-9)
    for k in (1, 2):
        wk = np.minimum(2 * p['kappa_d'] / (h[k] + h[k + 1]) * 86400.0, h[k + 1] - p['hmin'])
        np.testing.assert_allclose(wx[k], wk, rtol=1e-13)


@pytest.mark.parametrize('seq', [False, True])
def test_mixing_moves_keep_heat_salt_volume(grids, seq):
    g = grids(4)
    og = oracle.Grid(g)
    rng = np.random.default_rng(5)
    N, L = g.N, 3
    h, T, S = columns(rng, N, L)
    taum = 0.3 * rng.random(N)
    qns = -200 + 300 * rng.random(N)
    qsun = 200 * rng.random(N)
    qfw = 1e-5 * rng.standard_normal(N)
    p = mix_defaults(kappa_d=1e-4)
    wx = np.zeros((L - 1, N))
    A = g.areaCell
    H0, S0, V0 = np.sum(A * h * T), np.sum(A * h * S), np.sum(A * h)
    oo.column(N, L, p, h, T, S, taum, qns, qfw, qsun, 86400.0, wx)
    assert np.max(np.abs(wx[0])) > 1.0
    oo.xch_salt(N, L, wx, h, S, seq)
    lay = oracle.Layers(og, L, gprime=[0.016] * (L - 1), floor=10.0)
    u = rng.standard_normal((L, g.E)) * 0.1
    M = h * A
    lay.exchange(wx, h, u, M, T, seq=seq)
    assert abs(np.sum(A * h * T) - H0) / H0 <= 1e-13
    assert abs(np.sum(A * h * S) - S0) / S0 <= 1e-13
    assert abs(np.sum(A * h) - V0) / V0 <= 1e-14
    assert np.min(h) >= 10.0 - 1e-9


def test_sunlight_below_the_mixed_layer():
    rng = np.random.default_rng(6)
    N, L = 500, 4
    h, T, S = stable_columns(rng, N, L)
    h[0] = 5 + 40 * rng.random(N)
    qsw = 300 * rng.random(N)
    qoc = qsw + 50 * rng.standard_normal(N)
    p = mix_defaults()
    rc = p['rho_w'] * p['cw']
    T1, qa, sa = T.copy(), np.zeros(N), np.zeros(N)
    dt = 450.0
    oo.pen(N, L, p, h, T1, qoc, qsw, dt, qa, sa)
    np.testing.assert_allclose(qa, (qoc - qsw) * dt, rtol=1e-15)
    np.testing.assert_allclose(sa, qsw * dt, rtol=1e-15)
    z = np.cumsum(h, axis=0)
    f = light(z)
    f[-1] = 0.0
    # (to the rounding of T's small change: an ulp of T is ~1e-5 J/m^2 here)
    tol = 8 * np.spacing(np.max(T)) * rc * np.max(h)
    np.testing.assert_allclose((T1[0] - T[0]) * rc * h[0], -qsw * light(h[0]) * dt, rtol=1e-9, atol=tol)
    for k in range(1, L):
        np.testing.assert_allclose((T1[k] - T[k]) * rc * h[k], qsw * (f[k - 1] - f[k]) * dt, rtol=1e-8, atol=tol)
    np.testing.assert_allclose(np.sum(h * (T1 - T), axis=0) * rc, 0.0, atol=1e-6 * np.max(qsw) * dt)


@pytest.mark.parametrize('local', [0.0, 1.0])
def test_fresh_water(local):
    rng = np.random.default_rng(7)
    N = 1000
    h = 20 + 100 * rng.random(N)
    S = 35 + rng.standard_normal(N) * 0.1
    rain, evap = 1e-4 * rng.random(N), 1e-4 * rng.random(N)
    ice, ice_prev = 10 * rng.random(N), 10 * rng.random(N)
    p = mix_defaults(fw_local=local)
    dt = 450.0
    fw = (rain - evap) * dt - (ice - ice_prev)
    S1, ip, acc = S.copy(), ice_prev.copy(), np.zeros(N)
    hw, M, area = h.copy(), np.zeros(N), np.full(N, 2.0e11)
    oo.fresh(N, p, h, S1, evap, rain, ice, ip, dt, acc, hw=hw, M=M, area=area)
    np.testing.assert_allclose(acc, fw, rtol=1e-15)
    assert np.all(ip == ice)
    if local:
        np.testing.assert_allclose(hw, h + fw / p['rho_w'], rtol=1e-14)
        np.testing.assert_allclose(hw * S1, h * S, rtol=1e-14)          # the salt kept
        np.testing.assert_allclose(M, hw * area, rtol=1e-15)
    else:
        np.testing.assert_allclose(h * (S1 - S), -p['S_ref'] * fw / p['rho_w'], rtol=1e-10,
                                   atol=8 * np.spacing(np.max(S)) * np.max(h))      # (S's rounding)


def test_mixing_along_layers(grids):
    g = grids(4)
    og = oracle.Grid(g)
    rng = np.random.default_rng(8)
    L, N = 2, g.N
    h = 50 + 500 * rng.random((L, N))
    T = 273.15 + 20 * rng.random((L, N))
    S = 35 + rng.standard_normal((L, N))
    Tn, Sn = np.zeros_like(T), np.zeros_like(S)
    oo.diffuse(og, L, h, T, S, Tn, Sn, None, 500.0, 86400.0)
    A = g.areaCell
    for k in range(L):
        assert abs(np.sum(A * h[k] * Tn[k]) - np.sum(A * h[k] * T[k])) <= 1e-13 * np.sum(A * h[k] * T[k])
        assert abs(np.sum(A * h[k] * Sn[k]) - np.sum(A * h[k] * S[k])) <= 1e-13 * np.sum(A * h[k] * S[k])
    assert np.max(Tn) <= np.max(T) and np.min(Tn) >= np.min(T)
    assert np.std(Tn) < np.std(T)
    U = np.full((L, N), 280.0)
    oo.diffuse(og, L, h, U, U, Tn, Sn, None, 500.0, 86400.0)
    assert np.all(Tn == 280.0)
