"""The radiation oracle pinned on the CPU (oracle/radiation.c).

- LONGWAVE against line-by-line: the reference's own check (tests/test_radiation.py:95-146) on its fixtures
  (tests/golden/lbl: seven AFGL columns, fluxes per band from a line-by-line code) with its k-tables
  (tests/golden/ktables/lw_ktables.npz): OLR and surface down within 1%, heating within 0.2 K/day below 1 hPa,
  2 x CO2 forcing at the top and the tropopause within 10%. The oracle's lw_column runs each level as its own layer.
- The two-stream step against the radiative transfer equation solved by quadrature (a source linear in optical depth),
  the small-depth branch included; an isothermal column splits into layers without changing its fluxes.
- lw_setup's sub-levels against what they must hold: air masses summing to the surface pressure, edge temperatures
  continuous, the layer's vapor (mass-weighted) kept, the haze's shares summing to one, ozone on the table's
  log-pressure axis, the cloud in the sub-level holding the median of the vapor.
- SHORTWAVE: Delta-Eddington against the two-stream equations integrated with a matrix exponential (direct and
  diffuse), its limits (a clear layer transmits everything; conservative scattering absorbs nothing), and the column's
  shares closing to one; a transparent column gives the surface 1 - albedo; the sun's cosine from the orbit.
"""
import math
from pathlib import Path

import numpy as np
import pytest
from scipy import integrate, linalg

from oracle import radiation as orad

GOLD = Path(__file__).resolve().parent / 'golden'
COLUMNS = ['us_standard_400', 'us_standard_800', 'tropical_400', 'subarctic_winter_400', 'tropical_800',
           'subarctic_winter_400_h2o2', 'us_standard_400_h2o2_dT10']
GRAV, M_AIR, NA = 9.80665, 28.9647e-3, 6.02214076e23
SUB_HPA = (0, 1, 5, 20, 50, 100, 150, 200, 300, 400, 500, 600, 700, 800, 900, 1e4)


def par(L=3, **kw):
    """The column parameters (the radiation reads none of the interfaces' theta*, so any layer count's array)."""
    from insilica.atmosphere import par_array
    return par_array(len(kw.get('theta_stars', (380.0, 305.0))) + 1, kw)


@pytest.fixture(scope='module')
def lwk():
    return dict(np.load(GOLD / 'ktables' / 'lw_ktables.npz'))


def lbl_sub(t, r):
    """A line-by-line column's layers as sub-levels (each its own layer; no clouds), float32 as the kernels store
    them (the reference's afgl_terms, tests/test_radiation.py:74-87)."""
    p, T = r['p'], r['T']
    pm = (p[1:] - p[:-1]) / np.log(p[1:] / p[:-1])
    Tm = (T[1:] + T[:-1]) / 2
    air = np.diff(p) / GRAV / M_AIR * NA / 1e4
    mid = lambda g: (r['vmr_' + g][1:] + r['vmr_' + g][:-1]) / 2
    P, TT = t['P'], t['T']
    x = (np.log(P[0]) - np.log(pm)) / ((np.log(P[0]) - np.log(P[-1])) / (len(P) - 1))
    sub = np.zeros((len(pm), 15), np.float32)
    sub[:, 0], sub[:, 1] = T[:-1], T[1:]
    sub[:, 2] = np.clip(x, 0, t['logk'].shape[1] - 1 - 1e-4)
    sub[:, 3] = np.clip((Tm - TT[0]) / (TT[1] - TT[0]), 0, len(TT) - 1 - 1e-4)
    for c, g in enumerate(('H2O', 'CO2', 'O3', 'N2O', 'CH4')):
        sub[:, 4 + c] = mid(g) * air
    sub[:, 9] = mid('H2O') / 0.04
    return sub.reshape(-1, 1)


def heating(up, dn, p):
    net = up - dn
    return (net[1:] - net[:-1]) / np.diff(p) * GRAV / 1004.64 * 86400


def test_longwave_against_line_by_line(lwk):
    lwt = orad.LwTables(lwk)
    worst = dict(olr=0.0, down=0.0, heat=0.0, forcing=0.0)
    net = {}
    for name in COLUMNS:
        r = np.load(GOLD / 'lbl' / f'lbl_{name}.npz')
        L = len(r['p']) - 1
        radf = orad.lw_column(1, L, par(), lwt, lbl_sub(lwk, r), np.array([r['T'][-1]]), nsub=1)[:, 0]
        RW = L + 1
        U = radf[:RW] + np.pi * radf[2 * RW:3 * RW] * radf[3 * RW + 2]   # the surface black at the lowest level's T
        D = radf[RW:2 * RW]
        Ur, Dr = r['up'].sum(0), r['dn'].sum(0)
        net[name] = (U - D, Ur - Dr, r['p'])
        worst['olr'] = max(worst['olr'], abs(U[0] / Ur[0] - 1))
        worst['down'] = max(worst['down'], abs(D[-1] / Dr[-1] - 1))
        p = r['p']
        sub = sorted({int(np.argmin(np.abs(p - q * 100))) for q in SUB_HPA})
        e = np.abs(heating(U[sub], D[sub], p[sub]) - heating(Ur[sub], Dr[sub], p[sub]))
        worst['heat'] = max(worst['heat'], e[1:].max())
    for a, b in (('us_standard_400', 'us_standard_800'), ('tropical_400', 'tropical_800')):
        Fk, Fr = net[a][0] - net[b][0], net[a][1] - net[b][1]
        p = net[a][2]
        for j in (0, int(np.argmin(np.abs(p - (22700 if 'us' in a else 9400))))):
            worst['forcing'] = max(worst['forcing'], abs(Fk[j] / Fr[j] - 1))
    print('LBL worst:', worst)
    assert worst['olr'] <= 0.01 and worst['down'] <= 0.01 and worst['heat'] <= 0.2 and worst['forcing'] <= 0.10


def gray_tables(k0=1e-22, NG=1, a=100.0, b=2.0):
    """Tables with one constant ln k for every component, position and g-point, and a Planck table linear in T."""
    P = np.exp(np.linspace(np.log(1.1e5), 0.0, 31))
    T = np.arange(150.0, 341.0, 10.0)
    Tp = np.arange(140.0, 351.0, 1.0)
    logk = np.full((6, 31, len(T), NG), np.log(k0), np.float32)
    planck = np.repeat((a + b * (Tp - 140.0))[:, None], NG, 1).astype(np.float32)
    return dict(logk=logk, planck=planck, P=P, T=T, T_pl=Tp, o3=np.full(31, 1e-7, np.float32))


def one_sub(Tt, Tb, u, x=10.0, y=5.0, tc=0.0, fc=0.0):
    s = np.zeros(15, np.float32)
    s[0], s[1], s[2], s[3], s[4], s[10], s[11] = Tt, Tb, x, y, u, tc, fc
    return s


@pytest.mark.parametrize('tau', [2e-4, 0.05, 0.7, 4.0])
def test_two_stream_step_is_the_transfer_equation(tau):
    """One layer, source linear in optical depth: down flux at the bottom and up flux at the top (the air's emission)
    and the transmission, against the transfer equation integrated by quadrature along the diffuse path."""
    k0 = 1e-22
    t = gray_tables(k0=k0)
    Tt, Tb = 210.0, 290.0
    Bt, Bb = 100.0 + 2.0 * (Tt - 140.0), 100.0 + 2.0 * (Tb - 140.0)
    u = tau / k0
    sub = one_sub(Tt, Tb, u).reshape(-1, 1)
    radf = orad.lw_column(1, 1, par(1), orad.LwTables(t), sub, np.array([Tb]), nsub=1)[:, 0]
    x = 1.66 * float(np.float32(u)) * k0
    B = lambda s, b0, b1: b0 + (b1 - b0) * s / x          # linear in optical path from the far side
    dn = integrate.quad(lambda s: B(s, Bt, Bb) * math.exp(-(x - s)), 0, x, epsabs=0, epsrel=1e-12)[0]
    up = integrate.quad(lambda s: B(s, Bb, Bt) * math.exp(-(x - s)), 0, x, epsabs=0, epsrel=1e-12)[0]
    RW = 2
    tol = 2e-7 if x > 1e-3 else 2e-4 * x                   # float32 per-g storage; the small-depth series' x^3 term
    assert abs(radf[RW + 1] - np.pi * dn) <= tol * np.pi * max(Bt, Bb)
    assert abs(radf[0] - np.pi * up) <= tol * np.pi * max(Bt, Bb)
    assert abs(radf[2 * RW] - math.exp(-x)) <= 2e-7         # the surface's share reaching the top
    assert radf[3 * RW + 2] == pytest.approx(Bb, rel=1e-7)  # the Planck sum at the surface temperature


def test_isothermal_column_splits_freely():
    """An isothermal gray column: the fluxes at its top and bottom do not depend on how it is split into layers."""
    k0 = 1e-22
    t = orad.LwTables(gray_tables(k0=k0, NG=3))
    tau = 1.3
    out = []
    for L in (1, 2, 5):
        sub = np.stack([one_sub(250.0, 250.0, tau / L / k0) for _ in range(L)]).reshape(-1, 1)
        radf = orad.lw_column(1, L, par(L) if L > 1 else par(1), t, sub, np.array([250.0]), nsub=1)[:, 0]
        RW = L + 1
        out.append((radf[0], radf[RW + L], radf[2 * RW]))
    for o in out[1:]:
        assert o == pytest.approx(out[0], rel=2e-6)


def test_cloud_mixes_by_cover():
    """A cloud over the whole of a layer (cover 1) is the clear layer with the cloud's depth added along the diffuse
    path; at cover f the down flux is the cover-weighted mix."""
    k0 = 1e-22
    t = orad.LwTables(gray_tables(k0=k0))
    u, tcl = 0.4 / k0, 3.0
    def dn(sub):
        return orad.lw_column(1, 1, par(1), t, sub.reshape(-1, 1), np.array([280.0]), nsub=1)[3, 0]
    clear = dn(one_sub(220.0, 280.0, u))
    thick = dn(one_sub(220.0, 280.0, u + tcl / 1.66 / k0))
    cloud1 = dn(one_sub(220.0, 280.0, u, tc=tcl, fc=1.0))
    cloudh = dn(one_sub(220.0, 280.0, u, tc=tcl, fc=0.3))
    assert cloud1 == pytest.approx(thick, rel=1e-6)
    assert cloudh == pytest.approx(0.7 * clear + 0.3 * thick, rel=1e-6)


def seeded_columns(N=64, L=5, seed=3):
    from insilica.ktables import column_state
    rng = np.random.default_rng(seed)
    pos = rng.normal(size=(N, 3))
    pos /= np.linalg.norm(pos, axis=1)[:, None]
    return column_state(pos, layers=L, seed=seed), pos


def test_lw_setup_sub_levels(lwk):
    L, N = 5, 64
    c, _ = seeded_columns(N, L)
    p = par(L, edge_k=2.0, theta_stars=[500.0, 380.0, 330.0, 305.0])
    lwt = orad.LwTables(lwk)
    sub = orad.lw_setup(N, L, p, lwt, c['mu'], c['theta'], c['q'], c['cloudo']).reshape(4 * L, 15, N).astype(float)
    grav = p[13]
    ps = grav * c['mu'].sum(0)
    ua = sub[:, 12]
    # air molecules per cm^2 sum to the column's mass
    assert np.allclose(ua.sum(0) * 1e4 / 6.02214076e23 * grav * 28.9647e-3, ps, rtol=1e-5)
    # edge temperatures continuous inside each layer
    for k in range(L):
        for j in range(3):
            assert np.allclose(sub[4 * k + j, 1], sub[4 * k + j + 1, 0], rtol=0, atol=1e-4)
    # the layers above the edge isothermal
    for k in range(2):
        assert np.all(sub[4 * k:4 * k + 4, 0:2] == sub[4 * k, 0])
    # the layer's vapor kept: mass-weighted mean of the sub-levels' specific humidity is the layer's q
    x = sub[:, 4] / ua                                     # H2O volume mixing ratio
    qj = x * 18.01528 / 28.9647 / (1 + x * 18.01528 / 28.9647)
    w = ua.reshape(L, 4, N) / ua.reshape(L, 4, N).sum(1, keepdims=True)
    assert np.allclose((qj.reshape(L, 4, N) * w).sum(1), c['q'], rtol=2e-5)
    # the haze's shares sum to one
    assert np.allclose(sub[:, 13].sum(0), 1.0, atol=2e-6)
    # table positions: the sub-level's log-mean pressure lies inside it; temperature positions from the table's axis
    pe = np.concatenate([np.zeros((1, N)), np.cumsum(grav * c['mu'], 0)])
    P = lwk['P']
    pm = np.exp(np.log(P[0]) - sub[:, 2] * (np.log(P[0]) - np.log(P[-1])) / 30)
    ptop = np.repeat(pe[:-1], 4, 0)
    pbot = np.repeat(pe[1:], 4, 0)
    assert np.all(pm <= pbot * (1 + 1e-5)) and np.all(pm >= ptop * (1 - 1e-5))
    # ozone: the table's profile interpolated in log pressure; CO2 well mixed
    o3 = np.interp(-np.log(pm), -np.log(P), lwk['o3'].astype(float))
    assert np.allclose(sub[4:, 6] / ua[4:], o3[4:], rtol=2e-3)
    assert np.allclose(sub[:, 5] / ua, 400e-6, rtol=1e-6)
    # each layer's cloud in exactly one sub-level, with its longwave depth
    L3 = c['cloudo'][L + 3:2 * L + 3]
    for k in range(L):
        tc = sub[4 * k:4 * k + 4, 10]
        assert np.all((tc > 0).sum(0) == (L3[k] > 0))
        assert np.allclose(tc.sum(0), L3[k], rtol=1e-6)
        if k == 1:                                        # vapor even: the median on the boundary of the 2nd/3rd
            assert np.all(tc[[0, 3]] == 0)
        if k >= 2:                                        # vapor grows downward: its median in the lower half
            assert np.all(tc[:2] == 0)
        assert np.all(sub[4 * k:4 * k + 4, 11] == np.float32(c['cloudo'][k]))


@pytest.fixture(scope='module')
def swk():
    return dict(np.load(GOLD / 'ktables' / 'sw_ktables.npz'))


def two_stream_ode(tau, ssa, g, mu0):
    """Delta-scaled Eddington two-stream equations (Joseph, Wiscombe & Weinman 1976; Meador & Weaver 1980):
    dF+/dt = g1 F+ - g2 F- - S w g3 e^(-t/mu0), dF-/dt = g2 F+ - g1 F- + S w g4 e^(-t/mu0), S mu0 = 1; the linear
    system solved by a matrix exponential with the particular solution of the source, a black surface below."""
    f = g * g
    tau, ssa, g = (1 - ssa * f) * tau, (1 - f) * ssa / (1 - ssa * f), (g - f) / (1 - f)
    g1, g2 = (7 - ssa * (4 + 3 * g)) / 4, -(1 - ssa * (4 - 3 * g)) / 4
    g3 = (2 - 3 * g * mu0) / 4
    g4 = 1 - g3
    A = np.array([[g1, -g2], [g2, -g1]])
    M = linalg.expm(A * tau)
    S = 1 / mu0
    s = np.array([-S * ssa * g3, S * ssa * g4])
    C = -np.linalg.solve(A + np.eye(2) / mu0, s)         # particular solution C e^(-t/mu0)
    e = math.exp(-tau / mu0)
    # direct beam: F(0) = [x, 0]; F(tau) = M (F(0) - C) + C e; F+(tau) = 0
    r0 = M @ (np.array([0.0, 0.0]) - C) + C * e
    dx = M @ np.array([1.0, 0.0])
    x = -r0[0] / dx[0]
    Fb = r0 + x * dx
    Rs, Ts = x, Fb[1]
    # diffuse: F(0) = [y, 1], no source
    q0 = M @ np.array([0.0, 1.0])
    y = -q0[0] / dx[0]
    Rd, Td = y, (q0 + y * dx)[1]
    return np.array([Rd, Td, Rs, Ts, e])


@pytest.mark.parametrize('tau,ssa,g,mu0', [(0.1, 0.9, 0.5, 0.6), (2.0, 0.99, 0.85, 0.3), (0.5, 0.3, 0.1, 0.9),
                                           (8.0, 0.999, 0.7, 0.5), (0.02, 0.5, 0.0, 0.2)])
def test_delta_eddington_solves_the_two_stream_equations(tau, ssa, g, mu0):
    got = orad.sw_layer(tau, ssa, g, mu0)
    want = two_stream_ode(tau, ssa, g, mu0)
    assert np.allclose(got, want, rtol=1e-9, atol=1e-12), (got, want)


def test_delta_eddington_limits():
    Rd, Td, Rs, Ts, Tdd = orad.sw_layer(0.0, 0.5, 0.3, 0.5)
    assert (Rd, Td, Rs, Ts, Tdd) == pytest.approx((0.0, 1.0, 0.0, 0.0, 1.0), abs=1e-15)
    for tau in (0.3, 3.0, 30.0):                          # conservative scattering: nothing absorbed
        Rd, Td, Rs, Ts, Tdd = orad.sw_layer(tau, 1 - 1e-12, 0.6, 0.7)
        assert Rd + Td == pytest.approx(1.0, abs=1e-9)
        assert Rs + Ts + Tdd == pytest.approx(1.0, abs=1e-9)


def sw_sub(L, N, seed, clouds=False):
    """Sub-levels for the shortwave from lw_setup on seeded columns, with the reference's tables."""
    c, pos = seeded_columns(N, L, seed)
    lwt = orad.LwTables(dict(np.load(GOLD / 'ktables' / 'lw_ktables.npz')))
    p = par(L, edge_k=1.0, theta_stars=[380.0, 330.0, 305.0, 300.0][-(L - 1):])
    sub = orad.lw_setup(N, L, p, lwt, c['mu'], c['theta'], c['q'], c['cloudo'] if clouds else None)
    return sub, c, pos, p


def test_shortwave_shares_close(swk):
    """Clear sky: each layer's absorbed share, the surface's and the reflected share add up to the sunlight."""
    L, N = 4, 32
    sub, c, pos, p = sw_sub(L, N, 5)
    swf = orad.sw_column(N, L, p, orad.SwTables(swk), sub, c['alb'], mode=1)
    total = swf[:L].sum(0) + swf[L] + swf[L + 1]
    assert np.allclose(total, 1.0, atol=2e-6)
    assert np.all(swf[:L + 2] >= 0)
    # each layer's top and bottom sub-level absorb part of what the layer does
    for k in range(L):
        assert np.all(swf[L + 2 + 2 * k] + swf[L + 3 + 2 * k] <= swf[k] * (1 + 1e-6) + 1e-9)


def test_transparent_column(swk):
    """No gas, no Rayleigh scattering, no haze: the surface absorbs 1 - albedo, the rest leaves."""
    L, N = 3, 8
    sub, c, pos, p = sw_sub(L, N, 6)
    t = dict(swk, logk=np.full_like(swk['logk'], -200.0), ray=np.zeros_like(swk['ray']),
             haze_par=np.array([0.0, 0.9, 0.7, 1500.0]))
    swf = orad.sw_column(N, L, p, orad.SwTables(t), sub, c['alb'], mode=1)
    assert np.allclose(swf[:L], 0.0, atol=1e-12)
    assert np.allclose(swf[L], 1 - c['alb'], rtol=1e-6)
    assert np.allclose(swf[L + 1], c['alb'], rtol=1e-6)


def test_sun_geometry(swk):
    """mode 0 at the pole, a quarter year after the equinox with the axis tilted 30 degrees: the sun stands 30
    degrees up all day, so the mean cosine is 0.5, mode 1's fixed value."""
    L, N = 3, 4
    sub, c, _, p = sw_sub(L, N, 7)
    p = p.copy()
    p[17] = math.radians(30.0)
    pos = np.tile([0.0, 0.0, 1.0], (N, 1))
    t = orad.SwTables(swk)
    a = orad.sw_column(N, L, p, t, sub, c['alb'], mode=0, tnow=p[18] / 4, interval=1.0, pos=pos)
    b = orad.sw_column(N, L, p, t, sub, c['alb'], mode=1)
    assert np.allclose(a, b, rtol=1e-6, atol=1e-9)
