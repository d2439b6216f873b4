"""The oracle of the tides (oracle/tides.c and the tides' hooks in oracle.c's layers), pinned on closed forms on the
CPU: the degree-2 potential under a body (love GM R^2 / r^3) and a quarter turn away (half that, negative), the
Earth turning under the Sun (the sub-solar point moving west), the Moon's distance from Kepler's equation solved
independently, the mixing's exact heat and salt and its fraction of the buoyancy step, the internal tide's drag in
closed form, and a one-layer ocean under a fixed potential settling to the equilibrium tide eta = V / g."""
import math

import numpy as np
import pytest
from scipy.optimize import brentq

import oracle
from oracle import tides as ot
from insilica import grid as gridmod
from insilica.tides import TIDES

DAY, GRAV = 86400.0, 9.80616


def tp(tilt=0.0, year_s=1e12, day_s=DAY, ecc=0.0, lam_peri=0.0, sun_gm=TIDES['sun_gm'], sun_a=TIDES['sun_a'],
       moon_gm=0.0, moon_a=3.84748e8, moon_e=0.0549, moon_i=0.0, moon_T=27.321662 * DAY, node_T=-6798.38 * DAY,
       aps_T=3232.6 * DAY, L0=0.0, node0=0.0, aps0=0.0, love=0.69, radius=6.371e6):
    return np.array([tilt, year_s, day_s, ecc, lam_peri, sun_gm, sun_a, moon_gm, moon_a, moon_e, moon_i, moon_T,
                     node_T, aps_T, L0, node0, aps0, love, radius])


def pot(t, par, pts):
    og = oracle.Grid(gridmod.get(2))
    pts = np.ascontiguousarray(pts, dtype=float)
    V = np.zeros(len(pts))
    og_small = type('G', (), {'N': len(pts)})()
    ot.tide_pot(og_small, par, t, 0.0, pts, V)
    return V


def test_sun_potential_and_the_turning_earth():
    """The Sun alone, at the equinox (a year so long it stays put): under it V = love GM R^2 / a^3, a quarter turn
    away -V / 2; an eighth of a day later the sub-solar point is 45 degrees west."""
    par = tp()
    A = 0.69 * TIDES['sun_gm'] * 6.371e6 ** 2 / TIDES['sun_a'] ** 3
    pts = [[1, 0, 0], [0, 1, 0], [0, 0, 1], [0, -1, 0]]
    V = pot(0.0, par, pts)
    assert np.allclose(V, [A, -A / 2, -A / 2, -A / 2], rtol=1e-14, atol=0)
    # an eighth of a day later the Sun is over longitude -45 degrees (the ground turning east under it): full tide
    # there and at +135, low tide at +45 and -135
    r = math.sqrt(0.5)
    V = pot(DAY / 8, par, [[r, -r, 0], [-r, r, 0], [r, r, 0], [-r, -r, 0]])
    assert np.allclose(V, [A, A, -A / 2, -A / 2], rtol=1e-12, atol=1e-12 * A)


def test_moon_distance_by_kepler():
    """The Moon alone in the ecliptic (tilt 0, inclination 0): at perigee V under it is love GM R^2 / (a (1 - e))^3;
    a quarter month later its distance and direction follow Kepler's equation solved here by bisection."""
    e, T = 0.0549, 27.321662 * DAY
    par = tp(sun_gm=0.0, moon_gm=TIDES['moon']['gm'], moon_e=e, moon_T=T, node_T=1e30, aps_T=1e30)
    A = lambda r: 0.69 * TIDES['moon']['gm'] * 6.371e6 ** 2 / (3.84748e8 * r) ** 3
    V = pot(0.0, par, [[1, 0, 0]])
    assert abs(V[0] / A(1 - e) - 1) < 1e-13
    t = T / 4
    M = 2 * math.pi * t / T
    E = brentq(lambda x: x - e * math.sin(x) - M, 0, 2 * math.pi, xtol=1e-15)
    r = 1 - e * math.cos(E)
    nu = 2 * math.atan2(math.sqrt(1 + e) * math.sin(E / 2), math.sqrt(1 - e) * math.cos(E / 2))
    rot = 2 * math.pi * math.fmod(t, DAY) / DAY
    # (the Sun's right ascension with a year of 1e12 s: 2 pi t / year, its turn in the planet's frame)
    ras = 2 * math.pi * t / 1e12
    lon = -rot + (nu - ras)
    V = pot(t, par, [[math.cos(lon), math.sin(lon), 0.0]])
    assert abs(V[0] / A(r) - 1) < 1e-10


def _column(L, T, S, h, U, N=1):
    """Arrays for tide_mix on a level-2 mesh where every cell holds the same column and every edge the same bottom
    wind U."""
    g = gridmod.get(2)
    og = oracle.Grid(g)
    hh = np.repeat(np.asarray(h, float)[:, None], g.N, axis=1)
    TT = np.repeat(np.asarray(T, float)[:, None], g.N, axis=1)
    SS = np.repeat(np.asarray(S, float)[:, None], g.N, axis=1)
    u = np.zeros((L, g.E))
    u[L - 1] = U
    return g, og, hh, u, TT, SS


def test_mixing_keeps_heat_and_salt_and_narrows_the_step():
    """Two layers, the linear equation of state: heat and salt are kept exactly; the buoyancy step falls to
    (1 - f) of its value, f = gamma D dt e^(-h2 / zeta) / (rho h1 h2 db / 2), D = rho (c + cd |u_b|) |u_b|^2."""
    rho, c_bot, cd, gamma, dt, zeta = 1025.0, 2.5e-4, 2.5e-3, 0.2, 3600.0, 500.0
    eos = (GRAV, 2e-4, 0.0, 7.6e-4)                           # (alpha2 0: db linear in dT and dS)
    g, og, h, u, T, S = _column(2, [290.0, 280.0], [35.0, 35.2], [50.0, 60.0], 1.2)
    heat0, salt0 = np.sum(h * T, 0), np.sum(h * S, 0)
    db0 = GRAV * (2e-4 * (T[0] - T[1]) - 7.6e-4 * (S[0] - S[1]))
    dsum = np.zeros(g.N)
    ot.tide_mix(og, 2, h, u, T, S, None, c_bot, rho, gamma, dt, eos, dsum, cd, None, zeta)
    assert np.allclose(np.sum(h * T, 0), heat0, rtol=1e-15, atol=0)
    assert np.allclose(np.sum(h * S, 0), salt0, rtol=1e-15, atol=0)
    u2 = np.array([np.sum([g.le[e] * g.de[e] * 1.44 for e in g.cellEdge[i] if e >= 0]) / (2 * g.areaCell[i])
                   for i in range(g.N)])
    D = rho * (c_bot + cd * np.sqrt(u2)) * u2
    assert np.allclose(dsum, D * dt, rtol=1e-14)
    f = gamma * D * dt * math.exp(-60.0 / zeta) / (rho * 50.0 * 60.0 * db0 / 2)
    assert np.all(f < 1) and np.all(f > 1e-4)
    db1 = GRAV * (2e-4 * (T[0] - T[1]) - 7.6e-4 * (S[0] - S[1]))
    assert np.allclose(db1, (1 - f) * db0, rtol=1e-12)


def test_unstable_or_land_columns_are_left():
    rho, eos = 1025.0, (GRAV, 2e-4, 1e-5, 7.6e-4)
    g, og, h, u, T, S = _column(2, [280.0, 290.0], [35.0, 35.0], [50.0, 60.0], 1.0)    # warm under cold: db < 0
    T0, S0 = T.copy(), S.copy()
    ot.tide_mix(og, 2, h, u, T, S, None, 2.5e-4, rho, 0.2, 3600.0, eos, np.zeros(g.N), 2.5e-3, None, 500.0)
    assert np.array_equal(T, T0) and np.array_equal(S, S0)
    g, og, h, u, T, S = _column(2, [290.0, 280.0], [35.0, 35.0], [50.0, 60.0], 1.0)
    T0, dsum = T.copy(), np.zeros(g.N)
    ot.tide_mix(og, 2, h, u, T, S, np.ones(g.N), 2.5e-4, rho, 0.2, 3600.0, eos, dsum, 2.5e-3, None, 500.0)
    assert np.array_equal(T, T0) and not dsum.any()


def test_internal_tide_drag_closed_form():
    """cbot = the mean over the edge's cells of kappa h^2 N / 2: with the case's N, and from the two lowest layers'
    buoyancy step over the distance between their middles."""
    g = gridmod.get(2)
    og = oracle.Grid(g)
    rng = np.random.default_rng(1)
    h2 = rng.uniform(0, 1e4, g.N)
    nb = rng.uniform(1e-4, 2e-3, g.N)
    kap = 2 * math.pi / 1e4
    h = np.stack([np.full(g.N, 100.0), np.full(g.N, 3900.0)])
    T = np.stack([np.full(g.N, 290.0), np.full(g.N, 275.0)])
    S = np.full((2, g.N), 35.0)
    cb = np.zeros(g.E)
    eos = (GRAV, 2e-4, 0.0, 7.6e-4)
    ot.tide_itd(og, 2, h, T, S, h2, kap, eos, cb, nbot=nb)
    want = 0.5 * (0.5 * kap * h2[g.eI] * nb[g.eI] + 0.5 * kap * h2[g.eJ] * nb[g.eJ])
    assert np.allclose(cb, want, rtol=1e-14)
    ot.tide_itd(og, 2, h, T, S, h2, kap, eos, cb)
    Nb = math.sqrt(GRAV * 2e-4 * 15.0 / 2000.0)
    assert np.allclose(cb, 0.5 * (0.5 * kap * Nb * (h2[g.eI] + h2[g.eJ])), rtol=1e-13)


def test_equilibrium_tide():
    """One 4 km layer at rest without rotation under a fixed degree-2 potential V, damped by a linear drag, settles
    to the equilibrium tide: its surface eta = V / g (less V's mean over the area, the water being kept), within 2%
    of V's range at level 4 -- the sign and scale of the tide's force on the layers."""
    g = gridmod.get(4)
    og = oracle.Grid(g)
    core = oracle.Layers(og, 1, grav=GRAV, drag=2e-4, rotation=0.0, floor=1.0)
    V = 0.5 * (3 * np.asarray(g.pos)[:, 0] ** 2 - 1)            # (m^2/s^2: a ~0.5 m^2/s^2 tide, P2 under x)
    core.set_tides(np.ascontiguousarray(V), None, 0.0)
    h = np.full((1, g.N), 4000.0)
    u = np.zeros((1, g.E))
    dt = 600.0
    nfast = max(1, core.substeps(h, dt))
    for _ in range(300):                                     # (36 e-foldings of the drag)
        core.step(h, u, dt, nfast)
    A = np.asarray(g.areaCell)
    eta = h[0] - 4000.0
    want = (V - np.sum(A * V) / np.sum(A)) / GRAV
    assert np.max(np.abs(eta - want)) < 0.02 * (want.max() - want.min())
    assert abs(np.sum(A * eta)) < 1e-9 * np.sum(A) * 4000.0
