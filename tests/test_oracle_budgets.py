"""Pins of the oracle's exchange between layers, interfacial stress, heat return and one-layer SWE solver (CPU only).

Each check is computed here independently of the oracle's formulas, so that a dropped term, a wrong sign or a
swapped index in the restatement fails one of them:

  * the exchange (reference: kernels/layers.cu:454-565, layers.py:340-358; the reference's own budget check is
    tests/test_world.py:37-60): moving up to 40% of each donor above the floor between layers keeps the column's
    volume, each tracer's mass and the column's momentum sum_e l_e d_e sum_k h_e,k u_k to round-off, keeps M = h A,
    and for air books the kinetic energy it removes, 0.5 sum_k h_e u^2 before minus after (computed here from the
    arrays before and after), in diss, which vmix_heat (layers.cu:599-621) returns as heat: sum_i A_i sum_k h_k
    pibar_k dtheta_k, with pibar_k the layer's mean Exner function from the closed-form pressure integral, equals
    sum_e l_e d_e diss_e;
  * the interfacial stress (layers.cu:575-593): two layers relax as the ODE pair du_a/dt = -tau/h_a, du_b/dt =
    tau/h_b with tau = c (u_a - u_b) (linear) or cd |u_a - u_b| (u_a - u_b) (quadratic), integrated here by RK4 with
    small steps; the column's momentum is kept and diss is the kinetic energy removed;
  * the one-layer SWE step (reference: swe.py:124-132, kernels/fused.cu:27-120): Williamson 2 steady within 1e-3 at
    n = 4 and falling at least x3 to n = 5 (reference SPEC.md:207); mass exact; Williamson 5's energy drift < 1e-4 over
    15 days; a dam break onto nearly dry ground stays non-negative (the positivity limiter) with its mass exact; drag
    changes a steady flow's wind by the RK3 polynomial of e^(-r dt) and del^4 by -nu4 dt del^4 u, the Laplacian itself
    pinned on spherical harmonics (lap grad Y_l = -l (l + 1) / a^2 grad Y_l);
  * the floor limiter: two layers over a ridge higher than the lower one is deep keep every budget and every
    thickness at or above the floor, with the lower layer held at the floor over the crest.
"""
import math

import numpy as np
import pytest

import oracle
from insilica import cases


def edge_h(g, h):
    return 0.5 * (h[:, g.eI] + h[:, g.eJ])


def column_momentum(g, h, u):
    he = edge_h(g, h)
    lede = g.le * g.de
    return float(np.sum(lede * np.sum(he * u, axis=0))), float(np.sum(lede * np.sum(he * np.abs(u), axis=0)))


def kinetic(g, h, u):
    """0.5 sum_e l_e d_e sum_k h_e,k u_k^2: the exchange's own energy measure."""
    return float(np.sum(g.le * g.de * 0.5 * np.sum(edge_h(g, h) * u * u, axis=0)))


def random_exchange(g, h, floor, rng, frac=0.4):
    """wx (L-1, N): each interface moves up to frac of its donor's thickness above the floor, up or down."""
    L = h.shape[0]
    wx = np.zeros((L - 1, g.N))
    for k in range(L - 1):
        up, down = frac * (h[k + 1] - floor), frac * (h[k] - floor)
        wx[k] = np.where(rng.random(g.N) < 0.5, up, -down) * rng.random(g.N) / (L - 1)
    return wx


def two_layer_state(g, rng):
    """Two Boussinesq layers over a ridge with winds and a tracer."""
    b = 900.0 * np.exp(-(np.abs(np.angle(np.exp(1j * g.lon))) / math.radians(8)) ** 2) * np.cos(g.lat) ** 2
    lower = np.maximum(500.0 - b, 1.0)
    upper = 2500.0 - b - lower
    h = np.stack([upper, lower])
    u = rng.standard_normal((2, g.E)) * 0.3
    psi = rng.random((2, g.N))
    return h, u, psi, b


@pytest.mark.parametrize('seq', [False, True])
def test_exchange_budgets_boussinesq(grids, seq):
    g = grids(4)
    og = oracle.Grid(g)
    rng = np.random.default_rng(4)
    h, u, psi, b = two_layer_state(g, rng)
    lay = oracle.Layers(og, 2, b=b, gprime=0.02, floor=1.0)
    M = h * g.areaCell
    wx = random_exchange(g, h, 1.0, rng)
    v0, t0 = np.sum(M), np.sum(M * psi)
    m_top, m_low = np.sum(M[0]), np.sum(M[1])
    p0, gross = column_momentum(g, h, u)
    moved = float(np.sum(np.abs(wx) * g.areaCell) / np.sum(M[1]))
    assert moved > 0.05                       # a real exchange: several percent of the lower layer
    lay.exchange(wx, h, u, M, psi, seq=seq)
    assert abs(np.sum(M) - v0) / v0 <= 1e-13
    assert abs(np.sum(M * psi) - t0) / t0 <= 1e-13
    p1, _ = column_momentum(g, h, u)
    assert abs(p1 - p0) / gross <= 1e-13
    assert np.max(np.abs(M / g.areaCell - h)) / np.max(h) <= 1e-13
    # each layer's volume changes by what crossed its interface (wx > 0 moves the lower layer's water up)
    up = float(np.sum(wx[0] * g.areaCell))
    assert abs(np.sum(M[0]) - (m_top + up)) <= 1e-13 * v0
    assert abs(np.sum(M[1]) - (m_low - up)) <= 1e-13 * v0
    assert np.min(h) >= 1.0 - 1e-9


def air_state(g, rng, L=3):
    mu = np.stack([np.full(g.N, 2.2e4), np.full(g.N, 3.8e4), np.full(g.N, 4.0e4)])[:L] / 9.80616
    mu = mu * (1 + 0.1 * rng.random((L, g.N)))
    theta = np.stack([380.0 + 5 * rng.random(g.N), 320.0 + 5 * rng.random(g.N), 290.0 + 5 * rng.random(g.N)])[:L]
    u = rng.standard_normal((L, g.E)) * 10.0
    psi = rng.random((L, g.N))
    return mu, u, theta, psi


def pibar(mu, grav, cp, kappa, p0=1.0e5):
    """Each layer's mass-mean Exner function cp (p/p0)^kappa from the pressure integral over the layer."""
    p = np.cumsum(grav * mu, axis=0)
    pt = np.vstack([np.zeros((1, mu.shape[1])), p[:-1]])
    return cp * (p ** (kappa + 1) - pt ** (kappa + 1)) / ((kappa + 1) * (p - pt) * p0 ** kappa)


@pytest.mark.parametrize('thin_wind', [0.0, 3000.0])
def test_exchange_air_books_kinetic_energy_as_heat(grids, thin_wind):
    g = grids(4)
    og = oracle.Grid(g)
    rng = np.random.default_rng(11)
    mu, u, theta, psi = air_state(g, rng)
    grav, cp, kap = 9.80616, 1004.6, 287.04 / 1004.6
    floor = 10.0
    if thin_wind:
        mu[1, : g.N // 3] = floor * (1 + rng.random(g.N // 3))      # a middle layer nearly empty over a third
    lay = oracle.Layers(og, 3, grav=grav, gprime=[0.0, 0.0], floor=floor, air=dict(cp=cp, kappa=kap, rd=287.04),
                        thin_wind=thin_wind)
    M = mu * g.areaCell
    wx = random_exchange(g, mu, floor, rng)
    h0, u0, th0 = mu.copy(), u.copy(), theta.copy()
    ke0 = kinetic(g, h0, u0)
    p0, gross = column_momentum(g, h0, u0)
    lay.exchange(wx, mu, u, M, psi, theta=theta)
    diss = lay.diss.copy()
    lede = g.le * g.de
    ke1 = kinetic(g, mu, u)
    assert np.min(diss) >= -1e-9 * np.max(np.abs(diss))            # mixing winds only removes kinetic energy
    assert ke0 - ke1 > 0
    assert abs((ke0 - ke1) - float(np.sum(lede * diss))) <= 1e-10 * ke0
    p1, _ = column_momentum(g, mu, u)
    assert abs(p1 - p0) / gross <= 1e-13
    # the heat returned, with the new masses' Exner functions, is the kinetic energy removed
    heat = float(np.sum(g.areaCell * np.sum(mu * pibar(mu, grav, cp, kap) * (theta - th0), axis=0)))
    assert abs(heat - float(np.sum(lede * diss))) <= 1e-11 * abs(heat)
    # every layer of a column warms by the same amount per kilogram (to the rounding of theta's small change)
    dq = pibar(mu, grav, cp, kap) * (theta - th0)
    hot = np.abs(dq[0]) > 0
    assert hot.sum() > g.N // 2
    assert np.allclose(dq[0][hot], dq[2][hot], rtol=1e-9, atol=0)
    assert np.allclose(dq[1][hot], dq[2][hot], rtol=1e-9, atol=0)


def rk4_pair(ha, hb, ua, ub, c, cd, dt, steps=4000):
    """The two-layer interfacial stress as an ODE: du_a/dt = -tau/h_a, du_b/dt = tau/h_b, tau = (c + cd|du|) du."""
    s = dt / steps

    def f(a, b):
        du = a - b
        tau = (c + cd * np.abs(du)) * du
        return -tau / ha, tau / hb
    for _ in range(steps):
        k1 = f(ua, ub)
        k2 = f(ua + 0.5 * s * k1[0], ub + 0.5 * s * k1[1])
        k3 = f(ua + 0.5 * s * k2[0], ub + 0.5 * s * k2[1])
        k4 = f(ua + s * k3[0], ub + s * k3[1])
        ua = ua + s / 6 * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0])
        ub = ub + s / 6 * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1])
    return ua, ub


@pytest.mark.parametrize('c,cd', [(2e-2, 0.0), (0.0, 3e-3)])
def test_interfacial_stress_solves_its_ode(grids, c, cd):
    g = grids(3)
    og = oracle.Grid(g)
    rng = np.random.default_rng(5)
    h = np.stack([200 + 100 * rng.random(g.N), 50 + 50 * rng.random(g.N)])
    u = rng.standard_normal((2, g.E))
    dt = 3600.0
    he = edge_h(g, h)
    want_a, want_b = rk4_pair(he[0], he[1], u[0].copy(), u[1].copy(), c, cd, dt)
    uu, diss = u.copy(), np.zeros(g.E)
    oracle.lay_vmix(og, 2, np.ascontiguousarray(h), c, dt, uu, diss, cd)
    assert np.max(np.abs(uu[0] - want_a)) <= 1e-9 and np.max(np.abs(uu[1] - want_b)) <= 1e-9
    lede = g.le * g.de
    p0, gross = column_momentum(g, h, u)
    assert abs(column_momentum(g, h, uu)[0] - p0) / gross <= 1e-14
    assert abs(kinetic(g, h, u) - kinetic(g, h, uu) - float(np.sum(lede * diss))) <= 1e-11 * kinetic(g, h, u)
    # and the stress acts: the shear decays
    assert np.max(np.abs(uu[0] - uu[1])) < np.max(np.abs(u[0] - u[1]))


def test_interfacial_stress_three_layers_budgets(grids):
    g = grids(3)
    og = oracle.Grid(g)
    rng = np.random.default_rng(6)
    h = 20 + 300 * rng.random((3, g.N))
    u = rng.standard_normal((3, g.E)) * 5
    uu, diss = u.copy(), np.zeros(g.E)
    oracle.lay_vmix(og, 3, np.ascontiguousarray(h), 1e-2, 1800.0, uu, diss, 1e-3)
    p0, gross = column_momentum(g, h, u)
    assert abs(column_momentum(g, h, uu)[0] - p0) / gross <= 1e-14
    assert np.min(diss) >= 0
    assert abs(kinetic(g, h, u) - kinetic(g, h, uu) - float(np.sum(g.le * g.de * diss))) <= 1e-11 * kinetic(g, h, u)


# ---- the one-layer SWE solver ----

def swe_run(g, case, days, **kw):
    og = oracle.Grid(g)
    h, b, u = cases.williamson(g, case)
    h, u = h[None].copy(), u[None].copy()
    dt = cases.default_dt(g.n)
    h0, u0 = h.copy(), u.copy()
    for _ in range(int(round(days * 86400 / dt))):
        oracle.swe_step(og, 1, h, u, dt, b=b, **kw)
    return h0, u0, h, u, b


def l2rel(area, a, b):
    return float(np.sqrt(np.sum(area * (a - b) ** 2) / np.sum(area * b ** 2)))


def test_swe_williamson2_steady_and_converging(grids):
    e = {}
    for n in (4, 5):
        g = grids(n)
        h0, _, h, u, _ = swe_run(g, 'w2', 5)
        e[n] = l2rel(g.areaCell, h[0], h0[0])
        if n == 4:
            m0, m1 = oracle.dsum(h0[0], g.areaCell), oracle.dsum(h[0], g.areaCell)
            assert abs(m1 - m0) / m0 <= 1e-13
    assert e[4] < 1e-3, e
    assert e[4] / e[5] >= 3.0, e


def swe_energy(g, h, u, b, grav=9.80616):
    """sum_i A_i (h_i K_i + g (h_i^2 / 2 + h_i b_i)), K_i the cell's kinetic energy from its edges (as swe_cell forms
    it: sum over the cell's edges of l d u^2 / 4, over the cell's area)."""
    ce = g.cellEdge
    ok = ce >= 0
    cs = np.where(ok, ce, 0)
    K = np.sum(np.where(ok, g.le[cs] * g.de[cs] * 0.25 * u[cs] ** 2, 0.0), axis=1) / g.areaCell
    return float(np.sum(g.areaCell * (h * K + grav * (0.5 * h * h + h * b))))


def test_swe_williamson5_energy(grids):
    g = grids(4)
    h0, u0, h, u, b = swe_run(g, 'w5', 15)
    e0, e1 = swe_energy(g, h0[0], u0[0], b), swe_energy(g, h[0], u[0], b)
    assert abs(e1 - e0) / e0 < 1e-4
    assert np.all(np.isfinite(h)) and np.min(h) > 0


def test_swe_positivity_limiter_dam_break(grids):
    g = grids(4)
    og = oracle.Grid(g)
    h = np.where(g.lon > 0, 100.0, 0.5)[None].copy()          # half the sphere nearly dry (PV needs h > 0)
    u = np.zeros((1, g.E))
    dt = cases.default_dt(4)
    m0 = oracle.dsum(h[0], g.areaCell)
    lowest = 0.0
    for _ in range(30):
        oracle.swe_step(og, 1, h, u, dt)
        lowest = min(lowest, float(np.min(h)))
    assert lowest >= 0.0
    assert abs(oracle.dsum(h[0], g.areaCell) - m0) / m0 <= 1e-13
    # the water has spread onto the dry half
    assert float(np.sum(g.areaCell * h[0] * (g.lon <= 0))) > 0.01 * m0


def test_swe_drag_and_del4(grids):
    """Drag and del^4 enter the SWE tendency as -r u - nu4 del^4 u0 (del^4 from the step's start). On Williamson 2
    (steady: the other terms balance), one RK3 step with drag alone changes the wind by (1 - x + x^2/2 - x^3/6 - 1) u,
    x = r dt (the scheme's polynomial of e^-x), measured as its projection on u; with nu4 alone by -nu4 dt del^4 u0,
    del^4 = lap(lap(.)) being the layered core's operator, whose Laplacian is pinned here on spherical harmonics: for
    u = grad Y_l, lap(u) = -l (l + 1) / a^2 u up to the mesh's error."""
    from insilica.constants import R_PLANET
    g = grids(4)
    og = oracle.Grid(g)
    h, b, u = cases.williamson(g, 'w2')
    dt = cases.default_dt(4)

    def swe(r, v):
        hs, us = h[None].copy(), u[None].copy()
        oracle.swe_step(og, 1, hs, us, dt, drag=r, nu4=v)
        return us[0]
    u0 = swe(0.0, 0.0)
    r = 2e-5
    x = r * dt
    w = g.le * g.de
    proj = np.sum(w * (swe(r, 0.0) - u0) * u) / np.sum(w * u * u * (-x + x * x / 2 - x ** 3 / 6))
    assert abs(proj - 1) < 1e-3, proj
    lay = oracle.Layers(og, 1)
    for l, phi in ((1, g.pos[:, 2]), (2, g.pos[:, 0] * g.pos[:, 1])):
        ue = (phi[g.eJ] - phi[g.eI]) / g.de
        out = np.zeros((1, g.E))
        lay.lap(ue[None].copy(), out)
        want = -l * (l + 1) / R_PLANET ** 2 * ue
        assert np.linalg.norm(out[0] - want) / np.linalg.norm(want) < (2e-3 if l == 1 else 0.05)
    L1, L2 = np.zeros((1, g.E)), np.zeros((1, g.E))
    lay.lap(u[None].copy(), L1)
    lay.lap(L1, L2)
    nu = 1.0e15
    got, pred = swe(0.0, nu) - u0, -nu * dt * L2[0]
    big = np.abs(pred) > 0.1 * np.max(np.abs(pred))
    assert abs(np.median(got[big] / pred[big]) - 1) < 1e-3
    assert np.max(np.abs(got - pred)) < 0.02 * np.max(np.abs(pred))


# ---- the floor limiter on an outcropping layer (reference: tests/test_world.py:20-36, the 'outcrop' preset) ----

def test_outcrop_conserves_and_keeps_the_floor(grids):
    """Two layers over a ridge higher than the lower layer is deep, 4 days at n = 4 (dt 1 h): per-layer volume and
    tracer budgets to round-off, every thickness at or above the floor, and the lower layer held at the floor somewhere
    (the limiter acting: without it the layer would go negative over the crest)."""
    from insilica.ocean import visc4
    g = grids(4)
    og = oracle.Grid(g)
    h, b, u = cases.outcrop(g)
    floor, dt = 1.0, 3600.0
    lay = oracle.Layers(og, 2, b=b, gprime=0.02, stress=[0.0, 2.5e-4], nu4=visc4(g, 0.5), floor=floor)
    nf = oracle.nfast_for(lay, h, dt)
    A = g.areaCell
    psi = np.tile(cases.cosine_bell(g), (2, 1))
    M, M1, Fm = h * A, np.zeros((2, g.N)), np.zeros((2, g.E))
    v0 = [oracle.dsum(h[k], A) for k in range(2)]
    t0 = [oracle.dsum(M[k], psi[k]) for k in range(2)]
    low, at_floor = np.inf, 0
    for _ in range(96):
        lay.step(h, u, dt, nf, Fm)
        oracle.mass_update(og, 2, M, Fm, M1)
        oracle.transport(og, 2, M, M1, Fm, psi)
        M[...] = M1
        low = min(low, float(np.min(h)))
        at_floor = max(at_floor, int(np.sum(h[1] < floor + 1e-6)))
    for k in range(2):
        assert abs(oracle.dsum(h[k], A) - v0[k]) / v0[k] <= 1e-13
        assert abs(oracle.dsum(M[k], psi[k]) - t0[k]) / t0[k] <= 1e-13
    assert np.all(np.isfinite(h)) and np.all(np.isfinite(u))
    assert low >= floor - 1e-9, low
    assert at_floor > 0
