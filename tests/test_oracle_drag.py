"""The oracle of the gravity waves' drag (oracle/drag.c), pinned on closed forms and on the reference's own ledgers
(tests/test_world.py:231-244: the kinetic energy a drag removes is the heat it books; the waves keep each column's
momentum), on the CPU. Columns are set up on a level-2 mesh with every cell alike: Perot's coefficients are replaced
by one slot that reads the layer's wind (u = s_k on every edge) in a chosen direction, so each cell's wind per layer
is known exactly."""
import math

import numpy as np
import pytest
from scipy.integrate import quad

import oracle
from oracle import drag as od
from insilica import grid as gridmod
from insilica.orodrag import edge_normals
from insilica.atmosphere import perot_operators

GRAV, P0, KAPPA, RD = 9.80616, 1.0e5, 0.2857142857142857, 287.0


def oro_par(G=0.3, cd=1.0, hnc=0.5, ric=0.25, dt=900.0):
    return np.array([GRAV, P0, KAPPA, RD, G, cd, hnc, ric, dt, 1.0])


def gw_par(bt=0.004, cw=40.0, lam=3.0e5, psrc=31500.0, cmax=99.6, dc=1.2, dt=900.0, fc=0.0):
    return np.array([GRAV, P0, KAPPA, RD, bt, cw, 2 * math.pi / lam, psrc, cmax, dc, dt, fc])


def column(g, mu, theta, speeds, phi, sso=None):
    """Every cell the same column: masses mu (L), theta (L), the wind speed s_k of each layer in the direction phi
    (radians from east); sso (4,) the orography. Returns (og, mu, theta, u, op, nrm, sso)."""
    L, N, E = len(mu), g.N, g.E
    op = np.zeros((N, 6, 2))
    op[:, 0, 0], op[:, 0, 1] = math.cos(phi), math.sin(phi)
    u = np.repeat(np.asarray(speeds, float)[:, None], E, axis=1)
    M = np.repeat(np.asarray(mu, float)[:, None], N, axis=1)
    T = np.repeat(np.asarray(theta, float)[:, None], N, axis=1)
    s = None if sso is None else np.repeat(np.asarray(sso, float)[:, None], N, axis=1)
    return oracle.Grid(g), M, T, u, op, edge_normals(g), s


def heights(mu, theta):
    """Each layer's bottom and top height and its middle's, density and temperature at its middle pressure, from the
    hypsometric equation for layers at their middle's temperature (the ground at 0)."""
    L = len(mu)
    pt = np.concatenate([[0.0], np.cumsum(GRAV * np.asarray(mu))[:-1]])
    pb = pt + GRAV * np.asarray(mu)
    pm = 0.5 * (pt + pb)
    T = np.asarray(theta) * (pm / P0) ** KAPPA
    zb, zt, zm = np.zeros(L), np.zeros(L), np.zeros(L)
    z = 0.0
    for k in range(L - 1, -1, -1):
        Hs = RD * T[k] / GRAV
        zb[k], zm[k] = z, z + Hs * math.log(pb[k] / pm[k])
        zt[k] = z + Hs * math.log(pb[k] / max(pt[k], 1e-3 * pb[k])) if k > 0 else 1e30
        z = zt[k]
    return zb, zt, zm, pm / (RD * T), T


def test_blocked_profile_primitive():
    """blk_int is a primitive of sqrt((A - t) / t) on [0, A] (clamped outside): differences equal the integral."""
    for A in (0.5, 700.0, 2500.0):
        for a, b in ((0.0, 0.3 * A), (0.1 * A, 0.9 * A), (0.5 * A, A)):
            want = quad(lambda t: math.sqrt((A - t) / t), a, b, limit=200)[0]
            assert abs(od.blk_int(b, A) - od.blk_int(a, A) - want) <= 1e-7 * A
    assert od.blk_int(-5.0, 10.0) == od.blk_int(0.0, 10.0) and od.blk_int(12.0, 10.0) == od.blk_int(10.0, 10.0)


def test_edge_normals_and_perot(grids):
    """The edges' normals are unit vectors in each cell's tangent plane pointing from the first cell to the second,
    and agree with the edge's own normal to the mesh's resolution; Perot's reconstruction of a solid-body wind
    reproduces it at the cells."""
    g = grids(4)
    nrm = edge_normals(g)
    assert np.allclose(nrm[:, 0] ** 2 + nrm[:, 1] ** 2, 1, atol=1e-12)
    assert np.allclose(nrm[:, 2] ** 2 + nrm[:, 3] ** 2, 1, atol=1e-12)
    pos = np.asarray(g.pos)
    for col, cell in ((0, g.eI), (2, g.eJ)):
        lon, lat = np.arctan2(pos[cell, 1], pos[cell, 0]), np.arcsin(pos[cell, 2])
        east = np.stack([-np.sin(lon), np.cos(lon), 0 * lon], 1)
        north = np.stack([-np.sin(lat) * np.cos(lon), -np.sin(lat) * np.sin(lon), np.cos(lat)], 1)
        vec = nrm[:, col, None] * east + nrm[:, col + 1, None] * north
        assert np.all(np.sum(vec * (pos[g.eJ] - pos[g.eI]), 1) > 0)
        assert np.max(np.linalg.norm(vec - g.ne, axis=1)) < 0.05
    # a solid-body wind about the z axis: u = 20 cos(lat) east
    R = getattr(g, 'radius', 6.371e6)
    lon_e, lat_e = np.arctan2(g.xm[:, 1], g.xm[:, 0]), np.arcsin(np.clip(g.xm[:, 2], -1, 1))
    east_e = np.stack([-np.sin(lon_e), np.cos(lon_e), 0 * lon_e], 1)
    ue = 20 * np.cos(lat_e) * np.sum(east_e * g.ne, 1)
    op = perot_operators(g)
    ce = np.where(g.cellEdge >= 0, g.cellEdge, 0)
    w = np.einsum('imk,im->ik', op, np.where(g.cellEdge >= 0, ue[ce], 0.0))
    lat_c = np.arcsin(pos[:, 2])
    assert np.max(np.abs(w[:, 0] - 20 * np.cos(lat_c))) < 0.2 and np.max(np.abs(w[:, 1])) < 0.2
    assert R > 0


def ke(g, h, u):
    he = 0.5 * (h[:, g.eI] + h[:, g.eJ])
    return float(np.sum(g.le * g.de * np.sum(0.5 * he * u * u, 0)))


def test_no_orography_or_calm_changes_nothing():
    g = gridmod.get(2)
    og, mu, th, u, op, nrm, sso = column(g, [2000.0, 4000.0, 4000.0], [340.0, 310.0, 290.0], [10, 8, 6], 0.3,
                                          sso=[0.5, 0.02, 0.5, 0.2])            # mu 0.5 m <= mu_min
    u1, d1, beta, F = od.oro_drag(og, mu, th, u, op, nrm, sso, oro_par(), 900.0, np.zeros(g.E))
    assert np.array_equal(u1, u) and not d1.any() and not beta.any() and not F.any()
    og, mu, th, u, op, nrm, sso = column(g, [2000.0, 4000.0, 4000.0], [340.0, 310.0, 290.0], [0.0, 0.0, 0.05], 0.3,
                                          sso=[600.0, 0.02, 0.5, 0.2])          # low-level wind < 0.1 m/s
    u1, d1, beta, F = od.oro_drag(og, mu, th, u, op, nrm, sso, oro_par(), 900.0, np.zeros(g.E))
    assert np.array_equal(u1, u) and not d1.any() and not beta.any() and not F.any()


def test_launched_stress_goes_to_the_top_without_saturation():
    """No blocking (H_nc beyond reach) and no saturation (Ri_c tiny): the top layer takes the whole launched stress,
    Lott & Miller's G rho N |U| sigma |D| mu for an effective height 2 mu over the low-level layer (here the bottom
    layer alone), against the low-level wind for isotropic hills (gamma = 1, |D| = B = C = 0.78)."""
    g = gridmod.get(2)
    mu_l, th = [2000.0, 4000.0, 6000.0], [345.0, 315.0, 295.0]
    ms, sig, U = 200.0, 0.03, 20.0                         # (N 2 mu / U well below 1: the wave's Ri stays positive)
    og, mu, thm, u, op, nrm, sso = column(g, mu_l, th, [25.0, 22.0, U], 0.4, sso=[ms, sig, 1.0, 0.9])
    par = oro_par(G=0.3, hnc=1e9, ric=1e-9)
    u1, d1, beta, F = od.oro_drag(og, mu, thm, u, op, nrm, sso, par, 900.0, np.zeros(g.E))
    zb, zt, zm, rho, T = heights(mu_l, th)
    assert zt[2] > 2 * ms                                  # the bottom layer holds both mu and 2 mu
    n2 = GRAV * (th[1] - th[2]) / (0.5 * (th[1] + th[2]) * (zm[1] - zm[2]))
    stress = 0.3 * rho[2] * math.sqrt(n2) * U * sig * 0.78 * ms
    fx, fy = F[0, 0], F[1, 0]
    assert abs(math.hypot(fx, fy) / stress - 1) < 1e-12
    assert abs(fx / math.hypot(fx, fy) + math.cos(0.4)) < 1e-12 and abs(fy / math.hypot(fx, fy) + math.sin(0.4)) < 1e-12
    assert not F[2:].any() and not beta.any()


def test_anisotropic_hills_turn_the_stress():
    """Hills elongated across a wind at 45 degrees to their axis turn the stress by atan((B - C) / (B + C)) from the
    wind (Baines & Palmer's D = (B - (B - C) sin^2 psi, (B - C) sin psi cos psi))."""
    g = gridmod.get(2)
    gam, axis, phi = 0.3, 0.2, 0.2 + math.pi / 4       # wind 45 degrees from the axis: psi = -pi/4
    og, mu, thm, u, op, nrm, sso = column(g, [2000.0, 4000.0, 6000.0], [345.0, 315.0, 295.0], [12.0, 9.0, 7.0], phi,
                                          sso=[400.0, 0.03, gam, axis])
    u1, d1, beta, F = od.oro_drag(og, mu, thm, u, op, nrm, sso, oro_par(hnc=1e9, ric=1e-9), 900.0, np.zeros(g.E))
    B, C = 1 - 0.18 * gam - 0.04 * gam * gam, 0.48 * gam + 0.3 * gam * gam
    psi = axis - phi
    turn = math.atan2((B - C) * math.sin(psi) * math.cos(psi), B - (B - C) * math.sin(psi) ** 2)
    ang = math.atan2(-F[1, 0], -F[0, 0])
    assert abs(ang - (phi + turn)) < 1e-12


def test_blocking_height_and_blocked_drag():
    """One thick bottom layer under uniform stratification and wind: the flow is blocked below
    Z_blk = 2 mu - H_nc U / N (the N / U integral down from 2 mu reaching H_nc), and the layer's implicit factor is
    dt C_d B (sigma / 4 mu) |U| (1 / dz) int_0^Z_blk sqrt((Z_blk - z) / (z + mu)) dz (isotropic hills, the wind
    along the drag's direction; dz the layer's depth)."""
    g = gridmod.get(2)
    mu_l, th = [2000.0, 4000.0, 6000.0], [345.0, 315.0, 295.0]
    ms, sig, U, dt, hnc = 600.0, 0.02, 3.0, 900.0, 0.5
    og, mu, thm, u, op, nrm, sso = column(g, mu_l, th, [12.0, 9.0, U], -0.7, sso=[ms, sig, 1.0, 0.0])
    u1, d1, beta, F = od.oro_drag(og, mu, thm, u, op, nrm, sso, oro_par(hnc=hnc), dt, np.zeros(g.E))
    zb, zt, zm, rho, T = heights(mu_l, th)
    N2 = GRAV * (th[1] - th[2]) / (0.5 * (th[1] + th[2]) * (zm[1] - zm[2]))
    zblk = 2 * ms - hnc * U / math.sqrt(N2)
    assert 0 < zblk < 2 * ms
    I = quad(lambda z: math.sqrt((zblk - z) / (z + ms)), 0.0, zblk, limit=200)[0]
    want = dt * 1.0 * 0.78 * sig / (4 * ms) * U * I / (zt[2] - zb[2])
    assert abs(beta[2, 0] / want - 1) < 1e-9
    assert not beta[:2].any()


@pytest.mark.parametrize('n', [3, 4])
def test_oro_energy_booked(grids, n):
    """The reference's ledger on a seeded world: the kinetic energy the mountains' drag removes is the heat it books
    (<= 1e-13 of the kinetic energy), and it removes some."""
    from insilica import cases
    g = grids(n)
    built = cases.coupled_stack(g, seed=3)
    og = oracle.Grid(g)
    sso = cases.synthetic_sso(g, seed=3, land=built['land'])
    op, nrm = perot_operators(g), edge_normals(g)
    h, th, u = built['h'], built['theta'], 3 * built['u']
    u1, d1, beta, F = od.oro_drag(og, h, th, u, op, nrm, sso, oro_par(), 900.0, np.zeros(g.E))
    k0, k1 = ke(g, h, u), ke(g, h, u1)
    booked = float(np.sum(g.le * g.de * d1))
    assert k0 - k1 > 1e-8 * k0 and beta.any() and F.any()
    assert abs((k0 - k1) - booked) <= 1e-13 * k0


@pytest.mark.parametrize('n', [3, 4])
def test_waves_keep_momentum_and_book_energy(grids, n):
    """The reference's ledgers for the waves: each edge's column momentum sum_k h_e u is kept (<= 1e-13 of the
    gross), the kinetic energy removed is the heat booked (<= 1e-13), and the waves act."""
    from insilica import cases
    g = grids(n)
    built = cases.coupled_stack(g, seed=4)
    og = oracle.Grid(g)
    op, nrm = perot_operators(g), edge_normals(g)
    h, th, u = built['h'], built['theta'], 4 * built['u']
    par = gw_par(psrc=0.6 * GRAV * float(np.mean(h[0] + h[1])))      # the source inside the middle layer
    u1, d1, F = od.wave_drag(og, h, th, u, op, nrm, par, 900.0, np.zeros(g.E))
    he = 0.5 * (h[:, g.eI] + h[:, g.eJ])
    dm = np.sum(he * (u1 - u), 0)
    gross = np.sum(g.le * g.de * np.sum(he * np.abs(u), 0))
    assert F.any() and np.sum(np.abs(u1 - u)) > 0
    assert float(np.sum(g.le * g.de * np.abs(dm))) <= 1e-13 * gross
    assert np.max(np.abs(F[0::2].sum(0))) <= 1e-13 * np.max(np.abs(F))
    assert not F[1::2].any()
    k0, k1 = ke(g, h, u), ke(g, h, u1)
    assert abs((k0 - k1) - float(np.sum(g.le * g.de * d1))) <= 1e-13 * k0


def test_resting_air_feels_no_net_wave_force():
    """At rest the spectrum is symmetric about the source's wind (phase speeds -cmax .. cmax by dc): eastward and
    westward waves end in the same layers and cancel there (to round-off of the gross flux)."""
    g = gridmod.get(2)
    og, mu, thm, u, op, nrm, _ = column(g, [2000.0, 3000.0, 6000.0], [400.0, 320.0, 295.0], [0.0, 0.0, 0.0], 0.0)
    par = gw_par(psrc=0.5 * GRAV * 5000.0, dt=60.0)
    u1, d1, F = od.wave_drag(og, mu, thm, u, op, nrm, par, 60.0, np.zeros(g.E))
    assert np.max(np.abs(F)) <= 1e-15 * par[4] and np.max(np.abs(u1)) <= 1e-12


def test_a_layer_takes_no_more_than_brings_it_to_the_phase_speed():
    """One phase speed (c = 0, cmax 0) against a source wind U launches its whole flux B_t westward; a strong wind
    carries it unbroken to the top layer, whose deposit is capped at (c - U) mu / dt -- the wind brought exactly to
    the wave's speed at the cell -- and the source layer keeps the rest (its force equal and opposite)."""
    g = gridmod.get(2)
    mu_l, th, U = [1000.0, 3000.0, 6000.0], [600.0, 330.0, 295.0], 25.0
    og, mu, thm, u, op, nrm, _ = column(g, mu_l, th, [U, U, U], 0.0)
    dt = 900.0
    par = gw_par(bt=100.0, cmax=0.0, dc=1.2, lam=3.0e5, psrc=GRAV * (1000.0 + 1500.0), dt=dt)
    u1, d1, F = od.wave_drag(og, mu, thm, u, op, nrm, par, dt, np.zeros(g.E))
    cap = (0.0 - U) * mu_l[0] / dt
    assert abs(F[0, 0] / cap - 1) < 1e-14 and abs(F[2, 0] + cap) < 1e-12 * abs(cap) and F[4, 0] == 0
