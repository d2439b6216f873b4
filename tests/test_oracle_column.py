"""The Held-Suarez column and surface-exchange oracle pinned on the CPU (oracle/column.c).

- The relaxation reaches Held & Suarez's published equilibrium (Held & Suarez 1994, T_eq) at each layer's
  representative pressure, the one whose Exner function is the layer's mass-mean (computed here from the pressure
  integral), and relaxes at their published rate profile, the layer mean of max(0, (sigma - 0.7) / 0.3) by quadrature.
- The crossing conserves the column's energy (stored = heating x dt to rounding); a heated bottom layer sends up the
  triangle's share mu dth / (2 (theta* - theta)); a heated middle layer sends up mu dth f_b with f_b the edge density
  of the truncated exponential whose mean is the layer's theta (its shape solved here by Brent's method).
- The surface layer in neutral air over land is the log law's C_D = (k / ln(z_r / z0))^2; stable air exchanges less,
  unstable more.
- A solid-body wind's speed at cells from the edges' kinetic energy and from Perot's reconstruction; gas sources.
"""
import math

import numpy as np
import pytest
from scipy import integrate, optimize

from oracle import column as ocol

CP, KAPPA, P0, GRAV = 1004.6, 287.04 / 1004.6, 1.0e5, 9.80616


def par(L, **kw):
    from insilica.atmosphere import par_array
    stars = kw.pop('theta_stars', None)
    if stars is None:
        stars = [330.0, 300.0][-(L - 1):] if L <= 3 else list(np.linspace(360.0, 300.0, L - 1))
        if 'theta_star' in kw:                           # the lowest interface's (<= 0: no crossing)
            stars = stars[:-1] + [kw['theta_star']]
    return par_array(L, dict(kw, theta_stars=list(stars)))


def hs_teq(p, lat):
    """Held & Suarez (1994): T_eq = max(200, [315 - 60 sin^2 - 10 ln(p / p0) cos^2] (p / p0)^kappa)."""
    s2 = math.sin(lat) ** 2
    return max(200.0, (315.0 - 60.0 * s2 - 10.0 * math.log(p / P0) * (1 - s2)) * (p / P0) ** KAPPA)


def layer_pressures(mu):
    pb = np.cumsum(GRAV * mu)
    pt = np.concatenate([[0.0], pb[:-1]])
    # the mass-mean Exner function cp (p / p0)^kappa over [pt, pb], and the pressure that has it
    pib = CP * (pb ** (KAPPA + 1) - pt ** (KAPPA + 1)) / ((KAPPA + 1) * P0 ** KAPPA * (pb - pt))
    pbar = (pib / CP) ** (1 / KAPPA) * P0
    return pt, pb, pib, pbar


def column(L=3, lat=0.6):
    mu = np.array([2500.0, 3500.0, 4200.0][:L] if L <= 3 else np.full(L, 10000.0 / L))
    th = np.linspace(330.0, 285.0, L)
    return mu[:, None], th[:, None], np.array([lat])


def test_relaxes_to_held_suarez_equilibrium():
    L = 3
    mu, th, lat = column(L)
    p = par(L, theta_star=-1.0)                          # no crossing
    hs = np.array([1 / (40 * 86400.0), 1 / (4 * 86400.0)])
    th1, wx, qnet, stored = ocol.hs_column(p, hs, lat, mu, th, dt=1e12)
    _, _, pib, pbar = layer_pressures(mu[:, 0])
    want = np.array([hs_teq(pbar[k], lat[0]) * CP / pib[k] for k in range(L)])
    assert np.allclose(th1[:, 0], want, rtol=1e-12)
    assert np.all(wx == 0)


def test_relaxation_rate_profile():
    L = 4
    mu, th, lat = column(L, lat=0.3)
    p = par(L, theta_star=-1.0)
    hs = np.array([1 / (40 * 86400.0), 1 / (4 * 86400.0)])
    dt = 900.0
    th1, _, qnet, _ = ocol.hs_column(p, hs, lat, mu, th, dt)
    pt, pb, pib, pbar = layer_pressures(mu[:, 0])
    ps = pb[-1]
    c2 = math.cos(lat[0]) ** 2
    for k in range(L):
        st, sb = pt[k] / ps, pb[k] / ps
        f = integrate.quad(lambda s: max(0.0, (s - 0.7) / 0.3), st, sb, points=[0.7])[0] / (sb - st)
        kt = hs[0] + (hs[1] - hs[0]) * f * c2 * c2
        theq = hs_teq(pbar[k], lat[0]) * CP / pib[k]
        want = th[k, 0] + (theq - th[k, 0]) * (1 - math.exp(-kt * dt))
        assert th1[k, 0] == pytest.approx(want, rel=1e-12, abs=1e-12)
    assert qnet[0] == pytest.approx(np.sum(mu[:, 0] * pib * (th1[:, 0] - th[:, 0])) / dt, rel=1e-10)


@pytest.mark.parametrize('L', [2, 3, 5])
def test_crossing_conserves_energy(L):
    rng = np.random.default_rng(L)
    N = 64
    mu = rng.uniform(1500.0, 4000.0, (L, N))
    th = np.sort(rng.uniform(270.0, 360.0, (L, N)), axis=0)[::-1].copy()
    lat = rng.uniform(-1.4, 1.4, N)
    p = par(L, theta_star=300.0)
    hs = np.array([1 / 86400.0, 1 / 7200.0])
    th1, wx, qnet, stored = ocol.hs_column(p, hs, lat, mu, th, dt=3600.0)
    assert np.any(wx != 0)
    assert np.allclose(stored, qnet * 3600.0, rtol=1e-10, atol=1e-6 * np.abs(stored).max())
    # masses after the moves stay above the floor
    mu1 = mu.copy()
    for j in range(L - 1):
        mu1[j] += wx[j]
        mu1[j + 1] -= wx[j]
    assert np.all(mu1 >= p[11] - 1e-9)


def r_of(s):
    return 0.5 + s / 12 if abs(s) < 1e-8 else 1 - 1 / s + 1 / math.expm1(s)


def test_crossing_amounts():
    """Heated bottom layer below theta*: mu dth / (2 (theta* - theta_before)); heated middle layer: mu dth f_b."""
    L = 3
    ts_lo, ts_hi = 300.0, 330.0
    p = par(L, theta_stars=[ts_hi, ts_lo])
    mu = np.array([[2500.0], [3500.0], [4200.0]])
    th = np.array([[200.0], [305.0], [250.0]])            # all three below their equilibria: all heated
    lat = np.array([0.1])
    hs = np.array([1 / 86400.0, 1 / 86400.0])
    dt = 600.0
    th1, wx, _, _ = ocol.hs_column(p, hs, lat, mu, th, dt)
    _, _, pib, pbar = layer_pressures(mu[:, 0])
    dth = np.array([(hs_teq(pbar[k], lat[0]) * CP / pib[k] - th[k, 0]) * (1 - math.exp(-hs[0] * dt))
                    for k in range(L)])
    assert np.all(dth > 0)
    # the bottom layer (k = 2) below the lowest theta*: the triangle's share
    assert wx[1, 0] == pytest.approx(mu[2, 0] * dth[2] / (2 * (ts_lo - th[2, 0])), rel=1e-10)
    # the middle layer (k = 1) between theta* 300 and 330, mean 305: its upper edge's density
    x = (th[1, 0] - ts_lo) / (ts_hi - ts_lo)
    s = optimize.brentq(lambda s: r_of(s) - x, -50, 50, xtol=1e-14)
    fb = s * math.exp(s) / math.expm1(s) / (ts_hi - ts_lo)
    assert wx[0, 0] == pytest.approx(mu[1, 0] * dth[1] * fb, rel=1e-9)


def test_surface_layer_neutral_log_law():
    L, N = 2, 5
    p = par(L, sfc=1.0, bulk_gust=0.0, z0_land=0.05)
    mu = np.full((L, N), 5000.0)
    ps = GRAV * mu.sum(0)
    Ta = np.full(N, 288.0)
    theta = np.tile(Ta * (P0 / ps) ** KAPPA, (L, 1))
    uabs = np.linspace(2.0, 15.0, N)
    cd, ch = ocol.surface_layer(p, uabs, mu, theta, Ts=Ta, land=np.ones(N))
    depth = 287.0 * Ta / GRAV * np.log(ps / (ps - GRAV * mu[-1]))
    zr = np.maximum(10.0, 0.1 * depth)
    rho = ps / (287.0 * Ta)
    Cd = (0.4 / np.log(zr / 0.05)) ** 2
    Ch = 0.4 ** 2 / (np.log(zr / 0.05) * np.log(zr / 0.005))
    assert np.allclose(cd, rho * Cd * uabs, rtol=1e-12)
    assert np.allclose(ch, rho * Ch * uabs, rtol=1e-12)
    # stable (warm air over cold ground) exchanges less, unstable more
    cds, _ = ocol.surface_layer(p, uabs, mu, theta, Ts=Ta - 5.0, land=np.ones(N))
    cdu, _ = ocol.surface_layer(p, uabs, mu, theta, Ts=Ta + 5.0, land=np.ones(N))
    assert np.all(cds < cd) and np.all(cdu > cd)


def solid_body(g, U=20.0):
    """Normal winds of a solid-body rotation about the z axis, U at the equator, and its speed at the cells."""
    om = np.array([0.0, 0.0, U])
    u = np.sum(np.cross(om, g.xm) * g.ne, axis=1)
    speed = U * np.sqrt(1 - np.clip(g.pos[:, 2], -1, 1) ** 2)
    return u, speed


def test_cell_speeds(grids):
    from insilica.atmosphere import perot_operators
    g = grids(4)
    u, speed = solid_body(g)
    ke = ocol.cell_speed(g, u)
    pr = ocol.cell_speed_perot(g, u, perot_operators(g))
    band = np.abs(g.pos[:, 2]) < 0.9                     # away from the poles (small speeds)
    assert np.max(np.abs(pr[band] / speed[band] - 1)) < 0.01
    assert np.max(np.abs(ke[band] / speed[band] - 1)) < 0.05


def test_gas_src():
    rng = np.random.default_rng(0)
    L, G, N = 3, 2, 10
    src, h, x = rng.uniform(0, 1e-6, (G, N)), rng.uniform(1e3, 5e3, (L, N)), rng.uniform(0, 1e-3, (G * L, N))
    out = ocol.gas_src(src, h, 450.0, x, L)
    want = x.copy()
    for gg in range(G):
        want[gg * L + L - 1] += src[gg] * 450.0 / h[L - 1]
    assert np.array_equal(out, want)
