"""The dry air column's oracle (oracle/column.c og_air_column; reference air.cu:876-1972 with moist = 0) pinned on the
CPU by what the physics fixes, independently of the reference's loops:

- Energy: the column's stored energy (air enthalpy + slab heat) changes by exactly the absorbed sunlight less the
  outgoing longwave, sw_abs - olr, times dt (the sensible heat and the crossing's mass moves only move energy inside
  the column), with the gray scheme, the k-tables' fluxes, the bulk and surface-layer exchange and the ocean's cap.
- An isothermal gray column over a black surface at its temperature: the outgoing longwave is sigma T^4; each layer
  cools by sigma T^4 times the difference of its top's and bottom's transmission to space, exp(-tau(p)) with
  tau(p) = tau0 (f s + (1 - f) s^4); the surface loses sigma T^4 exp(-tau0).
- Sunlight: mode 1 absorbs asw (m / m_ref) I in the air (split by s^4 of pressure) and (I - that)(1 - alb) at the
  surface; mode 0's sun at the equinox stands over (1, 0, 0) at t = 0 and over (0, -cos tilt, sin tilt) at a quarter
  year, its mean over the sphere is S0 / 4, and Kepler's orbit gives the distance factor 1 / (1 -+ e)^2 at perihelion
  and aphelion.
- Sensible heat: h_sens (1 + |U| / u0)(Ts - theta_b (ps / p0)^kappa) out of the slab into the air touching the
  ground; the surface layer's implicit exchange brings that air at most to the surface's temperature in one step.
- Options with exact equivalents: theta_rise over elev is the raised theta*, the ocean's cap equal to c_slab is the
  slab, a firn layer's heat lands in the slab.
The k-tables' edge heating in the crossing (dtop/dbot) is pinned by the energy identity only.
"""
import math

import numpy as np
import pytest

from oracle import column as ocol
from oracle import lib

CP, KAPPA, P0, GRAV, SIGMA = 1004.6, 287.04 / 1004.6, 1.0e5, 9.80616, 5.67e-8


def par3(**kw):
    from insilica.atmosphere import par_array
    return par_array(3, dict(dict(theta_star=305.0, theta_star2=380.0), **kw))


def aqua(N=96, seed=0):
    """The dry aquaplanet's start (three layers with interfaces near 220 and 600 hPa), with seeded noise."""
    rng = np.random.default_rng(seed)
    pos = rng.normal(size=(N, 3))
    pos /= np.linalg.norm(pos, axis=1)[:, None]
    s2 = pos[:, 2] ** 2
    edges = np.array([0.0, 0.22e5, 0.60e5, 1.0e5])[:, None] * rng.uniform(0.97, 1.03, (1, N))
    mu = np.diff(edges, axis=0) / GRAV
    theta = np.stack([np.full(N, 415.0), np.full(N, 342.5), 290.0 - 30.0 * s2]) + rng.normal(0, 1.5, (3, N))
    Ts = 288.0 - 40.0 * (s2 - 1.0 / 3.0) + rng.normal(0, 2.0, N)
    uabs = rng.uniform(0.0, 15.0, N)
    return dict(pos=pos, mu=mu, theta=theta, Ts=Ts, uabs=uabs)


def pib_of(mu):
    """Each layer's mass-mean Exner function cp (p / p0)^kappa over its pressure span."""
    pb = np.cumsum(GRAV * mu, 0)
    pt = np.concatenate([np.zeros((1,) + mu.shape[1:]), pb[:-1]])
    return pt, pb, CP * (pb ** (KAPPA + 1) - pt ** (KAPPA + 1)) / ((KAPPA + 1) * P0 ** KAPPA * (pb - pt))


def run(c, p, mode=0, t=0.3 * 86400.0, dt=900.0, **kw):
    return ocol.air_column(p, mode, np.array([t]), c['pos'], c['uabs'], c['mu'], c['theta'], c['Ts'], dt, **kw)


def stored_matches(o, dt):
    np.testing.assert_allclose(o['stored'], (o['sw_abs'] - o['olr']) * dt, rtol=0, atol=2e-3)


@pytest.mark.parametrize('opts', [dict(), dict(sw_tau=0.5), dict(sw_tau=0.0), dict(bulk_c=1.2e-3, bulk_z=50.0),
                                  dict(bulk_c=1.2e-3, bulk_w=0.05), dict(bl_bot=0.5, mu_bl=3000.0),
                                  dict(theta_star=330.0, edge_heat=1.0)])
def test_energy_closes(opts):
    c = aqua()
    dt = 900.0
    o = run(c, par3(**opts), dt=dt)
    stored_matches(o, dt)
    assert np.all(o['sw_abs'] >= 0) and np.all(o['olr'] > 100)
    assert np.any(np.abs(o['wx']) > 0)                     # the crossing moved air
    assert np.all(o['rain'] == 0) and np.all(o['evap'] == 0) and np.all(o['conv'] == 0)
    assert np.all(o['rainp'] == 0) and np.all(o['flash'] == 0)
    np.testing.assert_array_equal(o['tsurf'], c['Ts'])


def test_energy_closes_mean_sun_ocean_cap_surface_layer():
    c = aqua(seed=4)
    N, dt = c['Ts'].size, 600.0
    rng = np.random.default_rng(5)
    insol = rng.uniform(0, 450, N)
    cap = rng.uniform(1e7, 4e8, N)
    ch = rng.uniform(1e-3, 3e-2, N)
    o = run(c, par3(sfc=1.0), mode=1, dt=dt, insol=insol, cap=cap, ch=ch, ocean=True)
    stored_matches(o, dt)
    # the ocean's forcing is the slab's net gain
    np.testing.assert_allclose((o['Ts'] - c['Ts']) * cap / dt, o['qocean'], rtol=1e-9, atol=1e-9)
    assert np.all(o['qsw'] >= 0)


def test_energy_closes_with_ktables():
    from oracle import radiation as orad
    from insilica.ktables import synthetic_lw, synthetic_sw
    c = aqua(N=48, seed=7)
    N, L, dt = 48, 3, 900.0
    p = par3(edge_k=1.0)
    lwt = orad.LwTables(synthetic_lw())
    swk = synthetic_sw()
    swt = orad.SwTables(swk)
    q = np.full((L, N), 1e-3)
    sub = orad.lw_setup(N, L, p, lwt, c['mu'], c['theta'], q)
    radf = orad.lw_column(N, L, p, lwt, sub, c['Ts'])
    swf = orad.sw_column(N, L, p, swt, sub, np.full(N, 0.06), mode=1)
    for opts in (dict(), dict(edge_heat=1.0)):
        o = run(c, par3(edge_k=1.0, **opts), dt=dt, radf=radf, swf=swf)
        stored_matches(o, dt)
        # the outgoing longwave is the k-tables' top-of-air up flux with the surface's emission carried up
        Es = 0.95 * SIGMA * c['Ts'] ** 4
        np.testing.assert_allclose(o['olr'], radf[0] + radf[2 * (L + 1)] * Es, rtol=1e-14)


def test_isothermal_gray_column():
    N, T0, dt, tau0, f = 5, 262.0, 1800.0, 1.7, 0.35
    mu = np.array([[2100.0], [3900.0], [4100.0]]) * np.linspace(0.9, 1.1, N)[None, :]
    pt, pb, pib = pib_of(mu)
    theta = T0 * CP / pib
    c = dict(pos=np.tile([0.0, 0.0, 1.0], (N, 1)), mu=mu, theta=theta, Ts=np.full(N, T0), uabs=np.full(N, 5.0))
    p = par3(eps_s=1.0, h_sens=0.0, tau0=tau0, tau_lin=f, theta_star=-1.0, theta_star2=-1.0)
    o = run(c, p, mode=1, dt=dt, insol=np.zeros(N))
    B = SIGMA * T0 ** 4
    np.testing.assert_allclose(o['olr'], B, rtol=1e-13)
    tau = lambda pp: tau0 * (f * pp / pb[-1] + (1 - f) * (pp / pb[-1]) ** 4)
    Q = -B * (np.exp(-tau(pt)) - np.exp(-tau(pb)))
    np.testing.assert_allclose((o['theta'] - theta) * mu * pib / dt, Q, rtol=1e-10)
    cw = p[8]
    np.testing.assert_allclose((o['Ts'] - T0) * cw / dt, -B * math.exp(-tau0), rtol=1e-9)
    assert np.all(o['wx'] == 0)


def test_transparent_air_and_black_surface():
    c = aqua(N=16, seed=2)
    o = run(c, par3(tau0=0.0, eps_s=1.0, S0=0.0, h_sens=0.0, theta_star=-1.0))
    np.testing.assert_allclose(o['olr'], SIGMA * c['Ts'] ** 4, rtol=1e-14)
    np.testing.assert_array_equal(o['theta'], c['theta'])


def test_sunlight_split():
    c = aqua(N=24, seed=3)
    N, dt, asw, alb = 24, 900.0, 0.23, 0.11
    I = np.linspace(0.0, 500.0, N)
    o = run(c, par3(asw=asw, alb=alb, tau0=0.0, eps_s=0.0, h_sens=0.0, theta_star=-1.0, sw_tau=2.0), mode=1,
            dt=dt, insol=I)
    m = c['mu'].sum(0)
    air = asw * m / (P0 / GRAV) * I
    sfc = (I - air) * (1 - alb)
    np.testing.assert_allclose(o['sw_abs'], air + sfc, rtol=1e-13, atol=1e-12)
    np.testing.assert_allclose((o['Ts'] - c['Ts']) * (1000.0 * 4186.0 * 100.0) / dt, sfc, rtol=1e-8, atol=1e-8)
    pt, pb, pib = pib_of(c['mu'])
    share = (pb / pb[-1]) ** 4 - (pt / pb[-1]) ** 4
    np.testing.assert_allclose((o['theta'] - c['theta']) * c['mu'] * pib / dt, air * share, rtol=1e-9, atol=1e-9)


def sun_I(g, t, **kw):
    N = len(g.pos)
    z = np.zeros((3, N))
    c = dict(pos=np.asarray(g.pos, float), mu=z + 3000.0, theta=z + 300.0, Ts=np.full(N, 280.0), uabs=np.zeros(N))
    return run(c, par3(asw=0.0, alb=0.0, theta_star=-1.0, **kw), mode=0, t=t)['sw_abs']


def test_sun_geometry():
    from insilica import grid as gridmod
    g = gridmod.get(4)
    pos, A = np.asarray(g.pos, float), np.asarray(g.areaCell, float)
    S0, tilt, year, day = 1361.0, math.radians(23.44), 365 * 86400.0, 86400.0
    I0 = sun_I(g, 0.0)
    np.testing.assert_allclose(I0, S0 * np.maximum(pos[:, 0], 0.0), rtol=1e-12, atol=1e-9)
    assert (A * I0).sum() / A.sum() == pytest.approx(S0 / 4, rel=2e-3)
    Iq = sun_I(g, year / 4)
    s = np.array([0.0, -math.cos(tilt), math.sin(tilt)])
    np.testing.assert_allclose(Iq, S0 * np.maximum(pos @ s, 0.0), rtol=1e-9, atol=1e-6)
    # Kepler's orbit: the distance factor at perihelion and aphelion, the longitude of perihelion
    e, lp = 0.0167, math.radians(282.9)
    p = par3(ecc=e, lam_peri=lp)
    E0 = 2 * math.atan(math.sqrt((1 - e) / (1 + e)) * math.tan(-lp / 2))
    tp = -(E0 - e * math.sin(E0)) * year / (2 * math.pi)
    fd = np.zeros(1)
    lon = lib().og_sun_lon(tp, p.ctypes.data, fd.ctypes.data)
    assert fd[0] == pytest.approx(1 / (1 - e) ** 2, rel=1e-12)
    assert math.remainder(lon - lp, 2 * math.pi) == pytest.approx(0.0, abs=1e-9)
    lib().og_sun_lon(tp + year / 2, p.ctypes.data, fd.ctypes.data)
    assert fd[0] == pytest.approx(1 / (1 + e) ** 2, rel=1e-12)


def test_sensible_heat():
    c = aqua(N=32, seed=6)
    N, dt, hsn, u0 = 32, 900.0, 12.0, 10.0
    p = par3(tau0=0.0, eps_s=0.0, h_sens=hsn, u0=u0, theta_star=-1.0)
    o = run(c, p, mode=1, dt=dt, insol=np.zeros(N))
    ps = GRAV * c['mu'].sum(0)
    Tsa = c['theta'][2] * (ps / P0) ** KAPPA                # the bottom layer is thicker than mu_bl
    H = hsn * (1 + c['uabs'] / u0) * (c['Ts'] - Tsa)
    np.testing.assert_allclose((c['Ts'] - o['Ts']) * p[8] / dt, H, rtol=1e-8, atol=1e-9)
    _, _, pib = pib_of(c['mu'])
    np.testing.assert_allclose((o['theta'][2] - c['theta'][2]) * c['mu'][2] * pib[2] / dt, H, rtol=1e-8, atol=1e-9)
    np.testing.assert_array_equal(o['theta'][:2], c['theta'][:2])
    # the surface layer's implicit exchange: a huge rho C_H U brings the air touching the ground (mu_bl) to the
    # surface's temperature within the step, never past it
    ch = np.full(N, 1e9)
    o2 = run(c, par3(tau0=0.0, eps_s=0.0, sfc=1.0, theta_star=-1.0), mode=1, dt=dt, insol=np.zeros(N), ch=ch)
    mb = P0 / GRAV * 0.1
    heat = (o2['theta'][2] - c['theta'][2]) * c['mu'][2] * pib[2]
    np.testing.assert_allclose(heat / (CP * mb * (c['Ts'] - Tsa)), 1.0, rtol=1e-6)


def test_bulk_stability():
    """Louis's stability: air colder than the surface takes more heat per kelvin than air warmer than it."""
    c = aqua(N=2, seed=1)
    c['uabs'][:] = 3.0
    ps = GRAV * c['mu'].sum(0)
    Tsa = c['theta'][2] * (ps / P0) ** KAPPA
    p = par3(tau0=0.0, eps_s=0.0, bulk_c=1.2e-3, bulk_z=50.0, theta_star=-1.0)
    rate = []
    for d in (4.0, -4.0):
        c['Ts'] = Tsa + d
        o = run(c, p, mode=1, insol=np.zeros(2))
        rate.append((c['Ts'] - o['Ts']) / d)
    assert np.all(rate[0] > 1.2 * rate[1])


def test_bottom_air_share():
    """bl_bot: a bottom layer thinner than bl_bot mu_bl shares the surface's heat with the air above it."""
    c = aqua(N=8, seed=9)
    c['mu'][1] += c['mu'][2] - 600.0
    c['mu'][2] = 600.0
    p = par3(tau0=0.0, eps_s=0.0, bl_bot=0.5, theta_star=-1.0)
    o = run(c, p, mode=1, insol=np.zeros(8))
    _, _, pib = pib_of(c['mu'])
    heat = (o['theta'] - c['theta']) * c['mu'] * pib
    mb = max(600.0, 0.5 * P0 / GRAV * 0.1)
    np.testing.assert_allclose(heat[1] / heat[2], (mb - 600.0) / 600.0, rtol=1e-9)


def test_exact_equivalents():
    c = aqua(N=40, seed=11)
    N = 40
    elev = np.where(np.arange(N) % 2 == 0, 300.0, -50.0)
    a = run(c, par3(theta_rise=0.01), elev=elev)
    for i in range(N):
        ci = {k: (v[..., i:i + 1] if k != 'pos' else v[i:i + 1]) for k, v in c.items()}
        ts = 305.0 + (0.01 * elev[i] if elev[i] > 0 else 0.0)
        b = run(ci, par3(theta_star=ts))
        np.testing.assert_array_equal(a['theta'][:, i], b['theta'][:, 0])
        np.testing.assert_array_equal(a['wx'][:, i], b['wx'][:, 0])
    s = run(c, par3())
    cs = run(c, par3(), cap=np.full(N, 1000.0 * 4186.0 * 100.0))
    for k in ('theta', 'Ts', 'wx', 'stored'):
        np.testing.assert_array_equal(s[k], cs[k])
    # a firn layer's heat lands in the slab before the surface's fluxes: the same step from the slab warmed by it
    Tfn = np.full(N, 265.0)
    f = run(c, par3(firn=10.0), firn=Tfn)
    Cf = 2090.0 * 450.0 * 10.0
    cw = dict(c, Ts=c['Ts'] + Cf * (265.0 - 273.15) / (1000.0 * 4186.0 * 100.0))
    w = run(cw, par3())
    for k in ('theta', 'Ts', 'wx', 'olr'):
        np.testing.assert_array_equal(f[k], w[k])
    np.testing.assert_array_equal(f['firn'], 273.15)
    stored_matches(f, 900.0)


def test_mixed_build_close():
    c = aqua(N=64, seed=12)
    a, b = run(c, par3()), run(c, par3(), mix=True)
    for k in ('theta', 'Ts', 'olr'):
        np.testing.assert_allclose(b[k], a[k], rtol=2e-6)
    assert np.any(b['olr'] != a['olr'])


# ---- the moist column (moist = 1; no clouds, no land) ----

def tetens(T, p, eps=0.622):
    """Saturation specific humidity over water (Tetens, v1's formula)."""
    return eps * 611.0 * np.exp(17.27 * (T - 273.15) / (T - 35.85)) / p


def moist_aqua(N=96, seed=0, rh=(0.1, 0.5, 1.05), warm=0.0, ice_share=0.3):
    c = aqua(N, seed)
    c['theta'][2] += warm
    pt, pb, _ = pib_of(c['mu'])
    pm = 0.5 * (pt + pb)
    c['q'] = np.array(rh)[:, None] * tetens(c['theta'] * (pm / P0) ** KAPPA, pm)
    rng = np.random.default_rng(seed + 100)
    c['ice'] = np.where(rng.uniform(size=N) < ice_share, rng.uniform(0, 200, N), 0.0)
    return c


def mrun(c, p, **kw):
    return run(c, p, q=c['q'], ice=c['ice'], **kw)


def new_masses(c, wx):
    mu = c['mu'].copy()
    mu[:-1] += wx
    mu[1:] -= wx
    return mu


@pytest.mark.parametrize('opts', [dict(), dict(lw_w0=5.0), dict(sat_layer=0.0), dict(q_sea=0.98, sfc=1.0),
                                  dict(bulk_c=1.2e-3, bulk_z=50.0)])
@pytest.mark.parametrize('mix', [False, True])
def test_moist_energy_and_water_close(opts, mix):
    """Energy: latent heat moves between the air's vapor and its enthalpy, the slab's heat into ice, nothing leaves
    but sunlight in and longwave out. Water: the column gains evaporation and loses rain."""
    c = moist_aqua(rh=(0.1, 0.8, 1.6), warm=18.0)
    dt = 900.0
    kw = dict(ch=np.full(96, 0.02)) if opts.get('sfc') else {}
    o = mrun(c, par3(moist=1.0, **opts), dt=dt, mix=mix, **kw)
    stored_matches(o, dt)
    W0 = (c['mu'] * c['q']).sum(0)
    W1 = (new_masses(c, o['wx']) * o['q']).sum(0)
    np.testing.assert_allclose(W1 - W0, (o['evap'] - o['rain']) * dt, rtol=0, atol=1e-12 * W0.max())
    assert o['rain'].max() > 0 and o['conv'].max() > 0 and o['evap'].max() > 0
    np.testing.assert_allclose(o['rainp'][:3].sum(0) - o['rainp'][2], o['rain'] - o['rainp'][2], rtol=1e-12,
                               atol=1e-18)
    assert np.all(o['q'] >= 0) and np.all(o['ice'] >= 0)


def test_condensation_saturates_the_layer():
    """A supersaturated layer (no sun, no longwave, no crossing or plumes) condenses in one Newton step to within a
    few per mille of its capacity, the layer's saturation averaged over its pressure span (adaptive quadrature here),
    and its latent heat warms it exactly as its vapor fell: L_v dq = pib dtheta."""
    from scipy import integrate
    c = moist_aqua(N=8, seed=3, ice_share=0.0)
    pt, pb, pib = pib_of(c['mu'])

    def capacity(th, k, i):
        return integrate.quad(lambda pp: tetens(th * (pp / P0) ** KAPPA, pp), pt[k, i], pb[k, i],
                              epsrel=1e-12)[0] / (pb[k, i] - pt[k, i])
    for k in (1, 2):                                     # 12% beyond the layer's capacity
        c['q'][k] = [1.12 * capacity(c['theta'][k, i], k, i) for i in range(8)]
    p = par3(moist=1.0, tau0=0.0, eps_s=0.0, h_sens=0.0, k_evap=0.0, theta_star=-1.0, theta_star2=-1.0)
    o = mrun(c, p, mode=1, insol=np.zeros(8))
    for k in (1, 2):
        dq = c['q'][k] - o['q'][k]
        assert np.all(dq > 0)
        np.testing.assert_allclose(2.5e6 * dq, pib[k] * (o['theta'][k] - c['theta'][k]), rtol=1e-12)
        for i in range(8):
            cap = capacity(o['theta'][k, i], k, i)
            assert o['q'][k, i] == pytest.approx(cap, rel=5e-3)
            # the Newton step removed nearly all of the excess: what is left is second order in it
            assert abs(o['q'][k, i] - cap) < 0.05 * (c['q'][k, i] - capacity(c['theta'][k, i], k, i))
    assert o['rain'] == pytest.approx((c['mu'][1:] * (c['q'][1:] - o['q'][1:])).sum(0) / 900.0, rel=1e-12)


def test_evaporation():
    """Over open water the air touching the sea takes k_evap (1 + |U| / u0) (q_s(Ts, ps) - q_b) kg/m^2/s, the slab
    pays Lv for it; under ice cover f only the open share 1 - f evaporates."""
    c = moist_aqua(N=16, seed=5, rh=(0.1, 0.3, 0.5), ice_share=0.0)
    c['ice'][:8] = np.linspace(0.0, 60.0, 8) * 917.0 / 1000.0          # 0..6 cm of ice: cover 0..1
    p = par3(moist=1.0, tau0=0.0, eps_s=0.0, h_sens=0.0, theta_star=-1.0, theta_star2=-1.0)
    dt = 600.0
    o = mrun(c, p, mode=1, insol=np.zeros(16), dt=dt)
    ps = GRAV * c['mu'].sum(0)
    fice = np.minimum(1.0, c['ice'] / 917.0 / 0.05)
    E = 4e-3 * (1 + c['uabs'] / 10.0) * (tetens(c['Ts'], ps) - c['q'][2]) * (1 - fice)
    thick = c['ice'] / 917.0 > 0.05
    np.testing.assert_allclose(o['evap'][~thick], E[~thick], rtol=1e-12)
    np.testing.assert_allclose((o['q'][2] - c['q'][2]) * c['mu'][2], o['evap'] * dt, rtol=1e-10)
    open_ = (c['ice'] == 0)
    np.testing.assert_allclose((o['Ts'] - c['Ts'])[open_ & (o['Ts'] > 273.15)] * p[8] / dt,
                               -2.5e6 * o['evap'][open_ & (o['Ts'] > 273.15)], rtol=1e-8)


def test_sea_ice():
    """A slab below freezing turns its heat below Tf into ice (Cw (Tf - T) / Lf kg/m^2) and stays at Tf; under thick
    ice the skin's temperature balances conduction k_ice (T_slab - x) / d against the surface's fluxes at x (solved here
    by Brent's method; transparent air, so no longwave comes down), capped at Tf; the conducted heat freezes ice under
    it; the ice's albedo covers the sea by the ice's cover."""
    from scipy import optimize
    N = 6
    c = moist_aqua(N=N, seed=8, ice_share=0.0)
    c['Ts'] = np.array([268.0, 270.0, 272.0, 275.0, 271.5, 273.15])
    c['ice'] = np.array([0.0, 0.0, 0.0, 0.0, 500.0, 300.0])
    c['theta'][2] = 250.0
    p = par3(moist=1.0, tau0=0.0, theta_star=-1.0, theta_star2=-1.0)
    dt, Cw, Lf, Tf = 900.0, p[8], 3.34e5, 273.15
    o = mrun(c, p, mode=1, insol=np.full(N, 300.0), dt=dt)
    stored_matches(o, dt)
    for i in range(3):                                   # open water below freezing: its heat below Tf is ice
        assert o['Ts'][i] == Tf
        assert o['ice'][i] == pytest.approx((Tf - o['tsurf'][i]) * Cw / Lf
                                            + (o['Ts'][i] - Tf) * 0, rel=0.02)
    assert o['Ts'][3] > Tf and o['ice'][3] == 0
    for i in (4, 5):
        d = c['ice'][i] / 917.0
        kc, tb = 2.2 / d, c['Ts'][i]
        ps = GRAV * c['mu'][:, i].sum()
        Tsa = c['theta'][2, i] * (ps / P0) ** KAPPA
        alb = 0.7 + (0.5 - 0.7) * min(1.0, max(0.0, (Tsa - Tf + 10.0) / 10.0))
        toAir = 0.2 * (c['mu'][:, i].sum() / (P0 / GRAV)) * 300.0
        swS = (300.0 - toAir) * (1 - alb)
        assert o['sw_abs'][i] == pytest.approx(toAir + swS, rel=1e-12)
        hs = 10.0 * (1 + c['uabs'][i] / 10.0)
        f = lambda xx: kc * (tb - xx) - 0.95 * SIGMA * xx ** 4 + swS - hs * (xx - Tsa)
        xb = min(optimize.brentq(f, 150.0, 400.0, xtol=1e-13), Tf)
        assert o['tsurf'][i] == pytest.approx(xb, abs=1e-8)
        assert o['Ts'][i] == Tf
        # the slab gave kc (tb - x) dt up through the ice and came back to Tf by freezing; the top melted what the
        # balance left over at Tf
        melt = f(xb) * dt / Lf
        Tn = tb - kc * (tb - xb) * dt / Cw
        assert o['ice'][i] == pytest.approx(c['ice'][i] - melt + (Tf - Tn) * Cw / Lf, rel=1e-9)


def test_plume_relaxes_over_tau_conv():
    """A conditionally unstable column (a warm, moist bottom layer whose lifted parcel is warmer than theta* at its top)
    sends up mc = m_max (1 - exp(-dt / tau_conv)): halving dt at a long tau_conv halves it, its rain is the parcel's
    condensate, it detrains where the parcel stops being buoyant."""
    c = moist_aqua(N=32, seed=9, rh=(0.1, 0.4, 1.0), warm=22.0, ice_share=0.0)
    p = par3(moist=1.0, tau0=0.0, eps_s=0.0, h_sens=0.0, k_evap=0.0, tau_conv=1e7)
    a = mrun(c, p, mode=1, insol=np.zeros(32), dt=900.0)
    b = mrun(c, p, mode=1, insol=np.zeros(32), dt=450.0)
    act = a['conv'] > 0
    assert act.sum() >= 8
    ratio = -math.expm1(-900.0 / 1e7) / -math.expm1(-450.0 / 1e7)
    np.testing.assert_allclose(a['conv'][act], ratio * b['conv'][act], rtol=1e-6)
    # the plume's mass crossed the interface above the bottom layer
    assert np.all(a['wx'][1][act] >= a['conv'][act] * (1 - 1e-12))
    assert np.all(a['rainp'][2][act] > 0)


# ---- clouds (cloud > 0, with the cloud water cw) ----

def cloudy(seed=21, N=64, cw=2e-5, **kw):
    c = moist_aqua(N=N, seed=seed, **kw)
    c['cw'] = np.full_like(c['q'], cw)
    return c


def crun(c, p, **kw):
    return run(c, p, q=c['q'], ice=c['ice'], cw=c['cw'], cloudo=True, **kw)


@pytest.mark.parametrize('opts', [dict(), dict(ice_air=1.0), dict(revap=1.0), dict(conv_revap=0.05),
                                  dict(cw_qs=1.0), dict(sat_layer=0.0, lw_w0=5.0)])
@pytest.mark.parametrize('mix', [False, True])
def test_clouds_energy_and_water_close(opts, mix):
    """With clouds the energy still closes (cloud water holds no energy of its own: its latent heat went into theta
    when it formed); vapor plus cloud water change by evaporation less rain (the plumes held off by a long tau_conv:
    a plume moves its base layer's air without its cloud water, as the reference's does)."""
    c = cloudy(rh=(0.3, 1.0, 1.3), warm=18.0)
    dt = 900.0
    o = crun(c, par3(moist=1.0, cloud=1.0, **opts), dt=dt, mix=mix)
    stored_matches(o, dt)
    assert o['rain'].max() > 0 and np.all(o['cw'] >= 0) and np.all(o['q'] >= 0)
    cover = o['cloudo'][:3]
    assert np.all((cover >= 0) & (cover <= 1)) and np.all(o['cloudo'][3] >= cover.max(0) - 1e-15)
    q = crun(c, par3(moist=1.0, cloud=1.0, tau_conv=1e12, **opts), dt=dt, mix=mix)
    W0 = (c['mu'] * (c['q'] + c['cw'])).sum(0)
    W1 = (new_masses(c, q['wx']) * (q['q'] + q['cw'])).sum(0)
    np.testing.assert_allclose(W1 - W0, (q['evap'] - q['rain']) * dt, rtol=0, atol=1e-11 * W0.max())


def test_cloud_statistics():
    """A layer's cloud water goes to the mean condensate of its total water spread evenly over +-(1 - rh_crit) q_s
    (the mean of max(qt + x - q_s, 0), integrated here), q_s the layer's capacity, with one Newton step for the latent
    heat's feedback: what is left is within a tenth of the initial departure. Rain, erosion, sunlight, longwave, the
    surface and the plumes are off."""
    from scipy import integrate
    c = cloudy(seed=5, N=6, cw=0.0, ice_share=0.0)
    pt, pb, pib = pib_of(c['mu'])
    k = 2
    capf = lambda th, i: integrate.quad(lambda pp: tetens(th * (pp / P0) ** KAPPA, pp), pt[k, i], pb[k, i],
                                        epsrel=1e-12)[0] / (pb[k, i] - pt[k, i])
    for i in range(6):
        c['q'][k, i] = capf(c['theta'][k, i], i) * (0.9 + 0.04 * i)      # around saturation: a partial cover
    p = par3(moist=1.0, cloud=1.0, c_auto=0.0, v_ice=0.0, tau_erode=1e30, tau0=0.0, eps_s=0.0, h_sens=0.0,
             k_evap=0.0, theta_star=-1.0, theta_star2=-1.0, rh_crit=0.8, liq_cold=1e-9)
    o = crun(c, p, mode=1, insol=np.zeros(6))
    for i in range(6):
        qs = capf(o['theta'][k, i], i)
        b = 0.2 * qs
        qt = o['q'][k, i] + o['cw'][k, i]
        mean = integrate.quad(lambda x: max(qt + x - qs, 0.0), -b, b, points=[qs - qt])[0] / (2 * b)
        qs0 = capf(c['theta'][k, i], i)
        mean0 = integrate.quad(lambda x: max(c['q'][k, i] + x - qs0, 0.0), -0.2 * qs0, 0.2 * qs0,
                               points=[qs0 - c['q'][k, i]])[0] / (0.4 * qs0)
        assert mean0 > 0 and abs(o['cw'][k, i] - mean) < 0.1 * mean0       # (without the feedback: ~2x off)
    # the condensate's latent heat: L_v dl = pib dtheta in the layer
    np.testing.assert_allclose(2.5e6 * o['cw'][k], pib[k] * (o['theta'][k] - c['theta'][k]), rtol=1e-10)


def test_cloud_radiation():
    """A thick cloud filling the top layer over transparent air is a black body at that layer's temperature for the
    outgoing longwave; the sunlight reaching a reflecting surface under a conservatively scattering cloud layer of
    transmission T is the series of reflections between them: (1 - alb) T / (1 - alb (1 - T)) of what enters."""
    c = cloudy(seed=8, N=8, cw=0.0, ice_share=0.0)
    c['q'][:] = 0.0
    c['cw'][0] = 5e-3                                    # ~11 kg/m^2 of cloud water in the top layer
    p = par3(moist=1.0, cloud=1.0, tau0=0.0, eps_s=1.0, h_sens=0.0, k_evap=0.0, theta_star=-1.0, theta_star2=-1.0,
             c_auto=0.0, v_ice=0.0, tau_erode=1e30, asw=0.0, alb=0.3)
    I = np.full(8, 400.0)
    o = crun(c, p, mode=1, insol=I)
    _, _, pib = pib_of(c['mu'])
    T0 = pib[0] * c['theta'][0] / CP
    np.testing.assert_allclose(o['olr'], SIGMA * T0 ** 4, rtol=1e-9)
    assert np.all(o['cloudo'][0] == 1.0)
    # the cloud's transmission T from a black surface (alb 0); over alb 0.3 the reflections between the surface and
    # the conservatively scattering cloud (reflectance 1 - T) sum as a series
    T = crun(c, par3(moist=1.0, cloud=1.0, tau0=0.0, eps_s=1.0, h_sens=0.0, k_evap=0.0, theta_star=-1.0,
                     theta_star2=-1.0, c_auto=0.0, v_ice=0.0, tau_erode=1e30, asw=0.0, alb=0.0),
             mode=1, insol=I)['sw_abs'] / I
    assert np.all((T > 0.005) & (T < 0.5))
    series = sum((1 - 0.3) * T * (0.3 * (1 - T)) ** n for n in range(400))
    np.testing.assert_allclose(o['sw_abs'], I * series, rtol=1e-12)


def test_autoconversion_limit():
    """Warm cloud far above the autoconversion threshold rains at Sundqvist's rate c_auto: a share
    1 - exp(-c_auto dt) of its water per step (evaporation of the rain held off: k_revap 0). The bottom layer is
    warmed above freezing: with liq_cold tiny its cloud is then all liquid (below freezing it would all be ice and
    fall at v_ice / depth instead)."""
    c = cloudy(seed=9, N=4, cw=0.0, ice_share=0.0, warm=30.0)
    c['q'][:] = 0.0
    c['cw'][2] = 5e-2                                    # 50 g/kg: (l / q_auto)^2 ~ 2.8e4
    p = par3(moist=1.0, cloud=1.0, tau0=0.0, eps_s=0.0, h_sens=0.0, k_evap=0.0, theta_star=-1.0, theta_star2=-1.0,
             tau_erode=1e30, k_revap=0.0, liq_cold=1e-9)
    dt = 600.0
    o = crun(c, p, mode=1, insol=np.zeros(4), dt=dt)
    cw1 = o['cw'][2] + o['rain'] * dt / c['mu'][2]      # the cloud water before it rained
    np.testing.assert_allclose(o['rain'] * dt / c['mu'][2], cw1 * -math.expm1(-1e-3 * dt), rtol=1e-9)


def test_ice_fall_rate():
    """All-ice cloud (bottom layer below freezing, liq_cold tiny) falls out at v_ice over the layer's depth: a share
    1 - exp(-v_ice dt / depth) per step, the depth mu R_d T / p at its mid pressure (hydrostatic, ideal gas; the air's R_d 287.0)."""
    c = cloudy(seed=9, N=4, cw=0.0, ice_share=0.0)
    c['q'][:] = 0.0
    c['cw'][2] = 5e-3
    p = par3(moist=1.0, cloud=1.0, tau0=0.0, eps_s=0.0, h_sens=0.0, k_evap=0.0, theta_star=-1.0, theta_star2=-1.0,
             tau_erode=1e30, k_revap=0.0, liq_cold=1e-9, v_ice=0.5, k_lw_liq=0.0, k_lw_ice=0.0)
    dt = 600.0
    pt, pb, pib = pib_of(c['mu'])
    T = pib[2] * c['theta'][2] / CP
    assert np.all(T < 272.0)
    depth = c['mu'][2] * 287.0 * T / (0.5 * (pt[2] + pb[2]))
    o = crun(c, p, mode=1, insol=np.zeros(4), dt=dt)
    cw1 = o['cw'][2] + o['rain'] * dt / c['mu'][2]
    np.testing.assert_allclose(o['rain'] * dt / c['mu'][2], cw1 * -np.expm1(-0.5 * dt / depth), rtol=1e-9)


# ---- land (air.cu air_column's land branches): the soil bucket, ponds, snow, the ground's skin, firn ----

def land_aqua(N=96, seed=3, warm=12.0, rh=(0.1, 0.8, 1.3), ice_share=0.3):
    """moist_aqua's columns with two thirds of them land: a soil bucket of 50-300 kg/m^2 a fifth to fully wet, ponds,
    snow on part of the land's ice, the ground's heat capacity c_land."""
    c = moist_aqua(N=N, seed=seed, rh=rh, warm=warm, ice_share=ice_share)
    rng = np.random.default_rng(seed + 200)
    c['land'] = (rng.uniform(size=N) < 2 / 3).astype(float)
    c['soilcap'] = rng.uniform(50.0, 300.0, N)
    c['soil'] = c['soilcap'] * rng.uniform(0.2, 1.0, N)
    c['pond'] = np.where(rng.uniform(size=N) < 0.4, rng.uniform(0.0, 60.0, N), 0.0)
    c['snow'] = np.minimum(c['ice'], rng.uniform(0.0, 80.0, N))
    c['cap'] = np.where(c['land'] > 0.5, 2.0e6, 4.186e8)
    return c


def lrun(c, p, **kw):
    kw.setdefault('cap', c['cap'])
    return run(c, p, q=c['q'], ice=c['ice'], land=c['land'], soilcap=c['soilcap'], soil=c['soil'], pond=c['pond'],
               snow=c['snow'], ocean=True, **kw)


def cover(N, seed):
    """A land cover's eight rows: z0, the dry and wet soil's albedo, the vegetated share, the vegetation's albedo,
    the share of ice that shows, the stomata's resistance and their sunlight scale."""
    rng = np.random.default_rng(seed)
    return np.stack([rng.uniform(0.01, 1.0, N), rng.uniform(0.25, 0.35, N), rng.uniform(0.1, 0.2, N),
                     rng.uniform(0.0, 1.0, N), rng.uniform(0.1, 0.2, N), rng.uniform(0.3, 1.0, N),
                     rng.uniform(50.0, 150.0, N), rng.uniform(100.0, 250.0, N)])


LAND_OPTS = [dict(), dict(land_rs=100.0, rs_max=5000.0), dict(c_deep=1.0), dict(c_wet=3.0e6, c_dry=1.3e6),
             dict(land_exch=2.0, revap=1.0, cloud=1.0), dict(firn=2.0, ice_air=1.0, cloud=1.0),
             dict(land_rs=100.0, bare_fx=2.0, lc=1)]


@pytest.mark.parametrize('opts', LAND_OPTS)
@pytest.mark.parametrize('mix', [False, True])
def test_land_energy_and_water_close(opts, mix):
    """Over land the water stays in the cell: the air's vapor (and cloud water) plus the soil's, the ponds' and the
    ice's change by nothing; over the sea by evaporation less rain. The energy closes as over the sea (the ground's
    heat C Ts, the firn's, the ice's -L_f m). The plumes are held off where clouds are on (a plume moves its base
    layer's air without its cloud water, as the reference's does)."""
    opts = dict(opts)
    c = land_aqua()
    N = c['land'].size
    kw = {}
    if opts.pop('lc', 0):
        kw['lc'] = cover(N, 5)
    if 'firn' in opts:
        kw['firn'] = np.random.default_rng(9).uniform(240.0, 273.15, N)
        c['ice'] = np.where(c['land'] > 0.5, np.maximum(c['ice'], 1500.0), c['ice'])
    cloudy = opts.get('cloud', 0) > 0
    if cloudy:
        c['cw'] = np.full_like(c['q'], 1e-4)
        kw['cw'] = c['cw']
    dt = 900.0
    p = par3(moist=1.0, **dict(opts, tau_conv=1e12) if cloudy else opts)
    o = lrun(c, p, dt=dt, mix=mix, **kw)
    stored_matches(o, dt)
    land = c['land'] > 0.5
    mu1 = new_masses(c, o['wx'])
    W0 = (c['mu'] * (c['q'] + (c['cw'] if cloudy else 0))).sum(0)
    W1 = (mu1 * (o['q'] + (o['cw'] if cloudy else 0))).sum(0)
    G0, G1 = c['soil'] + c['pond'] + c['ice'], o['soil'] + o['pond'] + o['ice']
    np.testing.assert_allclose((W1 - W0 + G1 - G0)[land], 0.0, atol=1e-11 * (W0 + G0).max())
    np.testing.assert_allclose((W1 - W0)[~land], ((o['evap'] - o['rain']) * dt)[~land], atol=1e-11 * W0.max())
    assert 'firn' in opts or o['evap'][land].max() > 0          # (with firn all the land is under ice)
    assert np.all(o['soil'] >= 0) and np.all(o['pond'] >= -1e-12)
    assert np.all(o['snow'][land] <= o['ice'][land] + 1e-12) and np.all(o['soil'] <= c['soilcap'] * (1 + 1e-15))
    assert np.all(o['soil'][~land] == c['soil'][~land]) and np.all(o['pond'][~land] == c['pond'][~land])
    if 'c_wet' in opts:                                   # the ground's capacity follows its wetness
        r = np.minimum(1.0, c['soil'] / c['soilcap'])
        np.testing.assert_allclose(o['cap'][land], (1.3e6 + 1.7e6 * r)[land], rtol=1e-15)


def test_land_evaporation_closed_form():
    """Bare wet ground below the bucket's soil_beta share, no ponds, no ice, no surface resistance: evaporation is the
    open water's kex (q_s(Ts, p_s) - q_a) times land_exch times wet / (soil_beta capacity); the surface resistance
    land_rs multiplies it by r_a / (r_a + r_s), r_a = rho / kex (the bottom layer holds the whole surface air here)."""
    c = land_aqua(ice_share=0.0)
    c['pond'][:] = 0.0
    c['soil'] = 0.5 * c['soilcap']
    N = c['land'].size
    for rs in (0.0, 100.0):
        p = par3(moist=1.0, land_exch=2.0, land_rs=rs, soil_beta=0.75, theta_star=-1.0, theta_star2=-1.0)
        o = lrun(c, p, mode=1, insol=np.zeros(N), dt=600.0)
        pt, pb, _ = pib_of(c['mu'])
        ps = pb[2]
        uabs = c['uabs']
        kex = np.where(c['land'] > 0.5, 2.0, 1.0) * 4.0e-3 * (1 + uabs / 10.0)
        qa = c['q'][2]
        Tsa = c['theta'][2] * (ps / P0) ** KAPPA
        E = kex * np.maximum(0.0, tetens(c['Ts'], ps) - qa)
        land = c['land'] > 0.5
        f = 0.5 / 0.75
        if rs > 0:
            ra = ps / (287.0 * Tsa) / kex
            f = f * ra / (ra + rs)
        np.testing.assert_allclose(o['evap'][land], (E * f)[land], rtol=1e-12)
        np.testing.assert_allclose(o['evap'][~land], E[~land], rtol=1e-12)


def test_land_infiltration_and_spring():
    """A pond on thawed ground with nothing evaporating soaks in at k_inf dt (1 - wet / capacity), at most the pond;
    a bucket filled past its capacity gives the rest back to the pond."""
    c = land_aqua(ice_share=0.0)
    c['q'][:] = 1e-9                                        # (no rain)
    c['Ts'] = np.maximum(c['Ts'], 280.0)
    N = c['land'].size
    c['pond'] = np.random.default_rng(4).uniform(0.0, 40.0, N)
    c['soil'] = np.random.default_rng(5).uniform(0.5, 1.0, N) * c['soilcap']
    dt = 1800.0
    p = par3(moist=1.0, k_evap=0.0, theta_star=-1.0, theta_star2=-1.0, k_inf=5.0 / 3600)
    o = lrun(c, p, mode=1, insol=np.zeros(N), dt=dt)
    land = c['land'] > 0.5
    room = 5.0 / 3600 * dt * (1 - c['soil'] / c['soilcap'])
    m = np.minimum(c['pond'], room)
    np.testing.assert_allclose(o['soil'][land], (c['soil'] + m)[land], rtol=1e-15)
    np.testing.assert_allclose(o['pond'][land], (c['pond'] - m)[land], rtol=1e-12, atol=1e-12)
    c['soil'] = c['soilcap'] * 1.1                          # over the capacity: a spring
    o = lrun(c, p, mode=1, insol=np.zeros(N), dt=dt)
    np.testing.assert_allclose(o['soil'][land], c['soilcap'][land], rtol=1e-15)
    np.testing.assert_allclose(o['pond'][land], (c['pond'] + 0.1 * c['soilcap'])[land], rtol=1e-12)


def test_ground_skin_implicit():
    """With c_deep the ground's skin is stepped implicitly: C (x - T_s) = dt F(x). Its energy then is the ground's
    change, so the ground ends at the skin's temperature (Newton converged: to 1e-9 K) wherever neither ponds freeze
    nor snow lands."""
    c = land_aqua(ice_share=0.0)
    c['Ts'] = np.maximum(c['Ts'], 285.0)
    c['q'] = 0.3 * c['q']                                   # (no rain)
    N = c['land'].size
    c['cap'] = np.where(c['land'] > 0.5, 3.0e5, 4.186e8)
    o = lrun(c, par3(moist=1.0, c_deep=1.0, theta_star=-1.0, theta_star2=-1.0), dt=900.0)
    land = c['land'] > 0.5
    assert o['rain'].max() == 0
    np.testing.assert_allclose(o['Ts'][land], o['tsurf'][land], atol=1e-9)
    assert np.abs(o['Ts'] - c['Ts'])[land].max() > 0.1      # (and it moved)


# ---- the mixed boundary layer (bl_mix: air.cu air_column's mixed bottom layer, its entrainment, night layer,
# dry convection and condensation level) ----

BLM_OPTS = [dict(), dict(bl_reach=1.0), dict(bl_lcl=1.0), dict(bl_lcl=1.0, ent_lcl=1.0, bl_reach=1.0),
            dict(moist=0.0), dict(cloud=1.0, bl_lcl=1.0), dict(land=1, bl_reach=1.0, mu_night=2000.0)]


@pytest.mark.parametrize('opts', BLM_OPTS)
@pytest.mark.parametrize('mix', [False, True])
def test_mixed_layer_closes(opts, mix):
    """With the mixed bottom layer the energy still closes, the water over the sea changes by evaporation less rain
    (over land by nothing with the ground's), the column's mass is only moved between layers and every layer keeps
    its floor."""
    opts = dict(opts)
    land = opts.pop('land', 0)
    c = land_aqua() if land else moist_aqua(rh=(0.1, 0.8, 1.3), warm=12.0)
    N = c['Ts'].size
    moist = opts.pop('moist', 1.0)
    cloudy = opts.get('cloud', 0) > 0
    kw = {}
    if cloudy:
        c['cw'] = np.full_like(c['q'], 1e-4)
        kw['cw'] = c['cw']
    p = par3(moist=moist, bl_mix=1.0, **dict(opts, tau_conv=1e12) if cloudy else opts)
    dt = 900.0
    if land:
        o = lrun(c, p, dt=dt, mix=mix, **kw)
    elif moist:
        o = run(c, p, q=c['q'], ice=c['ice'], dt=dt, mix=mix, **kw)
    else:
        o = run(c, p, dt=dt, mix=mix)
    stored_matches(o, dt)
    mu1 = new_masses(c, o['wx'])
    np.testing.assert_allclose(mu1.sum(0), c['mu'].sum(0), rtol=1e-14)
    assert np.all(mu1 >= 10.0 * (1 - 1e-9))
    assert np.abs(o['wx'][-1]).max() > 0
    if moist:
        W0 = (c['mu'] * (c['q'] + (c['cw'] if cloudy else 0))).sum(0)
        W1 = (mu1 * (o['q'] + (o['cw'] if cloudy else 0))).sum(0)
        sea = c['land'] < 0.5 if land else np.ones(N, bool)
        np.testing.assert_allclose((W1 - W0)[sea], ((o['evap'] - o['rain']) * dt)[sea], atol=1e-11 * W0.max())


def test_mixed_layer_night():
    """A mixed layer over a surface colder than its air, with no radiation and nothing else heating (so nothing
    crosses an interface), gives its top back to the layer above down to the night layer mu_night when the surface's
    buoyancy flux asks for more: the mass crossing the lowest interface is exactly mu_bot - mu_night."""
    c = aqua()
    N = c['Ts'].size
    c['theta'][1] = np.minimum(c['theta'][1], 360.0)          # (the middle layer below its top edge's 380 K)
    c['Ts'] = c['theta'][2] * (1.0e5 / P0) ** KAPPA - 30.0     # 30 K colder than the air above it
    mu_night = float(np.median(c['mu'][2])) - 150.0
    p = par3(bl_mix=1.0, mu_night=mu_night, tau0=0.0, eps_s=0.0, h_sens=200.0)
    o = run(c, p, mode=1, insol=np.zeros(N), dt=3600.0)
    capped = c['mu'][2] - mu_night
    big = capped > 0
    assert o['wx'][0].max() == 0 and big.sum() > 10
    wx = o['wx'][1]
    # where the flux's own ask exceeded the room, the layer ends at mu_night exactly
    ended = np.isclose(wx, capped, rtol=1e-14)
    assert ended[big].sum() > 0.3 * big.sum()
    assert np.all(wx[big] <= capped[big] * (1 + 1e-14))


def test_mixed_layer_condensation_level():
    """bl_lcl: the mixed layer ends at its condensation level. Its air lifted dry-adiabatically saturates at the
    pressure p_l where q_s(theta (p_l / p0)^kappa, p_l) = q (Tetens over water); the layer ends holding the mass
    (p_s - p_l) / g (bisection in numpy against the oracle's, to 1e-6 kg/m^2) where it held more."""
    c = moist_aqua(rh=(0.1, 0.5, 0.97), warm=6.0)
    N = c['Ts'].size
    p = par3(moist=1.0, bl_mix=1.0, bl_lcl=1.0, tau0=0.0, eps_s=0.0, h_sens=0.0, k_evap=0.0, mu_night=0.0,
             tau_conv=1e30)                                 # (the plumes, after the mixed layer's steps, held off)
    o = run(c, p, mode=1, insol=np.zeros(N), dt=900.0, q=c['q'], ice=np.zeros(N))
    mu1 = new_masses(c, o['wx'])
    ps = GRAV * c['mu'].sum(0)
    th, q = o['theta'][2], o['q'][2]
    sat = lambda pp: tetens(th * (pp / P0) ** KAPPA, pp) <= q
    lo, hi = 0.1 * ps, ps.copy()
    for _ in range(60):
        m = 0.5 * (lo + hi)
        s = sat(m)
        lo, hi = np.where(s, m, lo), np.where(s, hi, m)
    lcl = (ps - 0.5 * (lo + hi)) / GRAV
    left = c['mu'][2] > lcl + 1.0
    assert left.sum() > 10
    np.testing.assert_allclose(mu1[2][left], lcl[left], atol=1e-6)


# ---- the earth preset's moist options: sat_prof, sat_base, edge_vap, freeze_dry, q_shape, plume_root, root_q ----

EARTH_OPTS = [dict(sat_prof=1.0), dict(sat_base=1.0), dict(edge_vap=1.0), dict(freeze_dry=1.0, edge_vap=1.0),
              dict(q_shape=1.0), dict(plume_root=1.0), dict(plume_root=1.0, root_q=1.0, sfc=1.0),
              dict(sat_prof=1.0, edge_vap=1.0, freeze_dry=1.0, q_shape=1.0, plume_root=1.0, cloud=1.0),
              dict(edge_vap=1.0, bl_mix=1.0, bl_reach=1.0)]


@pytest.mark.parametrize('opts', EARTH_OPTS)
@pytest.mark.parametrize('mix', [False, True])
def test_earth_options_close(opts, mix):
    """Each option only moves vapor between the layers (edge vapor, freeze-drying), changes a capacity (sat_prof,
    sat_base, q_shape) or the plume's root: the energy still closes, the water changes by evaporation less rain."""
    c = moist_aqua(rh=(0.1, 0.8, 1.3), warm=12.0)
    N = c['Ts'].size
    cloudy = opts.get('cloud', 0) > 0
    kw = dict(ch=np.full(N, 0.02)) if opts.get('sfc') else {}
    if cloudy:
        c['cw'] = np.full_like(c['q'], 1e-4)
        kw['cw'] = c['cw']
    p = par3(moist=1.0, **dict(opts, tau_conv=1e12) if cloudy else opts)
    dt = 900.0
    o = run(c, p, q=c['q'], ice=c['ice'], dt=dt, mix=mix, **kw)
    stored_matches(o, dt)
    mu1 = new_masses(c, o['wx'])
    W0 = (c['mu'] * (c['q'] + (c['cw'] if cloudy else 0))).sum(0)
    W1 = (mu1 * (o['q'] + (o['cw'] if cloudy else 0))).sum(0)
    np.testing.assert_allclose(W1 - W0, (o['evap'] - o['rain']) * dt, atol=1e-11 * W0.max())
    base = run(c, par3(moist=1.0, **({'tau_conv': 1e12, 'cloud': 1.0} if cloudy else {}),
                       **({'sfc': 1.0} if opts.get('sfc') else {})), q=c['q'], ice=c['ice'], dt=dt, mix=mix, **kw)
    assert np.abs(o['q'] - base['q']).max() > 0                 # (the option changed something)


QUIET = dict(moist=1.0, tau0=0.0, eps_s=0.0, h_sens=0.0, k_evap=0.0, theta_star=-1.0, theta_star2=-1.0)


def newton_step(q, qs, dqs_dth, pib):
    """One saturation-adjustment Newton step: the condensate (q - q_s) / (1 + L_v / pi dq_s/dtheta)."""
    return (q - qs) / (1 + 2.5e6 / pib * dqs_dth)


def test_sat_base_closed_form():
    """sat_base: the third layer condenses against saturation at its base (its lower edge's pressure and Exner
    function), one Newton step with Clausius-Clapeyron's slope."""
    c = moist_aqua(rh=(0.1, 0.5, 1.3))
    N = c['Ts'].size
    o = run(c, par3(sat_base=1.0, **QUIET), mode=1, insol=np.zeros(N), q=c['q'], ice=np.zeros(N), dt=600.0)
    pt, pb, _ = pib_of(c['mu'])
    p, pi = pb[2], CP * (pb[2] / P0) ** KAPPA
    T = pi * c['theta'][2] / CP
    qs = tetens(T, p)
    dq = (c['q'][2] - qs) / (1 + 2.5e6 * 2.5e6 * qs / (CP * 461.5 * T * T))
    np.testing.assert_allclose(o['q'][2], c['q'][2] - np.maximum(dq, 0.0), rtol=1e-12)


def spread_shape(a, b, m):
    """The truncated-exponential shape s of a layer's theta between its edges a < b with mean m:
    1 - 1/s + 1/(e^s - 1) = (m - a) / (b - a) (brentq)."""
    from scipy.optimize import brentq
    x = (m - a) / (b - a)
    f = lambda s: (0.5 + s / 12 if abs(s) < 1e-4 else 1 - 1 / s + 1 / np.expm1(s)) - x
    return brentq(f, -700, 700, xtol=1e-15, rtol=1e-15)


def test_sat_prof_closed_form():
    """sat_prof: the middle layer (below edge_k's tropopause) condenses against its capacity with its theta spread
    between its edges by the truncated exponential of its mean (the mass share m from its base at theta
    a + (b - a) ln(1 + m (e^s - 1)) / s), averaged by 4-point Gauss-Legendre over its pressure span."""
    c = moist_aqua(N=8, rh=(0.1, 1.6, 0.5))
    N = 8
    # (theta* kept at 305 and 380 K, the spread's edges; the crossing then only lifts the middle layer's air, at its
    # own vapor, into the top layer; the plumes held off)
    p = par3(moist=1.0, sat_prof=1.0, tau0=0.0, eps_s=0.0, h_sens=0.0, k_evap=0.0, tau_conv=1e30)
    o = run(c, p, mode=1, insol=np.zeros(N), q=c['q'], ice=np.zeros(N), dt=600.0)
    pt, pb, pib = pib_of(c['mu'])
    x, w = np.polynomial.legendre.leggauss(4)
    a, b = 305.0, 380.0
    for i in range(N):
        s = spread_shape(a, b, c['theta'][1, i])
        qs = dqs = 0.0
        for xj, wj in zip(x, w):
            p = 0.5 * (pt[1, i] + pb[1, i]) + 0.5 * (pb[1, i] - pt[1, i]) * xj
            m = (pb[1, i] - p) / (pb[1, i] - pt[1, i])
            tn = a + (b - a) * np.log1p(m * np.expm1(s)) / s
            ex = (p / P0) ** KAPPA
            v = tetens(tn * ex, p)
            qs += 0.5 * wj * v
            dqs += 0.5 * wj * v * 2.5e6 / (461.5 * (tn * ex) ** 2) * ex
        dq = newton_step(c['q'][1, i], qs, dqs, pib[1, i])
        np.testing.assert_allclose(o['q'][1, i], c['q'][1, i] - max(dq, 0.0), rtol=1e-10)


def test_q_shape_closed_form():
    """q_shape: the surface air holds the bottom layer's vapor at its relative humidity there, q_b q_s(T_sa, p_s) /
    q_cap (q_cap the layer's 4-point Gauss-Legendre capacity), at most saturated; evaporation is kex (q_s(T_s) - q_a)."""
    c = moist_aqua(rh=(0.1, 0.5, 0.6))
    N = c['Ts'].size
    p = par3(moist=1.0, q_shape=1.0, theta_star=-1.0, theta_star2=-1.0)
    o = run(c, p, mode=1, insol=np.zeros(N), q=c['q'], ice=np.zeros(N), dt=600.0)
    pt, pb, _ = pib_of(c['mu'])
    x, w = np.polynomial.legendre.leggauss(4)
    pp = 0.5 * (pt[2] + pb[2])[None] + 0.5 * (pb[2] - pt[2])[None] * x[:, None]
    qcap = (0.5 * w[:, None] * tetens(c['theta'][2][None] * (pp / P0) ** KAPPA, pp)).sum(0)
    ps = pb[2]
    Tsa = c['theta'][2] * (ps / P0) ** KAPPA
    qss = tetens(Tsa, ps)
    qa = np.minimum(c['q'][2] * qss / qcap, qss)
    kex = 4.0e-3 * (1 + c['uabs'] / 10.0)
    np.testing.assert_allclose(o['evap'], kex * np.maximum(tetens(c['Ts'], ps) - qa, 0.0), rtol=1e-12)


def test_freeze_dry_closed_form():
    """freeze_dry: air rising into the stratosphere (across the interface below edge_k's layer) past the cold point
    at theta* holds at most saturation there, q_s(theta* (p_i / p0)^kappa, p_i); the top layer's vapor is then its
    own plus that, by mass. (The middle layer supersaturated: its condensation heats it across its top edge and the
    cap binds; only the crossing moves air into the top.)"""
    c = moist_aqua(rh=(0.01, 1.3, 0.5))
    N = c['Ts'].size
    c['theta'][1] = 381.0 + np.linspace(0, 5, N)                # above its top edge: it rises into the top layer
    c['q'][1] = 0.15                                            # (past its capacity, down to its ~330 K base)
    p = par3(moist=1.0, freeze_dry=1.0, edge_k=1.0, tau_conv=1e30)
    o = run(c, p, q=c['q'], ice=c['ice'], dt=600.0)
    wx = o['wx'][0]
    up = wx > 0
    assert up.sum() > N // 2
    pi = GRAV * c['mu'][0]
    qsi = tetens(380.0 * (pi / P0) ** KAPPA, pi)
    assert np.all(qsi[up] < o['q'][1][up])
    top = (c['mu'][0] * c['q'][0] + wx * qsi) / (c['mu'][0] + wx)
    np.testing.assert_allclose(o['q'][0][up], top[up], rtol=1e-12)


# ---- the deep plume (plume_deep: air.cu plume_ascent with plume_ent, ent_cloud, ent_det, l_det, ice_air, lightning,
# downdrafts) ----

DEEP_OPTS = [dict(), dict(plume_ent=1e-5), dict(ent_cloud=1e-4), dict(ent_cloud=1e-4, ent_det=1.0),
             dict(plume_ent=1e-5, ent_cloud=1e-4, cloud=1.0), dict(ice_air=1.0, cloud=1.0, downdraft=0.3),
             dict(plume_root=1.0, plume_ent=1e-5, downdraft=0.3, revap=1.0, cloud=1.0),
             dict(plume_root=1.0, root_q=1.0, sfc=1.0, ent_cloud=1e-4, sat_prof=1.0, edge_vap=1.0),
             dict(buoy_v=1.0, plume_ent=1e-5), dict(ent_cloud=1e-4, ent_ifs=1.0),
             dict(ent_cloud=1e-4, ent_ifs=1.0, det_all=1.0, ent_loc=1.0, plume_ent=1e-5, buoy_v=1.0, cloud=1.0,
                  ice_air=1.0),
             dict(plume_root=1.0, root_q=1.0, sfc=1.0, shallow=0.03, ent_cloud=1e-4, ent_ifs=1.0, cloud=1.0,
                  ice_air=1.0),
             dict(plume_root=1.0, root_q=1.0, sfc=1.0, shallow=0.03, ent_cloud=1e-4, plume_ent=1e-5)]


def deep_columns(N=96, L=4, seed=5):
    """Four layers (interfaces near 150, 400 and 800 hPa), a warm moist bottom layer under a conditionally unstable
    column, cold enough aloft for ice."""
    rng = np.random.default_rng(seed)
    pos = rng.normal(size=(N, 3))
    pos /= np.linalg.norm(pos, axis=1)[:, None]
    edges = np.array([0.0, 0.15e5, 0.40e5, 0.80e5, 1.0e5])[:, None] * rng.uniform(0.97, 1.03, (1, N))
    mu = np.diff(edges, axis=0) / GRAV
    theta = np.stack([np.full(N, 420.0), np.full(N, 345.0), np.full(N, 318.0), np.full(N, 303.0)])
    theta += rng.normal(0, 1.0, (4, N))
    pt, pb, _ = pib_of(mu)
    pm = 0.5 * (pt + pb)
    q = np.array([0.01, 0.3, 0.7, 0.9])[:, None] * tetens(theta * (pm / P0) ** KAPPA, pm)
    Ts = 302.0 + rng.normal(0, 1.0, N)
    return dict(pos=pos, mu=mu, theta=theta, Ts=Ts, uabs=rng.uniform(0, 15, N), q=q, ice=np.zeros(N))


def par4(**kw):
    from insilica.atmosphere import par_array
    return par_array(4, dict(dict(theta_stars=[380.0, 330.0, 305.0], edge_k=1.0), **kw))


@pytest.mark.parametrize('opts', DEEP_OPTS)
@pytest.mark.parametrize('mix', [False, True])
def test_deep_plume_closes(opts, mix):
    """The deep plume moves air, vapor and cloud water between the layers, rains its condensate (snow by its frozen
    share), keeps l_det of it as cloud, and its downdraft evaporates rain into the bottom layer: the energy closes,
    the water changes by evaporation less rain (cloud water included: the plume's air takes its cloud water along)."""
    c = deep_columns()
    N = c['Ts'].size
    cloudy = opts.get('cloud', 0) > 0
    kw = dict(ch=np.full(N, 0.02)) if opts.get('sfc') else {}
    if cloudy:
        c['cw'] = np.full_like(c['q'], 5e-5)
        kw['cw'] = c['cw']
    p = par4(moist=1.0, plume_deep=1.0, **opts)
    dt = 900.0
    o = run(c, p, q=c['q'], ice=c['ice'], dt=dt, mix=mix, **kw)
    stored_matches(o, dt)
    mu1 = new_masses(c, o['wx'])
    W0 = (c['mu'] * (c['q'] + (c['cw'] if cloudy else 0))).sum(0)
    W1 = (mu1 * (o['q'] + (o['cw'] if cloudy else 0))).sum(0)
    np.testing.assert_allclose(W1 - W0, (o['evap'] - o['rain']) * dt, atol=1e-11 * W0.max())
    assert o['conv'].max() > 0 and o['rainp'][2].max() > 0 and np.all(mu1 >= 10.0 * (1 - 1e-9))
    # (a plume that freezes in the layer where it stops, short of -40 C, takes the -40 C level at that layer's bottom
    # (air.cu:660): with ent_ifs' ~25 hPa sub-levels the 0 C level can lie above it, and the band's log goes negative,
    # as the reference's does)
    if not opts.get('ent_ifs'):
        assert np.all(o['flash'] >= 0)


def test_deep_plume_lightning_and_downdraft():
    """Lightning needs an updraft (half sqrt(2 CAPE)) faster than graupel falls: none with graupel_v beyond every
    storm's, some at 1 m/s, and the flash rate grows as the graupel's fall speed's share of the updraft shrinks
    ((1 - v^2 / w^2) v per storm); a downdraft carries the cloud layer's air down into the bottom layer, moister than
    it came: the rain it evaporated stays in the column's vapor, which gains exactly what the rain lost."""
    c = deep_columns()
    c['cw'] = np.full_like(c['q'], 5e-5)
    fl = {v: run(c, par4(moist=1.0, plume_deep=1.0, cloud=1.0, graupel_v=v), q=c['q'], ice=c['ice'], cw=c['cw'],
                 dt=900.0)['flash'] for v in (1e3, 1.0)}
    assert np.all(fl[1e3] == 0) and fl[1.0].max() > 0
    d = run(c, par4(moist=1.0, plume_deep=1.0, downdraft=0.3), q=c['q'], ice=c['ice'], dt=900.0)
    n = run(c, par4(moist=1.0, plume_deep=1.0), q=c['q'], ice=c['ice'], dt=900.0)
    go = n['conv'] > 0
    assert go.sum() > 10
    assert np.all(d['rain'][go] <= n['rain'][go] + 1e-15) and (d['rain'][go] < n['rain'][go]).sum() > go.sum() // 2
    Wd = (new_masses(c, d['wx']) * d['q']).sum(0)
    Wn = (new_masses(c, n['wx']) * n['q']).sum(0)
    np.testing.assert_allclose((Wd - Wn)[go], ((n['rain'] - d['rain']) * 900.0)[go], rtol=1e-9, atol=1e-12)


def test_deep_plume_ifs_and_virtual_buoyancy():
    """buoy_v counts the plume's vapor in its buoyancy (theta (1 + 0.608 q)): the plume leaving the moist bottom layer
    is moister than the air it meets, so more of it rises. ent_ifs entrains 1.75e-3 /m (1.3 - RH) of the cloud
    layer's air, ~20 times Tiedtke's 1e-4 in dry air: the diluted plume rains less and stops below the layer under the
    top one, which Tiedtke's plumes reach and warm by subsidence. det_all and ent_loc act only on the tropospheric
    layers (below edge_k): with edge_k past the cloud layer they change nothing."""
    c = deep_columns()
    base = dict(moist=1.0, plume_deep=1.0, ent_cloud=1e-4)
    run_ = lambda c_, **kw: run(c_, par4(**dict(base, **kw)), q=c_['q'], ice=c_['ice'], dt=900.0)
    n, v = run_(c), run_(c, buoy_v=1.0)
    assert v['conv'].sum() > 1.05 * n['conv'].sum() and (v['conv'] > n['conv']).sum() > (n['conv'] > 0).sum() // 2
    i = run_(c, ent_ifs=1.0)
    assert i['rain'].sum() < 0.8 * n['rain'].sum()
    d1 = lambda o: np.abs(o['theta'][1] - c['theta'][1]).mean()
    assert d1(i) < 0.05 * d1(n) and d1(n) > 0.05
    a = run_(c, ent_ifs=1.0, edge_k=3.0)
    b = run_(c, ent_ifs=1.0, edge_k=3.0, det_all=1.0, ent_loc=1.0)
    for k in ('q', 'theta', 'wx', 'rain', 'conv'):
        np.testing.assert_array_equal(a[k], b[k])
    assert not np.array_equal(run_(c, ent_ifs=1.0, ent_loc=1.0)['q'], i['q'])
    wet = dict(c, q=c['q'].copy())                   # (plumes through a moister cloud layer reach the layer above)
    wet['q'][2] *= 1.3
    assert not np.array_equal(run_(wet, ent_ifs=1.0, det_all=1.0)['q'], run_(wet, ent_ifs=1.0)['q'])


def test_shallow_plume():
    """The shallow plume (shallow: a second plume from the surface air's root with the thermals' excess, stopping in
    the cloud layer) moves air only from the bottom layer into the cloud layer above it: the layers above those and
    the interfaces between them are the deep plume's alone, bit for bit; it adds its base mass flux rho a w* (cut to
    what the layers hold) to the convection, growing with a; without the root's thermals (root_q off) it is off."""
    c = deep_columns()
    N = c['Ts'].size
    base = dict(moist=1.0, plume_deep=1.0, plume_root=1.0, root_q=1.0, sfc=1.0, ent_cloud=1e-4, ent_ifs=1.0)
    run_ = lambda **kw: run(c, par4(**dict(base, **kw)), q=c['q'], ice=c['ice'], dt=900.0, ch=np.full(N, 0.02))
    n, s1, s2 = run_(), run_(shallow=0.03), run_(shallow=0.06)
    for k in ('q', 'theta'):
        np.testing.assert_array_equal(s1[k][:2], n[k][:2])
    np.testing.assert_array_equal(s1['wx'][:2], n['wx'][:2])
    go = s1['conv'] > n['conv']
    assert go.sum() > N // 4 and np.all(s1['conv'] >= n['conv'])
    d1, d2 = s1['conv'] - n['conv'], s2['conv'] - n['conv']
    assert np.all(d2 >= d1 - 1e-18)
    # where no cap binds, the flux is linear in a: twice the fraction, twice the flux, to rounding
    lin = go & np.isclose(d2, 2 * d1, rtol=1e-9, atol=0)
    assert lin.sum() > go.sum() // 4
    assert not np.array_equal(s1['q'][2], n['q'][2])
    off = run(c, par4(**dict(base, root_q=0.0, shallow=0.03)), q=c['q'], ice=c['ice'], dt=900.0, ch=np.full(N, 0.02))
    ref = run(c, par4(**dict(base, root_q=0.0)), q=c['q'], ice=c['ice'], dt=900.0, ch=np.full(N, 0.02))
    np.testing.assert_array_equal(off['q'], ref['q'])


# ---- bl_q: the mixed bottom layer's even vapor and its cloud statistics (mixed_cloud, ml_clear, cu_cloud) ----

BLQ_OPTS = [dict(), dict(ml_clear=1.0), dict(ml_clear=1.0, cu_cloud=1.0), dict(ml_clear=1.0, cu_cloud=2.0),
            dict(plume_deep=1.0, downdraft=0.3, ent_cloud=1e-4), dict(plume_deep=1.0, ml_clear=1.0, cu_cloud=1.0,
                                                                      edge_vap=1.0, q_shape=1.0)]


@pytest.mark.parametrize('opts', BLQ_OPTS)
@pytest.mark.parametrize('mix', [False, True])
def test_bl_q_closes(opts, mix):
    """With the mixed layer's cloud statistics the energy closes and the water changes by evaporation less rain
    (the simple plume held off: it moves its base layer's air without its cloud water, as the reference's does)."""
    c = deep_columns()
    N = c['Ts'].size
    c['cw'] = np.full_like(c['q'], 5e-5)
    deep = opts.get('plume_deep', 0) > 0
    p = par4(moist=1.0, cloud=1.0, bl_mix=1.0, bl_q=1.0, **dict(opts, tau_conv=1e12) if not deep else opts)
    dt = 900.0
    o = run(c, p, q=c['q'], ice=c['ice'], cw=c['cw'], cloudo=True, dt=dt, mix=mix)
    stored_matches(o, dt)
    mu1 = new_masses(c, o['wx'])
    W0 = (c['mu'] * (c['q'] + c['cw'])).sum(0)
    W1 = (mu1 * (o['q'] + o['cw'])).sum(0)
    np.testing.assert_allclose(W1 - W0, (o['evap'] - o['rain']) * dt, atol=1e-11 * W0.max())


@pytest.mark.parametrize('ml_clear', [0.0, 1.0])
def test_mixed_layer_cover(ml_clear):
    """bl_q: the mixed layer's cover is where its even vapor (and cloud water) is most saturated among its capacity's
    four Gauss-Legendre levels: by the statistics' share f = (q_t - q_s + b) / 2b, b = (1 - rh_crit) q_s, clipped to
    [0, 1]; with ml_clear 1 where q_t > q_s, else 0 (no cloud water in it, so nothing else adds cover)."""
    c = deep_columns(seed=8)
    N = c['Ts'].size
    pt, pb, _ = pib_of(c['mu'])
    x, w = np.polynomial.legendre.leggauss(4)
    pp = 0.5 * (pt[3] + pb[3])[None] + 0.5 * (pb[3] - pt[3])[None] * x[:, None]
    qsj = tetens(c['theta'][3][None] * (pp / P0) ** KAPPA, pp)
    c['q'][3] = qsj.min(0) * np.linspace(0.85, 1.1, N)             # around the top level's saturation
    cw = np.zeros_like(c['q'])
    o = run(c, par4(moist=1.0, cloud=1.0, bl_q=1.0, ml_clear=ml_clear, rh_crit=0.9), q=c['q'], ice=c['ice'], cw=cw,
            cloudo=True, dt=600.0)
    qt = c['q'][3][None]
    if ml_clear:
        f = (qt > qsj).astype(float).max(0)
    else:
        b = 0.1 * qsj
        f = np.clip((qt - qsj + b) / (2 * b), 0.0, 1.0).max(0)
    assert 0 < f.mean() < 1
    np.testing.assert_allclose(o['cloudo'][3], f, rtol=1e-12, atol=1e-15)


# ---- the cumulus layer (cu_dth: air.cu's floating cumulus layer above the mixed layer, its top at cu_dth above theta*
# or where cu_dp lifts the mixed layer's air; cl_float: the cloud layer's lower edge above that; plume_out: the plumes'
# air taken in on the way moved with them; cu_sub: the mixed layer refilled from the cumulus layer; the gas tracers
# (xfer) carried by every move of the column's air) ----

EARTH7 = dict(moist=1.0, plume_deep=1.0, sat_prof=1.0, bl_bot=0.25, cloud=1.0, q_shape=1.0, freeze_dry=1.0,
              edge_vap=1.0, plume_root=1.0, plume_ent=1e-5, conv_revap=0.05, cw_qs=0.025, bl_mix=1.0, bl_q=1.0,
              sfc=1.0, downdraft=0.3, ent_cloud=1e-4, q_sea=0.98, bl_reach=1.0, dry_edge=1.0, bl_lcl=1.0, edge_k=2.0,
              ice_air=1.0, revap=1.0, root_q=1.0, ent_det=1.0, ent_ifs=1.0, shallow=0.03, cu_dth=4.0, cu_dp=2.0e4,
              buoy_v=1.0, cu_sub=1.0, det_all=1.0, ml_clear=1.0, cu_cloud=2.0)
EARTH7_STARS = [500.0, 380.0, 330.0, 305.0, 204.0, 200.0]


def cumulus_columns(N, L, seed):
    """Columns with a mixed layer under a cumulus layer under a cloud layer: seven layers (the earth7 layering's
    stratosphere, upper troposphere, two tropospheric layers, the cloud layer, the cumulus layer, the mixed layer) or
    five (the stratosphere, the troposphere above 305 K, the cloud, cumulus and mixed layers), a moist warm bottom,
    cloud water in most layers."""
    rng = np.random.default_rng(seed)
    pos = rng.normal(size=(N, 3))
    pos /= np.linalg.norm(pos, axis=1)[:, None]
    if L == 7:
        share = [0.0, 0.05, 0.13, 0.27, 0.5, 0.78, 0.88, 1.0]
        th = [560.0, 395.0, 345.0, 318.0, 306.0, 301.5, 299.0]
        rh = [0.01, 0.05, 0.3, 0.45, 0.7, 0.8, 0.85]
    else:
        share = [0.0, 0.12, 0.4, 0.75, 0.87, 1.0]
        th = [420.0, 340.0, 306.0, 301.5, 299.0]
        rh = [0.02, 0.35, 0.7, 0.8, 0.85]
    edges = np.array(share)[:, None] * 1.0e5 * rng.uniform(0.98, 1.02, (1, N))
    mu = np.diff(edges, axis=0) / GRAV
    theta = np.array(th)[:, None] + rng.normal(0, 0.6, (L, N))
    theta[-1] -= 8.0 * pos[:, 2] ** 2                      # (colder toward the poles)
    pt, pb, _ = pib_of(mu)
    pm = 0.5 * (pt + pb)
    q = np.array(rh)[:, None] * rng.uniform(0.9, 1.1, (L, N)) * tetens(theta * (pm / P0) ** KAPPA, pm)
    cw = np.where(rng.uniform(size=(L, N)) < 0.6, rng.uniform(0, 4e-4, (L, N)), 0.0)
    Ts = 301.0 - 10.0 * pos[:, 2] ** 2 + rng.normal(0, 0.5, N)
    return dict(pos=pos, mu=mu, theta=theta, Ts=Ts, uabs=rng.uniform(0, 15, N), q=q, ice=np.zeros(N), cw=cw)


def cu_par(L, **kw):
    from insilica.atmosphere import par_array
    stars = EARTH7_STARS if L == 7 else [380.0, 305.0, 204.0, 200.0]
    return par_array(L, dict(dict(theta_stars=stars), **kw))


CU_OPTS = [(5, dict(moist=1.0, plume_deep=1.0, cu_dth=4.0, ent_cloud=1e-4, ent_ifs=1.0, bl_mix=1.0)),
           (5, dict(moist=1.0, plume_deep=1.0, cu_dth=4.0, cu_dp=2.0e4, cl_float=2.0, ent_cloud=1e-4, ent_ifs=1.0,
                    buoy_v=1.0, cloud=1.0, ice_air=1.0, downdraft=0.3, bl_mix=1.0)),
           (5, dict(moist=1.0, plume_deep=1.0, cu_dth=4.0, cu_sub=1.0, sub_mix=0.5, bl_mix=1.0, bl_lcl=1.0,
                    plume_root=1.0, root_q=1.0, sfc=1.0, shallow=0.03, ent_cloud=1e-4, ent_ifs=1.0, cloud=1.0)),
           (7, EARTH7), (7, dict(EARTH7, cl_float=2.0, sub_mix=0.5))]


@pytest.mark.parametrize('L,opts', CU_OPTS)
@pytest.mark.parametrize('mix', [False, True])
def test_cumulus_layer_closes(L, opts, mix):
    """With the cumulus layer (its floating edges, the plumes' air taken in through it moved with them, the shallow
    plume stopping in it or the cloud layer above, the mixed layer refilled from it) the energy closes, the water
    changes by evaporation less rain (cloud water included), the mass is only moved between layers, every layer keeps
    its floor, and the gas tracers carried by the moves keep each gas's column amount (sum mu x) while the cumulus
    options move air."""
    c = cumulus_columns(96, L, seed=200 + L)
    N = c['Ts'].size
    cloudy = opts.get('cloud', 0) > 0
    kw = dict(ch=np.full(N, 0.02)) if opts.get('sfc') else {}
    if cloudy:
        kw['cw'] = c['cw']
    rng = np.random.default_rng(7)
    xf = rng.uniform(0.5, 1.5, (3 * L, N)) * np.repeat([400e-6, 0.32e-6, 1.8e-6], L)[:, None]
    dt = 900.0
    o = run(c, cu_par(L, **opts), q=c['q'], ice=c['ice'], dt=dt, mix=mix, xfer=xf, **kw)
    stored_matches(o, dt)
    mu1 = new_masses(c, o['wx'])
    np.testing.assert_allclose(mu1.sum(0), c['mu'].sum(0), rtol=1e-14)
    assert np.all(mu1 >= 10.0 * (1 - 1e-9))
    W0 = (c['mu'] * (c['q'] + (c['cw'] if cloudy else 0))).sum(0)
    W1 = (mu1 * (o['q'] + (o['cw'] if cloudy else 0))).sum(0)
    np.testing.assert_allclose(W1 - W0, (o['evap'] - o['rain']) * dt, atol=1e-11 * W0.max())
    assert o['conv'].max() > 0
    for g in range(3):
        x0 = (c['mu'] * xf[g * L:(g + 1) * L]).sum(0)
        x1 = (mu1 * o['xfer'][g * L:(g + 1) * L]).sum(0)
        np.testing.assert_allclose(x1, x0, rtol=1e-13)
    assert np.any(o['xfer'] != xf)


def test_cumulus_options_act():
    """Each cumulus option changes the column: the floating edges (cu_dth, cu_dp, cl_float), the cumulus layer's
    subsidence (cu_sub, sub_mix), against the same columns without it."""
    c = cumulus_columns(96, 5, seed=211)
    base = dict(moist=1.0, plume_deep=1.0, ent_cloud=1e-4, ent_ifs=1.0, bl_mix=1.0, bl_lcl=1.0, cloud=1.0)
    out = lambda **kw: run(c, cu_par(5, **dict(base, **kw)), q=c['q'], ice=c['ice'], cw=c['cw'], dt=900.0)
    ref = out()
    steps = [dict(cu_dth=4.0), dict(cu_dth=4.0, cu_dp=2.0e4), dict(cu_dth=4.0, cu_dp=2.0e4, cl_float=2.0),
             dict(cu_dth=4.0, cu_sub=1.0), dict(cu_dth=4.0, cu_sub=1.0, sub_mix=0.5)]
    prev = ref
    for kw in steps:
        o = out(**kw)
        assert np.any(o['wx'] != prev['wx']) or np.any(o['theta'] != prev['theta']), kw
        prev = o


# ---- the diagnostics rows (dgo: air.cu:1286-1290, 1944-1950): per layer the heating Q and the latent heating (W/m^2),
# per interface the dry crossing's gross up and down masses, per layer the convective mass delivered into it ----

def test_diagnostics_rows():
    """The rows add up to what the column did: without clouds or plumes the latent heating sums to Lv times the rain
    (every gram condensed rains, each releasing Lv); where no floor limits the crossing each interface's net crossing
    is gross up less gross down; the parcel plume's (and the deep plume's, without the cloud layer's entrainment)
    delivered mass sums to the convective mass conv; the rows are zero where nothing happens (a dry column's latent
    heating and convective mass)."""
    c = moist_aqua(rh=(0.1, 0.8, 1.3), warm=12.0)
    N, dt = c['Ts'].size, 900.0
    o = run(c, par3(moist=1.0, tau_conv=1e30), q=c['q'], ice=c['ice'], dt=dt, dgo=True)
    L = 3
    assert o['dgo'].shape == (5 * L - 2, N)
    Q, dgl, gross, cin = o['dgo'][:L], o['dgo'][L:2 * L], o['dgo'][2 * L:4 * L - 2], o['dgo'][4 * L - 2:]
    assert o['rain'].max() > 0 and o['conv'].max() == 0
    from insilica.atmosphere import PAR_ORDER
    lv = par3()[PAR_ORDER.index('Lv')]
    np.testing.assert_allclose(dgl.sum(0), lv * o['rain'], rtol=1e-12, atol=1e-12 * lv * o['rain'].max())
    assert np.all(cin == 0)
    net = gross[0::2] - gross[1::2]
    mu1 = new_masses(c, o['wx'])
    free = np.all(mu1 > 10.0 * 1.01, axis=0)            # (no layer at its floor: the crossing was not limited)
    assert free.sum() > N // 2
    np.testing.assert_allclose(o['wx'][:, free], net[:, free], rtol=1e-12, atol=1e-12 * np.abs(net).max())
    # the plumes: the parcel plume, then the deep plume without the cloud layer's entrainment
    for extra in (dict(), dict(plume_deep=1.0)):
        d = deep_columns()
        p = par4(moist=1.0, **extra)
        o = run(d, p, q=d['q'], ice=d['ice'], dt=dt, dgo=True)
        L = 4
        cin = o['dgo'][4 * L - 2:]
        assert o['conv'].max() > 0
        np.testing.assert_allclose(cin.sum(0), o['conv'], rtol=1e-13, atol=0)
    # a dry column: no latent heating, no convective mass
    a = aqua()
    o = run(a, par3(), dt=dt, dgo=True)
    L = 3
    assert np.all(o['dgo'][L:2 * L] == 0) and np.all(o['dgo'][4 * L - 2:] == 0) and np.any(o['dgo'][:L] != 0)
