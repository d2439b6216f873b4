"""The oracle's plume ascent (CPU) against closed forms: the reference's deep and shallow plume (air.cu
plume_ascent, air.cu:540-666) with each of its options alone, on columns built so that the answer follows from
the physics the options state, not from the code's own steps.

The columns: four layers whose two middle ones hold theta linear in mass between their edges (a layer's mean at
the midpoint of its edges makes the truncated-exponential profile's shape exactly 0, so the environment at mass
share x is a + (b - a) x), and a top layer of uniform theta. A sub-level's thickness is the hypsometric dz =
R_d T dp / (g p) at its midpoint, T = theta (p / p0)^kappa. What each check pins:

  * no option: theta conserved (dry adiabatic); the buoyancy A = mu_bot (theta_p - theta*) / 2 (the bottom layer's
    stretch a triangle) + sum over the buoyant sub-levels of (mass) (theta_p - theta_e); which sub-levels are buoyant
    fixes the layers' orientation (their lower edge theta* below);
  * buoy_v: the same in virtual theta, theta (1 + 0.608 q), the environment's vapor the layer's;
  * plume_ent (mixing per unit pressure, lambda): on a uniform layer the plume's theta relaxes as e^(-lambda dp)
    (a geometric series for A); the mixing exchanges equal masses, so the plume's vapor plus what it gave the layers
    is its starting vapor;
  * ent_cloud (entrainment per metre, eps, in the layer above the bottom one): the plume's mass grows as
    e^(eps z); its vapor and theta fluxes (mass times value) change by exactly the entrained mass times the
    environment's; with ent_det the mass stays 1 and the exchange is equal-mass;
  * ent_ifs (Bechtold et al. 2008, 2014, as the reference states them): entrainment 1.75e-3 /m (1.3 - RH)
    (q_s / q_s at cloud base)^3 and detrainment min(that, 0.75e-4 /m), the mass growing as e^((eps - delta) dz),
    over sub-levels of at most ~25 hPa (at least 3), the first from cloud base to its middle (half a dz); RH the
    environment's vapor over its saturation (Tetens over water); the shallow plume's rate 3.5e-3 and its stop in
    the cloud layer; det_all: every layer from edge_k down entraining; ent_loc: the air taken in at the layer's
    relative humidity, so RH is the layer's vapor over the mean of its sub-levels' saturation.

The mixed build (float32 saturation, exponentials and powers) is held to the same forms at 1e-5.
"""
import math

import numpy as np
import pytest

from oracle import column as ocol

from insilica.atmosphere import PAR_ORDER, par_array

G_ = 9.80616
P0 = 1.0e5
KAPPA = 0.285725661955007
RD = 287.0
assert par_array(4, dict(theta_stars=[380.0, 350.0, 305.0]))[PAR_ORDER.index('grav')] == G_
assert par_array(4, dict(theta_stars=[380.0, 350.0, 305.0]))[PAR_ORDER.index('kappa')] == KAPPA


def tetens(T, p, eps=0.622):
    return eps * 611.0 * math.exp(17.27 * (T - 273.15) / (T - 35.85)) / p


def par(stars, **kw):
    return par_array(4, dict(dict(theta_stars=list(stars), edge_k=1.0, moist=1.0), **kw))


def column(stars, mu, top, th_bot):
    """theta of the four layers: the top uniform, the middle two at their edges' midpoint (linear in mass)."""
    s0, s1, s2 = stars
    return np.array([top, 0.5 * (s1 + s0), 0.5 * (s2 + s1), th_bot]), np.asarray(mu, float)


def sub_levels(mu, k, stars, n, pt):
    """Layer k's n sub-levels from the bottom up: (pressure, Exner, theta_e, dz) at each midpoint, pt the pressure
    at the layer's bottom edge (k = 1, 2: theta linear from its lower edge theta* to its upper)."""
    a, b = (stars[2], stars[1]) if k == 2 else (stars[1], stars[0])
    out = []
    for j in range(n):
        x = (j + 0.5) / n
        p = pt - x * G_ * mu[k]
        ex = (p / P0) ** KAPPA
        te = a + (b - a) * x
        out.append((p, ex, te, RD * te * ex * (G_ * mu[k] / n) / (G_ * p)))
    return out


def run(p, th, mu, q, th0, q0, shal=False, mix=False):
    pi_ = G_ * float(np.sum(mu[:3]))
    return ocol.plume_ascent(p, mu, th, q, th0, q0, pi_, shal=shal, mix=mix), pi_


TOL = {False: 1e-12, True: 1e-5}


@pytest.mark.parametrize('mix', [False, True])
def test_adiabatic_buoyancy(mix):
    """No option: a dry plume keeps its theta; A sums its excess by mass up to the last buoyant sub-level (three per
    layer), the bottom layer's half; it stops in the layer whose air beats it."""
    stars = (380.0, 350.0, 305.0)
    th, mu = column(stars, [2000.0, 3000.0, 3000.0, 2000.0], 390.0, 300.0)
    q = np.zeros(4)
    p = par(stars)
    # stops in layer 2: of its sub-levels' 312.5, 327.5, 342.5 K it beats the first two
    o, _ = run(p, th, mu, q, 340.0, 0.0, mix=mix)
    A = 0.5 * mu[3] * (340.0 - 305.0) + mu[2] / 3 * ((340.0 - 312.5) + (340.0 - 327.5))
    assert o['kd'] == 2 and o['thd'] == 340.0 and o['G'] == 1.0 and not np.any(o['qx'])
    assert abs(o['A'] - A) <= TOL[mix] * A
    # beats every layer: the whole column's excess
    o, _ = run(p, th, mu, q, 400.0, 0.0, mix=mix)
    A = 0.5 * mu[3] * (400 - 305.0) + mu[2] * (400 - 327.5) + mu[1] * (400 - 365.0) + mu[0] * (400 - 390.0)
    assert o['kd'] == 0 and o['thd'] == 400.0
    assert abs(o['A'] - A) <= TOL[mix] * A


@pytest.mark.parametrize('mix', [False, True])
def test_virtual_buoyancy(mix):
    """buoy_v: theta (1 + 0.608 q) on both sides (the bottom triangle against the layer above's vapor). A plume at
    342.3 K with 5 g/kg beats the 342.5 K sub-level of air holding 3 g/kg only in virtual theta."""
    stars = (380.0, 350.0, 305.0)
    th, mu = column(stars, [2000.0, 3000.0, 3000.0, 2000.0], 390.0, 300.0)
    q = np.array([0.0, 1e-3, 3e-3, 4e-3])
    q0, th0 = 5e-3, 342.3
    tes = (312.5, 327.5, 342.5)
    for bv in (0.0, 1.0):
        o, _ = run(par(stars, buoy_v=bv), th, mu, q, th0, q0, mix=mix)
        vf = (lambda qq: 1 + 0.608 * qq) if bv else (lambda qq: 1.0)
        tpv = th0 * vf(q0)
        A = 0.5 * mu[3] * (tpv - 305.0 * vf(q[2])) + sum(mu[2] / 3 * (tpv - te * vf(q[2])) for te in tes
                                                         if tpv > te * vf(q[2]))
        assert o['kd'] == 2 and o['qd'] == q0          # (unsaturated all the way)
        assert abs(o['A'] - A) <= TOL[mix] * abs(A), (bv, o['A'], A)
    # the virtual form counts the third sub-level, the dry form does not
    assert th0 * (1 + 0.608 * q0) > 342.5 * (1 + 0.608 * q[2]) and th0 < 342.5


@pytest.mark.parametrize('mix', [False, True])
def test_mixing_relaxes_and_exchanges_equal_masses(mix):
    """plume_ent: through a uniform top layer (two layers: the plume starts under the top one) the plume's theta
    relaxes as e^(-lambda dp) over its eight sub-levels, A the geometric series; in the four-layer column the
    vapor it holds where it stops plus what it gave each layer is its starting vapor."""
    lam = 2e-5
    p2 = par_array(2, dict(theta_star=300.0, moist=1.0, plume_ent=lam))
    mu2, th2 = np.array([4000.0, 2000.0]), np.array([310.0, 290.0])
    pi2 = G_ * mu2[0]
    o = ocol.plume_ascent(p2, mu2, th2, np.zeros(2), 330.0, 0.0, pi2, mix=mix)
    r = math.exp(-lam * G_ * mu2[0] / 8)
    A = 0.5 * mu2[1] * (330.0 - 300.0) + mu2[0] / 8 * 20.0 * r * (1 - r ** 8) / (1 - r)
    thd = 310.0 + 20.0 * math.exp(-lam * G_ * mu2[0])
    assert o['kd'] == 0
    assert abs(o['thd'] - thd) <= TOL[mix] * thd and abs(o['A'] - A) <= TOL[mix] * A
    stars = (380.0, 350.0, 305.0)
    th, mu = column(stars, [2000.0, 3000.0, 3000.0, 2000.0], 390.0, 300.0)
    q = np.array([0.0, 1e-3, 3e-3, 4e-3])
    o, _ = run(par(stars, plume_ent=lam), th, mu, q, 345.0, 5e-3, mix=mix)
    assert o['kd'] == 2 and 3e-3 < o['qd'] < 5e-3
    assert abs(o['qd'] + o['qx'].sum() - 5e-3) <= 1e-15 + TOL[mix] * 5e-3


def _entraining(mix, ent_det):
    stars = (380.0, 350.0, 305.0)
    th, mu = column(stars, [2000.0, 3000.0, 3000.0, 2000.0], 390.0, 300.0)
    q = np.array([0.0, 1e-4, 2e-4, 3e-4])
    eps = 1e-4
    o, pi_ = run(par(stars, ent_cloud=eps, ent_det=ent_det), th, mu, q, 375.0, 5e-4, mix=mix)
    return o, sub_levels(mu, 2, stars, 3, pi_), eps, q


@pytest.mark.parametrize('mix', [False, True])
def test_cloud_entrainment_grows_the_plume(mix):
    """ent_cloud: through the layer above the bottom one the plume's mass grows as e^(eps z) (hypsometric z), and
    mass times vapor and mass times theta change by exactly the entrained mass times the environment's."""
    o, subs, eps, q = _entraining(mix, 0.0)
    assert o['kd'] == 1                                # (buoyant past the entraining layer)
    z = sum(s[3] for s in subs)
    G = math.exp(eps * z)
    assert abs(o['G'] - G) <= 10 * TOL[mix] * G
    assert abs(o['G'] * o['qd'] - (5e-4 + (o['G'] - 1) * q[2])) <= TOL[mix] * 5e-4
    M, flux = 1.0, 375.0
    for s in subs:
        M1 = M * math.exp(eps * s[3])
        flux += (M1 - M) * s[2]
        M = M1
    assert abs(o['G'] * o['thd'] - flux) <= TOL[mix] * flux
    # ent_det: the same exchange at constant mass, equal masses in and out
    o, _, _, _ = _entraining(mix, 1.0)
    assert o['G'] == 1.0 and o['kd'] == 1
    assert abs(o['qd'] + o['qx'].sum() - 5e-4) <= TOL[mix] * 5e-4


def _ifs_growth(subs, qk, qsb, rate, loc=False, pick=None):
    """The mass growth through the sub-levels subs under the IFS's rates."""
    qse = [tetens(te * ex, p) for p, ex, te, _ in subs]
    qsl = sum(qse) / len(qse)
    lnG = 0.0
    for j, ((p, ex, te, dz), qs) in enumerate(zip(subs, qse)):
        if pick is not None and j > pick:
            break
        rh = min(1.0, max(0.0, qk / qsl if loc else qk / qs))
        fs = min(1.0, qs / qsb)
        e = rate * (1.3 - rh) * fs ** 3
        d = min(e, 0.75e-4)
        lnG += (e - d) * (0.5 * dz if j == 0 else dz)
    return math.exp(lnG)


IFS = (312.0, 310.0, 305.0)                            # a near-neutral cloud layer under a near-neutral layer


def _ifs_column(mu2, mu1=600.0, qk=0.0, qk1=0.0, th_bot=150.0):
    th, mu = column(IFS, [3000.0, mu1, mu2, 2000.0], 330.0, th_bot)
    q = np.array([0.0, qk1, qk, 0.0])
    return th, mu, q


@pytest.mark.parametrize('mix', [False, True])
def test_ifs_entrainment_dry(mix):
    """ent_ifs in dry air under a cold cloud base (q_s / q_s at base >= 1, so the factor is 1): eps = 1.75e-3 * 1.3,
    delta = 0.75e-4, over ceil(g mu / 2500 Pa) = 4 sub-levels of the 800 kg/m^2 cloud layer."""
    th, mu, q = _ifs_column(800.0)
    o, pi_ = run(par(IFS, ent_cloud=1e-4, ent_ifs=1.0), th, mu, q, 340.0, 0.0, mix=mix)
    assert o['kd'] <= 1
    n = math.ceil(G_ * mu[2] / 2500.0)
    assert n == 4
    G = _ifs_growth(sub_levels(mu, 2, IFS, n, pi_), 0.0, 1e-30, 1.75e-3)
    assert abs(o['G'] - G) <= 10 * TOL[mix] * G, (o['G'], G)
    assert abs(math.log(G) - (1.75e-3 * 1.3 - 0.75e-4) * sum(s[3] for s in sub_levels(mu, 2, IFS, n, pi_)) * (
        1 - 0.5 * sub_levels(mu, 2, IFS, n, pi_)[0][3] / sum(s[3] for s in sub_levels(mu, 2, IFS, n, pi_)))) < 1e-12


@pytest.mark.parametrize('mix', [False, True])
@pytest.mark.parametrize('loc', [False, True])
def test_ifs_entrainment_moist(mix, loc):
    """ent_ifs in moist air under a warm base: RH of the environment's vapor (level by level, or with ent_loc the
    layer's vapor over its sub-levels' mean saturation) and the cube of q_s / q_s at cloud base."""
    th, mu, q = _ifs_column(1600.0, qk=4e-4, th_bot=302.0)
    o, pi_ = run(par(IFS, ent_cloud=1e-4, ent_ifs=1.0, ent_loc=float(loc)), th, mu, q, 345.0, 0.0, mix=mix)
    assert o['kd'] <= 1
    n = math.ceil(G_ * mu[2] / 2500.0)
    subs = sub_levels(mu, 2, IFS, n, pi_)
    qsb = tetens(th[3] * (pi_ / P0) ** KAPPA, pi_)
    fs = [tetens(te * ex, p) / qsb for p, ex, te, _ in subs]
    rh = [q[2] / tetens(te * ex, p) for p, ex, te, _ in subs]
    assert n == 7 and min(fs) < 0.5 < 1 < max(fs) and 0.2 < min(rh) < 1 < max(rh)   # (both sides of the caps)
    G = _ifs_growth(subs, q[2], qsb, 1.75e-3, loc=loc)
    assert abs(o['G'] - G) <= 10 * TOL[mix] * G, (o['G'], G)
    other = _ifs_growth(subs, q[2], qsb, 1.75e-3, loc=not loc)
    assert abs(other - G) > 1e-6 * G                   # (the two RH rules differ here)


@pytest.mark.parametrize('mix', [False, True])
def test_shallow_plume(mix):
    """The shallow plume: twice the deep one's entrainment (3.5e-3), and it stops in the cloud layer though buoyant
    above it."""
    th, mu, q = _ifs_column(800.0, qk=4e-4, th_bot=302.0)
    p = par(IFS, ent_cloud=1e-4, ent_ifs=1.0)
    deep, pi_ = run(p, th, mu, q, 345.0, 0.0, mix=mix)
    o, _ = run(p, th, mu, q, 345.0, 0.0, shal=True, mix=mix)
    assert deep['kd'] <= 1 and o['kd'] == 2
    subs = sub_levels(mu, 2, IFS, 4, pi_)
    qsb = tetens(th[3] * (pi_ / P0) ** KAPPA, pi_)
    G = _ifs_growth(subs, q[2], qsb, 3.5e-3)
    assert abs(o['G'] - G) <= 10 * TOL[mix] * G, (o['G'], G)


@pytest.mark.parametrize('mix', [False, True])
def test_det_all(mix):
    """det_all: with edge_k = 1 both tropospheric layers entrain (each in its own sub-levels, each starting with half
    a dz); with edge_k = 2 only the cloud layer does, as without det_all."""
    th, mu, q = _ifs_column(800.0, mu1=1100.0, qk=4e-4, qk1=2e-4, th_bot=302.0)
    th[0] = 312.5
    o, pi_ = run(par(IFS, ent_cloud=1e-4, ent_ifs=1.0, det_all=1.0), th, mu, q, 360.0, 0.0, mix=mix)
    assert o['kd'] == 0                                # (buoyant past both entraining layers)
    qsb = tetens(th[3] * (pi_ / P0) ** KAPPA, pi_)
    n2, n1 = math.ceil(G_ * mu[2] / 2500.0), max(3, math.ceil(G_ * mu[1] / 2500.0))
    assert (n2, n1) == (4, 5)
    G = (_ifs_growth(sub_levels(mu, 2, IFS, n2, pi_), q[2], qsb, 1.75e-3)
         * _ifs_growth(sub_levels(mu, 1, IFS, n1, pi_ - G_ * mu[2]), q[1], qsb, 1.75e-3))
    assert abs(o['G'] - G) <= 10 * TOL[mix] * G, (o['G'], G)
    o2, _ = run(par(IFS, ent_cloud=1e-4, ent_ifs=1.0, det_all=1.0, edge_k=2.0), th, mu, q, 360.0, 0.0, mix=mix)
    G2 = _ifs_growth(sub_levels(mu, 2, IFS, n2, pi_), q[2], qsb, 1.75e-3)
    assert o2['kd'] <= 1 and abs(o2['G'] - G2) <= 10 * TOL[mix] * G2


@pytest.mark.parametrize('mix', [False, True])
def test_deep_plume_mass_closure(mix):
    """The column's deep plume (plume_deep) moves the mass that would remove its buoyancy, A / (theta_d - theta*)
    (theta_d - theta* at least 1 K), at most the bottom layer's air above its floor, relaxed over tau_conv by
    1 - e^(-dt / tau_conv), from the bottom layer to the layer it detrains into: with nothing else acting (no
    radiation, no surface exchange, no crossing), the mass across every interface from there down is that mass,
    the plume's A and theta_d those of its ascent from the bottom layer's air."""
    rng = np.random.default_rng(3)
    N = 8
    pos = rng.normal(size=(N, 3))
    pos /= np.linalg.norm(pos, axis=1)[:, None]
    edges = np.array([0.0, 0.15e5, 0.40e5, 0.80e5, 1.0e5])[:, None] * rng.uniform(0.97, 1.03, (1, N))
    mu = np.diff(edges, axis=0) / G_
    theta = np.array([420.0, 345.0, 318.0, 303.0])[:, None] + rng.normal(0, 1.0, (4, N))
    pm = 0.5 * (edges[:-1] + edges[1:])
    T = theta * (pm / P0) ** KAPPA
    q = np.array([0.01, 0.3, 0.7, 0.9])[:, None] * np.vectorize(tetens)(T, pm)
    dt, tau = 900.0, 7200.0
    p = par_array(4, dict(theta_stars=[380.0, 330.0, 305.0], edge_k=1.0, moist=1.0, plume_deep=1.0, tau0=0.0,
                          eps_s=0.0, h_sens=0.0, k_evap=0.0, tau_conv=tau))
    o = ocol.air_column(p, 1, np.array([0.0]), pos, np.zeros(N), mu, theta, np.full(N, 250.0), dt,
                        insol=np.zeros(N), q=q, ice=np.zeros(N), mix=mix)
    floor = p[PAR_ORDER.index('floor_')]
    fired = 0
    for i in range(N):
        a = ocol.plume_ascent(p, mu[:, i], theta[:, i], q[:, i], theta[3, i], q[3, i], G_ * mu[:3, i].sum(), mix=mix)
        if a['kd'] < 0 or a['A'] <= 0:
            assert not np.any(o['wx'][:, i])
            continue
        mc = min(a['A'] / max(a['thd'] - 305.0, 1.0), mu[3, i] - floor) * (1 - math.exp(-dt / tau))
        assert abs(o['wx'][2, i] - mc) <= TOL[mix] * mc, (i, o['wx'][:, i], mc)
        for j in range(3):
            assert o['wx'][j, i] == (o['wx'][2, i] if j >= a['kd'] else 0.0)
        fired += 1
    assert fired >= N // 2
