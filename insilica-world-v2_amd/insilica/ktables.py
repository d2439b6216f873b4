"""Synthetic k-tables in the reference's table format, and the column states of the radiation workload.

The reference fits its tables to line-by-line spectra (its tools/make_ktables.py, not part of the model's step);
these tables are seeded synthetic ones of any size, for the 40,962-column workload with 16 x 16 longwave and 14 x 16
shortwave g-points: absorption coefficients log-uniform between 1e-6 and 1e2 m^2/kg, stored as the reference stores
them, ln of cm^2 per molecule (k[cm^2/molecule] = k[m^2/kg] * M / N_A * 1e4), with a mild pressure and temperature
dependence; the Planck table per g-point is the band's Planck radiance (W/m^2/sr) split evenly over its g-points, so
that pi times its sum over g-points is sigma T^4 over the bands' span; the solar weights are a 5778 K sun's share per
g-point. Axes as the reference's: 31 pressures from 1100 hPa to 1 Pa (log-spaced), temperatures 150-340 K by 10 K
(longwave) or 150-330 K by 20 K (shortwave), the Planck table 140-350 K by 1 K.
"""
import numpy as np

NA = 6.02214076e23
H, C, KB = 6.62607015e-34, 2.99792458e8, 1.380649e-23
# molar masses (kg/mol) of the table components: H2O (two tables), CO2, O3, N2O, CH4 and, in the shortwave, O2
M_COMP = (18.015e-3, 18.015e-3, 44.01e-3, 48.0e-3, 44.013e-3, 16.04e-3, 32.0e-3)


def _planck_band(T, lo, hi, n=96):
    """Band-integrated Planck radiance (W/m^2/sr) between lo and hi cm^-1 at temperatures T."""
    nu = np.linspace(lo, hi, n) * 100.0                          # m^-1
    x = np.outer(np.atleast_1d(T), np.ones_like(nu))
    b = 2 * H * C ** 2 * nu ** 3 / np.expm1(H * C * nu / (KB * x))
    return np.trapezoid(b, nu, axis=1)


def _axes():
    P = np.exp(np.linspace(np.log(1.1e5), 0.0, 31))
    return P


def _logk(rng, ncomp, NP, NT, ng_total, P, T):
    """ln k (cm^2/molecule): per component and g-point a base k log-uniform in [1e-6, 1e2] m^2/kg, sorted within each
    run of 16 g-points (a k-distribution), times (p / 1e4 Pa)^a (a in 0.3-0.9) and exp(b (T - 250) / 100)."""
    out = np.empty((ncomp, NP, NT, ng_total), np.float32)
    lp = np.log(P / 1.0e4)[:, None, None]
    tt = ((T - 250.0) / 100.0)[None, :, None]
    for c in range(ncomp):
        base = np.exp(rng.uniform(np.log(1e-6), np.log(1e2), ng_total))
        base = np.sort(base.reshape(-1, 16), axis=1).reshape(-1) if ng_total % 16 == 0 else base
        kcm = base * M_COMP[c] / NA * 1e4
        a = rng.uniform(0.3, 0.9, ng_total)[None, None, :]
        b = rng.uniform(-1.0, 1.0, ng_total)[None, None, :]
        out[c] = (np.log(kcm)[None, None, :] + a * lp + b * tt).astype(np.float32)
    return out


def synthetic_lw(bands=16, g_per_band=16, seed=0, lo=10.0, hi=3000.0):
    """Longwave tables: logk (6, 31, 20, NG), planck (211, NG), P, T, T_pl, o3 (31), ng, bands."""
    rng = np.random.default_rng(seed)
    ng = bands * g_per_band
    P, T, Tp = _axes(), np.arange(150.0, 341.0, 10.0), np.arange(140.0, 351.0, 1.0)
    edges = np.linspace(lo, hi, bands + 1)
    planck = np.empty((len(Tp), ng), np.float32)
    for b in range(bands):
        pb = _planck_band(Tp, edges[b], edges[b + 1])
        planck[:, b * g_per_band:(b + 1) * g_per_band] = (pb / g_per_band)[:, None]
    # ozone: a peak of 8 ppm near 10 hPa in log pressure
    o3 = (8e-6 * np.exp(-0.5 * ((np.log(P) - np.log(1.0e3)) / 1.2) ** 2) + 3e-8).astype(np.float32)
    return dict(logk=_logk(rng, 6, len(P), len(T), ng, P, T), planck=planck, P=P, T=T, T_pl=Tp, o3=o3,
                ng=np.full(bands, g_per_band), bands=edges)


def synthetic_sw(bands=14, g_per_band=16, seed=1, lo=2600.0, hi=50000.0):
    """Shortwave tables: logk (7, 31, 10, NG), P, T, sun, ray, haze (NG), ng, haze_par, band_top, cloud_ssa."""
    rng = np.random.default_rng(seed)
    ng = bands * g_per_band
    P, T = _axes(), np.arange(150.0, 331.0, 20.0)
    edges = np.exp(np.linspace(np.log(lo), np.log(hi), bands + 1))
    mid = 0.5 * (edges[:-1] + edges[1:])
    sun_b = np.array([_planck_band(5778.0, edges[b], edges[b + 1])[0] for b in range(bands)])
    lam_um = 1e4 / mid
    ray_b = 4.02e-28 / lam_um ** 4                               # cm^2 per molecule (Rayleigh, ~lambda^-4)
    haze_b = (lam_um / 0.55) ** -1.3
    rep = lambda a: np.repeat(a, g_per_band) / (g_per_band if a is sun_b else 1)
    cloud_ssa = tuple(float(x) for x in np.clip(1 - 0.4 * np.exp(-(mid - lo) / 1500.0), 0.6, 0.99999))
    return dict(logk=_logk(rng, 7, len(P), len(T), ng, P, T), P=P, T=T, sun=rep(sun_b), ray=rep(ray_b),
                haze=rep(haze_b), ng=np.full(bands, g_per_band), haze_par=np.array([0.12, 0.95, 0.7, 1500.0]),
                band_top=tuple(float(x) for x in edges[1:]), cloud_ssa=cloud_ssa)


def column_state(pos, layers=7, seed=0, grav=9.80616, kappa=287.04 / 1004.6, p0=1.0e5):
    """Seeded column states of the radiation workload at cell positions pos (N, 3): surface pressure 980-1030 hPa,
    layer edges at fixed shares of it, temperature by lapse rate (surface 250-305 K by latitude and noise, 6.5 K/km up
    to a 180-210 K tropopause, isothermal above), vapor q = q_s exp(-z / 2 km) with q_s 80% of saturation at the
    surface, clouds in about a third of the layers. Returns host float64 arrays: mu (L, N) kg/m^2, theta (L, N),
    q (L, N), Ts (N), alb (N), cloudo (4L + 3, N)."""
    rng = np.random.default_rng(seed)
    N, L = len(pos), layers
    lat = np.arcsin(np.clip(pos[:, 2], -1, 1))
    ps = rng.uniform(9.8e4, 1.03e5, N)
    share = np.concatenate([[0.0], np.geomspace(0.02, 1.0, L)])   # edges from the top
    pe = share[:, None] * ps[None, :]
    mu = np.diff(pe, axis=0) / grav
    Ts = 250.0 + 55.0 * np.cos(lat) ** 2 + rng.normal(0.0, 2.0, N)
    Ttp = rng.uniform(180.0, 210.0, N)
    pm = 0.5 * (pe[1:] + pe[:-1])
    z = 7000.0 * np.log(ps[None, :] / pm)                          # height of the layer middles (scale height 7 km)
    T = np.maximum(Ts[None, :] - 6.5e-3 * z, Ttp[None, :])
    theta = T * (p0 / pm) ** kappa
    qs = 0.622 * 611.0 * np.exp(17.27 * (Ts - 273.15) / (Ts - 35.85)) / ps
    q = 0.8 * qs[None, :] * np.exp(-z / 2000.0)
    alb = rng.uniform(0.06, 0.3, N)
    cloudo = np.zeros((4 * L + 3, N))
    cov = np.where(rng.uniform(size=(L, N)) < 0.35, rng.uniform(0.1, 0.9, (L, N)), 0.0)
    cov[0] = 0.0
    cloudo[:L] = cov
    cloudo[L + 3:2 * L + 3] = np.where(cov > 0, rng.uniform(0.5, 20.0, (L, N)), 0.0)       # longwave optical depth
    cloudo[2 * L + 3:3 * L + 3] = np.where(cov > 0, rng.uniform(1.0, 40.0, (L, N)), 0.0)   # shortwave optical depth
    cloudo[3 * L + 3:4 * L + 3] = np.where(cov > 0, rng.uniform(0.0, 1.0, (L, N)), 1.0)    # the liquid's share
    return dict(mu=mu, theta=theta, q=q, Ts=Ts, alb=alb, cloudo=cloudo)
