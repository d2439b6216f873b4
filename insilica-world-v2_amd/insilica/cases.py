"""Initial states: (h (L, N), b (N), u (L, E)) NumPy arrays, plus extras (theta, Ts, q) for air.

The one-layer benchmarks are Williamson et al. (1992) 2 and 5 and the Galewsky et al. (2004) jet, built the way the
reference builds them (src/insilica/cases.py:15-104): winds sampled at the Voronoi edge midpoint xm and projected on
the edge normal; the same closed forms. The layered starts (a gravity-wave bump, a two-layer internal wave, the
balanced two-layer jet of the Phillips problem) and the three-layer aquaplanet air (src/insilica/cases.py:227-275)
likewise. The two benchmark workloads of this package are built here too:

  * `jet_bumps`: one layer 8 km deep in a balanced 20 cos(lat) m/s zonal jet with three seeded 100 m, 500 km
    Gaussian bumps;
  * `coupled_stack`: three air layers over two ocean layers with a seeded 30% land mask and random-phase
    spherical-harmonic perturbations of degree <= 20 (+-2 K on theta, +-5 m/s on the wind).
"""
import math

import numpy as np

from .constants import CP_DRY, GRAV, KAPPA, OMEGA, P0, R_PLANET
from .grid import planet_of

MU_REF = 1.0e5 / GRAV             # kg/m^2: the column mass of 1000 hPa
AQUA_P = (0.22e5, 0.60e5)         # Pa: the three air layers' starting interfaces
SURFACE_DRAG = 0.15 * MU_REF / 86400.0


def edge_east_north(g):
    """Latitude, longitude and the east and north components of each edge normal, at the edge midpoints xm."""
    x = g.xm
    lat = np.arcsin(np.clip(x[:, 2], -1, 1))
    lon = np.arctan2(x[:, 1], x[:, 0])
    en = -np.sin(lon) * g.ne[:, 0] + np.cos(lon) * g.ne[:, 1]
    nn = (-np.sin(lat) * np.cos(lon) * g.ne[:, 0] - np.sin(lat) * np.sin(lon) * g.ne[:, 1] + np.cos(lat) * g.ne[:, 2])
    return lat, lon, en, nn


def great_circle(g, lon_c, lat_c):
    c = np.array([math.cos(lat_c) * math.cos(lon_c), math.cos(lat_c) * math.sin(lon_c), math.sin(lat_c)])
    return R_PLANET * np.arccos(np.clip(g.pos @ c, -1, 1))


def cosine_bell(g, lon_c=3 * math.pi / 2, lat_c=0.0, radius=R_PLANET / 3, height=1000.0):
    """Williamson 1's cosine bell: height/2 (1 + cos(pi r / R)) within great-circle distance R."""
    d = great_circle(g, lon_c, lat_c)
    return np.where(d < radius, 0.5 * height * (1 + np.cos(math.pi * d / radius)), 0.0)


def default_dt(n):
    """450 s at n = 4, halved per level."""
    return 450.0 * 2.0 ** (4 - n)


def williamson(g, case):
    """One layer (h, b, u) of Williamson 2 (steady zonal flow) or 5 (zonal flow over a mountain)."""
    a, Om, G = planet_of(g)
    lat_e, _, en, _ = edge_east_north(g)
    s = np.sin(g.lat)
    b = np.zeros(g.N)
    if case == 'w2':
        u0, gh0 = 2 * math.pi * a / (12 * 86400), 2.94e4
        h = (gh0 - (a * Om * u0 + u0 * u0 / 2) * s * s) / G
    elif case == 'w5':
        u0, h0, Rm, lc, pc = 20.0, 5960.0, math.pi / 9, -math.pi / 2, math.pi / 6
        dl = np.arctan2(np.sin(g.lon - lc), np.cos(g.lon - lc))
        b = 2000 * (1 - np.minimum(Rm, np.hypot(dl, g.lat - pc)) / Rm)
        h = h0 - (a * Om * u0 + u0 * u0 / 2) * s * s / G - b
    else:
        raise ValueError(f'unknown case {case!r}')
    return h, b, u0 * np.cos(lat_e) * en


def jet_bumps(g, seed=0, H0=8000.0, u0=20.0, amp=100.0, radius=5.0e5, nbumps=3):
    """One layer in a balanced zonal jet u = u0 cos(lat), h = H0 - (a Om u0 + u0^2/2) sin^2(lat) / g (the Williamson 2
    form), plus `nbumps` Gaussian bumps of `amp` m and e-folding `radius` m at seeded centres."""
    a, Om, G = planet_of(g)
    rng = np.random.default_rng(seed)
    s = np.sin(g.lat)
    h = H0 - (a * Om * u0 + u0 * u0 / 2) * s * s / G
    for _ in range(nbumps):
        lon_c, lat_c = rng.uniform(-math.pi, math.pi), math.asin(rng.uniform(-0.9, 0.9))
        d = great_circle(g, lon_c, lat_c) * (a / R_PLANET)
        h = h + amp * np.exp(-(d / radius) ** 2)
    lat_e, _, en, _ = edge_east_north(g)
    return h[None], np.zeros(g.N), (u0 * np.cos(lat_e) * en)[None]


def gravity_wave(g, H, amp=10.0, radius=1.0e6):
    r = great_circle(g, 0.0, 0.0)
    return (H + amp * np.exp(-(r / radius) ** 2))[None], np.zeros(g.N), np.zeros((1, g.E))


def internal_wave(g, h1, h2, amp=50.0, radius=1.0e6):
    d = amp * np.exp(-(great_circle(g, 0.0, 0.0) / radius) ** 2)
    return np.stack([h1 - d, h2 + d]), np.zeros(g.N), np.zeros((2, g.E))


def outcrop(g, H=(2000.0, 500.0), U=0.5, ridge=900.0, gprime=0.02, floor=1.0):
    """Two Boussinesq layers in balanced solid-body rotation, u = U cos(lat) in both, over a meridional ridge at 0 E
    (crest `ridge` m, e-folding half-width 8 degrees of longitude, fading as cos^2(lat)) that is higher than the lower
    layer is deep: over the crest the lower layer starts at the floor and the upper layer meets the ridge (reference
    preset 'outcrop', cases.py:194-201, 216). Balance: a layer's Montgomery potential in solid-body rotation is
    C_k - (a Om U + U^2 / 2) sin^2(lat); with the same U in both layers the interface between them is level (at H[1]
    above the flat bottom) and the free surface carries the whole potential. Area means: H[0] + H[1] for the column."""
    a, Om, G = planet_of(g)
    wrapped = np.arctan2(np.sin(g.lon), np.cos(g.lon))
    b = ridge * np.exp(-(np.abs(wrapped) / math.radians(8.0)) ** 2) * np.cos(g.lat) ** 2
    w = g.areaCell / np.sum(g.areaCell)
    surf = -(a * Om * U + 0.5 * U * U) * np.sin(g.lat) ** 2 / G
    eta = surf + (H[0] + H[1]) - np.sum(w * surf)
    lower = np.maximum(H[1] - b, floor)
    upper = eta - b - lower
    lat_e, _, en, _ = edge_east_north(g)
    ue = U * np.cos(lat_e) * en
    return np.stack([upper, lower]), b, np.stack([ue, ue])


def internal_wave_speed(gprime, h1, h2, grav=GRAV):
    """The internal mode's exact speed for two Boussinesq layers under a free surface."""
    H = h1 + h2
    return math.sqrt(grav * H / 2 * (1 - math.sqrt(1 - 4 * gprime * h1 * h2 / (grav * H * H))))


def zonal_jet_layers(g, H, ujets, gprime, M=20000):
    """Two layers with zonal winds ujets[k](lat), each in gradient-wind balance integrated from the south pole."""
    a = R_PLANET
    dphi = math.pi / M
    ph = -math.pi / 2 + (np.arange(1, M + 1) - 0.5) * dphi
    Ms = []
    for uj in ujets:
        uu = uj(ph)
        Ms.append(np.concatenate([[0.0], np.cumsum(-a * uu * (2 * OMEGA * np.sin(ph) + np.tan(ph) * uu / a) * dphi)]))
    x = (g.lat + math.pi / 2) / dphi
    k = np.minimum(M - 1, np.floor(x).astype(int))
    m0, m1 = (m[k] + (x - k) * (m[k + 1] - m[k]) for m in Ms)
    A = g.areaCell / np.sum(g.areaCell)
    eta_v, z1_v = m0 / GRAV, (m1 - m0) / gprime
    eta = eta_v + H[0] + H[1] - np.sum(A * eta_v)
    z1 = z1_v + H[1] - np.sum(A * z1_v)
    lat_e, _, en, _ = edge_east_north(g)
    return np.stack([eta - z1, z1]), np.zeros(g.N), np.stack([uj(lat_e) * en for uj in ujets])


def phillips(g, shear_factor, H=(5000.0, 5000.0), Ld=5.0e5, width_deg=7.0, bump=0.0):
    """The balanced Phillips jet: equal layers, the upper a Gaussian jet at 45N of shear_factor times the critical
    shear 2 beta Ld^2, the lower at rest; with bump > 0 the interface is displaced by bump H[1] cos(6 lon) within 10
    degrees of 45N (the seed of the instability; reference preset 'phillips', cases.py:174-191). Returns
    (h, b, u, gprime)."""
    lat0 = math.pi / 4
    f0, beta = 2 * OMEGA * math.sin(lat0), 2 * OMEGA * math.cos(lat0) / R_PLANET
    He = H[0] * H[1] / (H[0] + H[1])
    gprime = (Ld * f0) ** 2 / He
    U0 = shear_factor * 2 * beta * Ld ** 2
    sig = math.radians(width_deg)
    jet = lambda phi: U0 * np.exp(-((np.asarray(phi) - lat0) / sig) ** 2)
    h, b, u = zonal_jet_layers(g, H, (jet, lambda phi: 0.0 * np.asarray(phi)), gprime)
    if bump:
        d = bump * H[1] * np.cos(6 * g.lon) * np.exp(-((g.lat - lat0) / math.radians(10.0)) ** 2)
        h = np.stack([h[0] - d, h[1] + d])
    return h, b, u, gprime


def aqua_air(g):
    """Three air layers at rest over a slab aquaplanet: interfaces at 220 and 600 hPa, theta 415 / 342.5 /
    290 - 30 sin^2(lat) K, surface 288 - 40 (sin^2(lat) - 1/3) K (no colder than 274 K), vapor 10/50/70% of
    saturation at each layer's mid-pressure."""
    s2 = np.sin(g.lat) ** 2
    edges = np.array([0.0, *AQUA_P, MU_REF * GRAV])
    h = np.repeat((np.diff(edges) / GRAV)[:, None], g.N, axis=1)
    theta = np.stack([np.full(g.N, 415.0), np.full(g.N, 342.5), 290.0 - 30.0 * s2])
    Ts = np.maximum(288.0 - 40.0 * (s2 - 1.0 / 3.0), 274.0)
    pm = (0.5 * (edges[:-1] + edges[1:]))[:, None]
    T = theta * (pm / P0) ** KAPPA
    qs = 0.622 * 611.0 * np.exp(17.27 * (T - 273.15) / (T - 35.85)) / pm
    q = np.array([0.1, 0.5, 0.7])[:, None] * qs
    return h, np.zeros(g.N), np.zeros((3, g.E)), dict(theta=theta, Ts=Ts, q=q)


def aqua_layers(g, stars, edges, q_top=0.1, q_bot=0.7):
    """The aquaplanet's start in L = len(stars) + 1 layers (reference: cases.py:257-272): stars the interfaces'
    theta* from the top, edges their starting pressures (Pa); each layer's theta midway between its edges, the top
    layer's 35 K above its edge and the bottom's 290 - 30 sin^2(lat); vapor from q_top to q_bot of saturation at each
    layer's mid-pressure down the column; the surface 288 - 40 (sin^2(lat) - 1/3) K, no colder than 274 K."""
    L = len(stars) + 1
    s2 = np.sin(g.lat) ** 2
    e = np.array([0.0, *edges, MU_REF * GRAV])
    h = np.repeat((np.diff(e) / GRAV)[:, None], g.N, axis=1)
    th = [np.full(g.N, stars[0] + 35.0)] + [np.full(g.N, 0.5 * (stars[k - 1] + stars[k])) for k in range(1, L - 1)]
    theta = np.stack(th + [290.0 - 30.0 * s2])
    Ts = np.maximum(288.0 - 40.0 * (s2 - 1.0 / 3.0), 274.0)
    pm = (0.5 * (e[:-1] + e[1:]))[:, None]
    T = theta * (pm / P0) ** KAPPA
    q = np.linspace(q_top, q_bot, L)[:, None] * (0.622 * 611.0 * np.exp(17.27 * (T - 273.15) / (T - 35.85)) / pm)
    return h, np.zeros(g.N), np.zeros((L, g.E)), dict(theta=theta, Ts=Ts, q=q)


# The aquaplanet in five and seven layers (reference presets 'aqua5', 'aqua7', cases.py:278-279): theta* of the
# interfaces from the top and their starting pressures
AQUA_N = {'aqua5': ([400.0, 350.0, 325.0, 305.0], [1.0e4, 2.5e4, 4.5e4, 8.0e4]),
          'aqua7': ([420.0, 370.0, 345.0, 330.0, 317.0, 305.0], [0.8e4, 1.8e4, 3.0e4, 4.5e4, 6.2e4, 8.0e4])}


def aqua_dry(g):
    """The dry aquaplanet's start (reference: cases.py:244-247): the three air layers of aqua_air at rest, the surface
    at 288 - 40 (sin^2(lat) - 1/3) K, no vapor. Returns h (3, N), b, u (3, E), dict(theta, Ts)."""
    h, b, u, x = aqua_air(g)
    return h, b, u, dict(theta=x['theta'], Ts=288.0 - 40.0 * (np.sin(g.lat) ** 2 - 1.0 / 3.0))


def harmonic_fields(rng, pos, count, lmax=20):
    """`count` random-phase fields of spherical-harmonic degree <= lmax at unit positions pos, each scaled to a
    maximum of 1: sum over l, m of a_lm P_l^m(sin lat) cos(m lon + phase_lm), Gaussian a_lm ~ 1/(l+1), phases uniform.
    Each (l, m) basis function is evaluated once for all the fields."""
    P, lon = _legendre(pos, lmax)
    A = np.zeros((lmax + 1, count, len(pos)))                       # per order m: sum_l a cos(phase) P_lm
    B = np.zeros((lmax + 1, count, len(pos)))                       # per order m: sum_l a sin(phase) P_lm
    for l in range(1, lmax + 1):
        for m in range(0, l + 1):
            a = rng.standard_normal(count) / (l + 1)
            ph = rng.uniform(0, 2 * math.pi, count)
            A[m] += (a * np.cos(ph))[:, None] * P[(l, m)][None, :]
            B[m] += (a * np.sin(ph))[:, None] * P[(l, m)][None, :]
    f = np.zeros((count, len(pos)))
    for m in range(lmax + 1):                                       # cos(m lon + ph) = cos m lon cos ph - sin m lon sin ph
        f += A[m] * np.cos(m * lon)[None, :] - B[m] * np.sin(m * lon)[None, :]
    return f / np.max(np.abs(f), axis=1, keepdims=True)


_LEGENDRE = {}


def _legendre(pos, lmax):
    """The orthonormal associated Legendre functions P_lm(sin lat) (Condon-Shortley phase, so that P_lm(cos theta)
    e^(i m lon) is the unit-norm spherical harmonic Y_l^m) at unit positions, cached per position set, by the standard
    recurrences in l at fixed m: P_mm = -sqrt((2m + 1) / 2m) sqrt(1 - x^2) P_(m-1)(m-1) from P_00 = 1 / sqrt(4 pi),
    P_(m+1)m = sqrt(2m + 3) x P_mm, P_lm = a_lm (x P_(l-1)m - b_lm P_(l-2)m) with a_lm = sqrt((4l^2 - 1) / (l^2 - m^2))
    and b_lm = sqrt(((l - 1)^2 - m^2) / (4 (l - 1)^2 - 1)); x = sin lat."""
    key = (len(pos), lmax, pos[0].tobytes(), pos[-1].tobytes())
    if key not in _LEGENDRE:
        x = np.clip(pos[:, 2], -1.0, 1.0)
        s = np.sqrt((1.0 - x) * (1.0 + x))
        P, pmm = {}, np.full(len(pos), 1.0 / math.sqrt(4 * math.pi))
        for m in range(0, lmax + 1):
            if m > 0:
                pmm = -math.sqrt((2 * m + 1) / (2.0 * m)) * s * pmm
            prev2, prev = None, pmm
            if m >= 1:
                P[(m, m)] = pmm
            if m + 1 <= lmax:
                cur = math.sqrt(2 * m + 3) * x * pmm
                P[(m + 1, m)] = cur
                prev2, prev = pmm, cur
            for l in range(m + 2, lmax + 1):
                a = math.sqrt((4.0 * l * l - 1) / (l * l - m * m))
                b = math.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1) ** 2 - 1))
                cur = a * (x * prev - b * prev2)
                P[(l, m)] = cur
                prev2, prev = prev, cur
        if len(_LEGENDRE) > 8:
            _LEGENDRE.clear()
        _LEGENDRE[key] = (P, np.arctan2(pos[:, 1], pos[:, 0]))
    return _LEGENDRE[key]


def land_mask(g, rng, frac=0.3, smooth=2):
    """A seeded land mask covering `frac` of the cells: a random field smoothed over `smooth` neighbour rings, the
    top `frac` of cells by value."""
    v = rng.standard_normal(g.N)
    nb = np.where(g.nbr >= 0, g.nbr, np.arange(g.N)[:, None])
    for _ in range(smooth):
        v = (v + v[nb].sum(1)) / (1 + nb.shape[1])
    cut = np.quantile(v, 1 - frac)
    return (v > cut).astype(float)


def coupled_stack(g, seed=0, lmax=20, dtheta=2.0, du=5.0, land_frac=0.3):
    """Three air layers (the aquaplanet start) over two ocean layers (a 100 m mixed layer over the abyss, 4 km deep),
    a seeded 30% land mask, and random-phase harmonic perturbations of degree <= lmax: +-dtheta K on every air
    layer's theta and +-du m/s on every air layer's wind (sampled at the edge midpoints along the normals)."""
    rng = np.random.default_rng(seed)
    h, b, u, x = aqua_air(g)
    theta = x['theta'] + dtheta * harmonic_fields(rng, g.pos, 3, lmax)
    lat_e, _, en, nn = edge_east_north(g)
    w = harmonic_fields(rng, g.xm, 6, lmax)
    u = du * (w[:3] * en + w[3:] * nn) / math.sqrt(2.0)
    land = land_mask(g, rng, land_frac)
    return dict(h=h, b=b, u=u, theta=theta, q=x['q'], Ts=x['Ts'], land=land)


def synthetic_sso(g, seed=0, land=None):
    """Seeded sub-grid orography statistics per cell for the mountains' drag (the reference reads Earth's from
    ETOPO1, cases.earth_sso): (mu, sigma, gamma, theta) -- the heights' standard deviation (m; 0 at sea, up to about
    1.5 km over a few ranges), the slope along the principal axis (0.005-0.05), the anisotropy (0.1-1) and the axis's
    angle from east (radians, -pi/2 .. pi/2). land: the land mask (mountains only where > 0.5), else every cell."""
    rng = np.random.default_rng(seed)
    ranges = rng.normal(size=(6, 3))
    ranges /= np.linalg.norm(ranges, axis=1)[:, None]
    # each range a Gaussian ridge of 15 degrees' width around a seeded centre
    d = np.arccos(np.clip(g.pos @ ranges.T, -1.0, 1.0))
    mu = 1500.0 * np.max(np.exp(-(d / math.radians(15.0)) ** 2), axis=1)
    if land is not None:
        mu = np.where(np.asarray(land) > 0.5, mu, 0.0)
    sigma = np.where(mu > 0, 0.005 + 0.045 * rng.random(g.N), 0.0)
    gamma = 0.1 + 0.9 * rng.random(g.N)
    theta = (rng.random(g.N) - 0.5) * math.pi
    return np.stack([mu, sigma, gamma, theta])


def held_suarez(g, seed=1):
    """Held & Suarez (1994) in two air layers, started as their protocol starts (reference: cases.py:657-670): at
    rest and horizontally uniform, each layer at the area-mean equilibrium theta of its half's mass-mean pressure
    (250 and 750 hPa), the lower capped 1 K below theta*, with 0.1 K of seeded noise. Returns h (2, N), b, u (2, E),
    theta (2, N)."""
    from .atmosphere import held_suarez_theta
    h = np.full((2, g.N), MU_REF / 2)
    A = g.areaCell / np.sum(g.areaCell)
    lat = np.arcsin(np.clip(g.pos[:, 2], -1, 1))
    rng = np.random.default_rng(seed)
    means = [float(np.sum(A * held_suarez_theta(p, lat))) for p in (0.25e5, 0.75e5)]
    means[1] = min(means[1], held_suarez_theta_star(g) - 1.0)
    theta = np.stack([np.full(g.N, m) for m in means]) + 0.1 * rng.standard_normal((2, g.N))
    return h, np.zeros(g.N), np.zeros((2, g.E)), theta


def held_suarez_theta_star(g):
    """The Held-Suarez interface's theta*: the area mean of the equilibrium theta at 500 hPa (reference:
    cases.py:673-675)."""
    from .atmosphere import held_suarez_theta
    lat = np.arcsin(np.clip(g.pos[:, 2], -1, 1))
    return float(np.sum(g.areaCell * held_suarez_theta(0.5e5, lat)) / np.sum(g.areaCell))
