"""The air physics' parameters and drivers (reference: src/insilica/atmosphere.py).

`PAR_ORDER` and `DEFAULTS` keep the reference's names, order and default values (atmosphere.py:26-43): the kernels
read the parameters by position, so the array `par_array` builds is the reference's `AirPhysics.par` value for value
(132 parameters, then the interfaces' theta* from the top, padded to LMAX). `AirPhysics` drives the column kernels
through the C ABI: the surface layer and the wind speed, the Held-Suarez column with its interface crossing, the
radiative air column (dry, or moist over the sea), the k-table radiation (the longwave's set-up and column pass,
then the shortwave) and the device clock.
"""
import ctypes
import math

import numpy as np
import torch

from .constants import CP_DRY, GRAV, KAPPA, P0

LMAX = 8
SIGMA = 5.670e-8
DAY_S, YEAR_S = 86400.0, 365 * 86400.0
TILT = math.radians(23.44)

PAR_ORDER = (
    'S0', 'asw', 'tau0', 'mu_ref', 'eps_s', 'h_sens', 'u0', 'alb', 'c_slab', 'mu_bl', 'theta_star', 'floor_', 'sigma',
    'grav', 'cp', 'kappa', 'p0', 'tilt', 'year_s', 'day_s', 'tau_lin', 'theta_star2',
    'moist', 'Lv', 'Lf', 'Tf', 'k_evap', 'tau_conv', 'alb_ice', 'rho_ice', 'ice_cover', 'k_ice', 'Rv', 'sat_layer',
    'sw_tau', 'alb_dry', 'alb_wet', 'soil_beta', 'k_inf', 'pond_cover', 'alb_snow', 'alb_ice_melt', 'alb_snow_melt',
    'alb_glacier', 'snow_cover', 'alb_ramp', 'land_exch', 'plume_deep', 'lw_w0', 'c_dry', 'c_wet', 'sat_prof',
    'land_rs', 'sat_base', 'bl_bot', 'cloud', 'rh_crit', 're_liq', 're_ice', 'cloud_g', 'k_lw_liq', 'k_lw_ice',
    'c_auto', 'q_auto', 'v_ice', 'l_det', 'tau_erode', 'k_revap', 'mu_cloud', 'theta_rise', 'bl_mix', 'mu_night',
    'bl_ent', 'bl_cap', 'bulk_c', 'bulk_gust', 'bulk_z', 'bulk_w', 'plume_ent', 'l_ice', 'storm_area', 'graupel_v',
    'ice_small', 'q_shape', 'edge_heat', 'freeze_dry', 'edge_vap', 'conv_revap', 'plume_root', 'bl_q', 'sfc',
    'z0_land', 'downdraft', 'ent_cloud', 'c_deep', 'rs_max', 'rs_sun', 'q_sea', 'bl_reach', 'dry_edge', 'bl_lcl',
    'edge_k', 'ice_air', 'revap', 'revap_rhs', 'revap_rhl', 'revap_area', 'root_q', 'ent_det', 'ent_ifs', 'shallow',
    'cu_dth', 'cu_dp', 'buoy_v', 'cu_sub', 'bare_fx', 'det_all', 'cl_float', 'ml_clear', 'cu_cloud', 'firn', 'k_firn',
    'rho_firn', 'c_ice', 'lw_refl', 'ent_loc', 'ecc', 'lam_peri', 'liq_cold', 'cw_qs', 'ent_lcl', 'sub_mix')

DEFAULTS = dict(
    S0=1361.0, asw=0.2, tau0=0.9, mu_ref=P0 / GRAV, eps_s=0.95, h_sens=10.0, u0=10.0, alb=0.06,
    c_slab=1000.0 * 4186.0 * 100.0, mu_bl=1.0e4 / GRAV, theta_star=300.0, floor_=10.0, sigma=SIGMA, grav=GRAV,
    cp=CP_DRY, kappa=KAPPA, p0=P0, tilt=TILT, year_s=YEAR_S, day_s=DAY_S, tau_lin=0.5, theta_star2=380.0,
    moist=0.0, Lv=2.5e6, Lf=3.34e5, Tf=273.15, k_evap=4.0e-3, tau_conv=7200.0, alb_ice=0.7, rho_ice=917.0,
    ice_cover=0.05, k_ice=2.2, Rv=461.5, sat_layer=1.0, sw_tau=2.0, alb_dry=0.30, alb_wet=0.18, soil_beta=0.75,
    k_inf=5.0 / 3600.0, pond_cover=50.0, alb_snow=0.8, alb_ice_melt=0.50, alb_snow_melt=0.55, alb_glacier=0.35,
    snow_cover=30.0, alb_ramp=10.0, land_exch=1.0, plume_deep=0.0, lw_w0=0.0, c_dry=1.3e6, c_wet=0.0, sat_prof=0.0,
    land_rs=0.0, sat_base=0.0, bl_bot=0.0, cloud=0.0, rh_crit=0.8, re_liq=10e-6, re_ice=30e-6, cloud_g=0.85,
    k_lw_liq=130.0, k_lw_ice=65.0, c_auto=1e-3, q_auto=0.3e-3, v_ice=0.5, l_det=0.5e-3, tau_erode=1e5,
    k_revap=1e-5, mu_cloud=500.0, theta_rise=0.0, bl_mix=0.0, mu_night=250.0, bl_ent=0.2, bl_cap=0.0, bulk_c=0.0,
    bulk_gust=1.0, bulk_z=0.0, bulk_w=0.0, plume_ent=0.0, l_ice=0.1e-3, storm_area=1.0e8, graupel_v=5.0,
    ice_small=0.1, q_shape=0.0, edge_heat=0.0, freeze_dry=0.0, edge_vap=0.0, conv_revap=0.0, plume_root=0.0,
    bl_q=0.0, sfc=0.0, z0_land=0.1, downdraft=0.0, ent_cloud=0.0, c_deep=0.0, rs_max=0.0, rs_sun=180.0, q_sea=0.0,
    bl_reach=0.0, dry_edge=0.0, bl_lcl=0.0, edge_k=1.0, ice_air=0.0, revap=0.0, revap_rhs=0.9, revap_rhl=0.7,
    revap_area=0.05, root_q=0.0, ent_det=0.0, ent_ifs=0.0, shallow=0.0, cu_dth=0.0, cu_dp=0.0, buoy_v=0.0,
    cu_sub=0.0, bare_fx=0.0, det_all=0.0, cl_float=0.0, ml_clear=0.0, cu_cloud=0.0, firn=0.0, k_firn=0.34,
    rho_firn=450.0, c_ice=2090.0, lw_refl=0.0, ent_loc=0.0, ecc=0.0, lam_peri=0.0, liq_cold=20.0, cw_qs=0.0,
    ent_lcl=0.0, sub_mix=0.0, cover=0.0, sfc_wind=0.0,
    # the drags (insilica.orodrag; not in PAR_ORDER: the column does not read them)
    oro=0.0, oro_g=0.3, oro_cd=1.0, oro_hnc=0.5, oro_ric=0.25,
    gw=0.0, gw_bt=0.004, gw_cw=40.0, gw_lambda=3.0e5, gw_psrc=31500.0, gw_cmax=99.6, gw_dc=1.2, gw_conv=0.0)
assert len(PAR_ORDER) == 132 and set(PAR_ORDER) <= set(DEFAULTS)


def annual_mean_insolation(lat, S0=1361.0, tilt=TILT, days=365):
    """The annual mean of the daily-mean insolation (W/m^2) at latitudes lat (radians) on a circular orbit, averaged
    over `days` days at mid-day declinations (reference: atmosphere.py:48-57): S0 / pi (h0 sin(lat) sin(dec) +
    cos(lat) cos(dec) sin(h0)), h0 the sunset hour angle, arccos(-tan(lat) tan(dec)) clipped to [0, pi]."""
    lat = np.asarray(lat, dtype=float)
    dec = np.arcsin(math.sin(tilt) * np.sin(2 * math.pi * (np.arange(days) + 0.5) / days))[:, None]
    h0 = np.arccos(np.clip(-np.tan(lat)[None, :] * np.tan(dec), -1, 1))
    day = S0 / math.pi * (h0 * np.sin(lat)[None, :] * np.sin(dec) + np.cos(lat)[None, :] * np.cos(dec) * np.sin(h0))
    return day.mean(0).reshape(lat.shape)


def par_array(layers, params=None):
    """The reference's AirPhysics.par (atmosphere.py:84-90) as a host float64 array: PAR_ORDER's values, then the
    interfaces' theta* from the top (`theta_stars`, or theta_star2/theta_star), zero-padded to LMAX."""
    p = dict(DEFAULTS, **(params or {}))
    stars = list(p.get('theta_stars') or ([p['theta_star2'], p['theta_star']] if layers > 2 else [p['theta_star']]))
    if len(stars) != layers - 1:
        raise ValueError('one theta* per interface')
    p['theta_star'] = stars[-1]
    return np.array([float(p[k]) for k in PAR_ORDER] + [float(x) for x in stars] + [0.0] * (LMAX - len(stars)))


def perot_operators(g):
    """Perot's reconstruction of a cell's wind (east, north) from its edges' normal winds: (N, 6, 2) coefficients per
    edge slot (reference: orodrag.operators, orodrag.py:20-28), u_cell = (1/A) sum_e sign_e le_e u_e (x_e - x_c)."""
    from .constants import R_PLANET
    pos = np.asarray(g.pos, float)
    lon, lat = np.arctan2(pos[:, 1], pos[:, 0]), np.arcsin(np.clip(pos[:, 2], -1, 1))
    east = np.stack([-np.sin(lon), np.cos(lon), np.zeros_like(lon)], 1)
    north = np.stack([-np.sin(lat) * np.cos(lon), -np.sin(lat) * np.sin(lon), np.cos(lat)], 1)
    ce = np.where(g.cellEdge >= 0, g.cellEdge, 0)
    c = np.where(g.cellEdge >= 0, g.cellSign * g.le[ce] * getattr(g, 'radius', R_PLANET), 0.0) / g.areaCell[:, None]
    d = g.xm[ce] - pos[:, None, :]
    return np.stack([c * np.sum(d * east[:, None, :], 2), c * np.sum(d * north[:, None, :], 2)], 2)


def held_suarez_theta(p, lat):
    """Held & Suarez's equilibrium potential temperature at pressure p (Pa) and latitude lat (radians) (reference:
    atmosphere.py:161-165)."""
    s2 = np.sin(lat) ** 2
    T = np.maximum(200.0, (315.0 - 60.0 * s2 - 10.0 * np.log(p / P0) * (1 - s2)) * (p / P0) ** KAPPA)
    return T * (P0 / p) ** KAPPA


class AirPhysics:
    """The air physics on a DeviceGrid (reference: atmosphere.py:56-158): `surface` (the surface layer's exchange
    coefficients), `column` (the Held-Suarez forcing with the interfaces' crossing), `radiate` (the k-tables'
    longwave and shortwave), and the device clock. Parameters in the reference's PAR_ORDER array (`par`, host).

    forcing: 'held_suarez' steps the Held & Suarez column (air.cu hs_column); 'radiation' the reference's air column
    (air.cu air_column), dry, or moist over the sea and land, with or without clouds. mean_sun: the
    annual-mean insolation instead of the diurnal and seasonal sun. longwave='ktables' builds `lw`, `sw` from
    lw_tables, sw_tables (npz paths or dicts). `cloudo` ((4L + 3), N): per layer cover, total cover, the clouds'
    effect at the top, per layer longwave and shortwave in-cloud optical depths and liquid share."""

    def __init__(self, dg, layers, params=None, mean_sun=False, forcing='radiation', longwave='gray', lw_tables=None,
                 sw_tables=None, air=None):
        self.dg, self.L, self.forcing = dg, layers, forcing
        self.p = dict(DEFAULTS, **(params or {}))
        self.par = par_array(layers, params)
        self._par = self.par.ctypes.data_as(ctypes.c_void_p)
        self.mode = 1 if mean_sun else 0
        from .radiation import _air
        self.air = _air(air)
        N, dev = dg.N, dg.dev
        z = lambda *s: torch.zeros(*s, dtype=torch.float64, device=dev)
        self.pos = torch.as_tensor(np.ascontiguousarray(dg.host.pos, dtype=np.float64), device=dev)
        self.lat = torch.as_tensor(np.arcsin(np.clip(np.asarray(dg.host.pos)[:, 2], -1, 1)), device=dev)
        self.uabs, self.wx, self.qnet, self.stored = z(N), z(max(layers - 1, 1), N), z(N), z(N)
        self.olr, self.sw_abs, self.rain, self.evap, self.conv, self.tsurf = (z(N) for _ in range(6))
        self.rainp, self.flash = z(4, N), z(N)
        self.dgo = None                     # the diagnostics rows ((5 L - 2, N), air_column), set by a caller
        self.insol = None
        if mean_sun:
            lat = np.arcsin(np.clip(np.asarray(dg.host.pos)[:, 2], -1, 1))
            self.insol = torch.as_tensor(annual_mean_insolation(lat, self.p['S0'], self.p['tilt']), device=dev)
        self.moist = bool(self.p['moist'])
        self.hs = np.array([1.0 / (40 * DAY_S), 1.0 / (4 * DAY_S)])          # k_a, k_s (1/s)
        self.perot = None
        if self.p.get('sfc_wind', 0.0) > 0:
            self.perot = torch.as_tensor(np.ascontiguousarray(perot_operators(dg.host)).reshape(-1), device=dev)
        self.sfc = self.p['sfc'] > 0
        self.cd, self.ch = (z(N), z(N)) if self.sfc else (None, None)
        self.cloudo = z(4 * layers + 3, N)
        self.cover = None                                  # the land cover's rows (8, N), when a world gives them
        self.lw = self.sw = None
        if longwave == 'ktables':
            from .radiation import Longwave, Shortwave
            self.lw = Longwave(dg, layers, self.par, tables=lw_tables, air=air, precision=dg.precision, device=dev,
                               **{k: self.p[k] for k in ('co2', 'n2o', 'ch4', 'tmode') if k in self.p})
            ice_re = self.p['re_ice'] if self.p.get('ice_sw', 0) > 0 else None
            self.sw = Shortwave(dg, layers, self.par, self.lw, tables=sw_tables, ice_re=ice_re, air=air,
                                precision=dg.precision, device=dev,
                                **({'inhom': self.p['cloud_inhom']} if 'cloud_inhom' in self.p else {}))
            self.swalb = torch.full((N,), float(self.p['alb']), dtype=torch.float64, device=dev)
        self._mixed = 1 if dg.precision == 'mixed' else 0

    def _speed(self, u):
        """The surface's wind speed at cells from the bottom layer's edge winds u (L, E), into self.uabs (reference:
        atmosphere.py:109-115)."""
        from . import _hip
        ub = u[self.L - 1]
        if self.perot is not None:
            _hip.call('isw_cell_speed_perot', self.dg.ref, _hip.ptr(ub), _hip.ptr(self.perot), _hip.ptr(self.uabs),
                      _hip.stream())
        else:
            _hip.call('isw_cell_speed', self.dg.ref, _hip.ptr(ub), _hip.ptr(self.uabs), _hip.stream())

    def radiate(self, mu, theta, Ts, t, interval, q=None, cw=None, gx=None):
        """The k-tables' longwave and shortwave from the present state, the sun over the coming `interval` (s) from
        the device clock t (1 double); the clouds' cloudo used when cw (the cloud water) is given."""
        if self.lw is not None:
            co = self.cloudo if cw is not None else None
            self.lw(mu, theta, q, co, Ts, gx)
            self.sw(self.mode, t, interval, self.pos, self.swalb, co)

    def surface(self, mu, theta, Ts, u, q=None, ice=None, land=None, lc=None):
        """The surface layer's rho C_D U and rho C_H U into self.cd, self.ch (reference: atmosphere.py:125-132)."""
        from . import _hip
        if not self.sfc:
            raise ValueError("surface(): the surface layer is off (params sfc = 0)")
        self._speed(u)
        _hip.call('isw_surface_layer', self.dg.N, self.L, self._par, ctypes.byref(self.air), _hip.ptr(self.uabs),
                  _hip.ptr(mu), _hip.ptr(theta), _hip.ptr(q), _hip.ptr(Ts), _hip.ptr(ice), _hip.ptr(land),
                  _hip.ptr(self.cd), _hip.ptr(self.ch), _hip.ptr(lc if lc is not None else self.cover), _hip.stream())

    def column(self, mu, theta, Ts, u, t, dt, q=None, ice=None, cap=None, qocean=None, qsw=None, land=None,
               soilcap=None, soil=None, pond=None, snow=None, cw=None, elev=None, firn=None, gx=None):
        """One column step (reference: atmosphere.py:134-158), then the device clock t advances by dt. With forcing
        'held_suarez': the relaxation and the crossing into theta (in place) and self.wx (the mass to move up, for
        Layers.exchange), self.qnet, self.stored. With forcing 'radiation': the sun, the gray or k-table radiation
        (radf/swf of the last radiate()), the slab surface at Ts (cap its heat capacity per cell when given; qocean,
        qsw the ocean's forcing), the surface's sensible heat (the surface layer's ch with sfc > 0) and the crossing;
        for moist air (moist > 0; q (L, N) and ice (N) updated in place) evaporation, sea ice, condensation and the
        parcel plume; with clouds (cloud > 0 and the cloud water cw (L, N), updated in place) the cloud statistics,
        autoconversion, ice fall and rain evaporation, the clouds' reflection and emission, and their cover and
        optical depths into self.cloudo for the next radiate(); over land (the mask land (N), the bucket's capacity
        soilcap; soil, pond and snow updated in place; self.cover the land cover when set) the ground's water,
        snow and heat. Into theta, Ts, self.wx, self.olr, self.sw_abs, self.stored, self.rain, self.evap,
        self.conv, self.tsurf, self.rainp, self.flash. gx (3, L, N; the gas tracers' mole fractions per layer, the
        reference's -DGASX column, world.py:124-125), when given, is carried by every move of the column's air
        (updated in place)."""
        from . import _hip
        if self.forcing == 'held_suarez':
            _hip.call('isw_hs_column', self.dg.N, self.L, self._par, self.hs.ctypes.data_as(ctypes.c_void_p),
                      _hip.ptr(self.lat), _hip.ptr(mu), _hip.ptr(theta), float(dt), _hip.ptr(self.wx),
                      _hip.ptr(self.qnet), _hip.ptr(self.stored), _hip.stream())
            self.advance_time(t, dt)
            return
        if self.moist and (q is None or ice is None):
            raise ValueError('the moist column needs the vapor q and the sea ice')
        if gx is not None and (not gx.is_contiguous() or tuple(gx.shape) != (3, self.L, self.dg.N)):
            raise ValueError('gx: a contiguous (3, L, N) array of the gas tracers per layer')
        if self.dgo is not None and (not self.dgo.is_contiguous() or
                                     tuple(self.dgo.shape) != (5 * self.L - 2, self.dg.N)):
            raise ValueError('dgo: a contiguous (5 L - 2, N) array for the diagnostics rows')
        self._speed(u)
        P = _hip.ptr
        lw, sw = self.lw, self.sw
        _hip.call('isw_air_column', self.dg.N, self.L, self._par, ctypes.byref(self.air), self.mode, P(t),
                  P(self.insol), P(self.pos), P(self.uabs), P(mu), P(theta), P(q) if self.moist else None, P(Ts),
                  P(ice) if self.moist else None, float(dt),
                  P(self.wx), P(self.olr), P(self.sw_abs), P(self.stored), P(self.rain), P(self.evap), P(self.conv),
                  P(self.tsurf), P(self.rainp), P(self.dgo), P(cap), P(qocean), P(qsw), P(land), P(soilcap), P(soil), P(pond),
                  P(snow), P(cw), P(self.cloudo) if cw is not None else None, P(elev), P(self.flash),
                  P(lw.radf) if lw is not None else None,
                  P(sw.swf) if sw is not None else None, P(self.swalb) if sw is not None else None,
                  P(self.ch) if self.sfc else None, P(self.cover), P(firn), P(gx), self._mixed, _hip.stream())
        self.advance_time(t, dt)

    @staticmethod
    def advance_time(t, dt):
        from . import _hip
        _hip.call('isw_advance_time', _hip.ptr(t), float(dt), _hip.stream())


def gas_sources(src, h, dt, x, layers):
    """World._tick's gas sources (reference: world.py:343-346): src (G, N) into the bottom layer of x (G*L, N)."""
    from . import _hip
    G, N = src.shape
    _hip.call('isw_gas_src', N, layers, G, _hip.ptr(src), _hip.ptr(h), float(dt), _hip.ptr(x), _hip.stream())


