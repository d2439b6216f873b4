"""The air column's parameters and the radiation pass of the air physics (reference: src/insilica/atmosphere.py).

`PAR_ORDER` and `DEFAULTS` keep the reference's names, order and default values (atmosphere.py:26-43): the kernels
read the parameters by position, so the array `par_array` builds is the reference's `AirPhysics.par` value for value
(132 parameters, then the interfaces' theta* from the top, padded to LMAX). `AirPhysics.radiate` is the reference's
k-table radiation call (atmosphere.py:117-123): the longwave's set-up and column pass, then the shortwave.
"""
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
    ent_lcl=0.0, sub_mix=0.0)
assert len(PAR_ORDER) == 132 and set(PAR_ORDER) <= set(DEFAULTS)


def par_array(layers, params=None):
    """The reference's AirPhysics.par (atmosphere.py:84-90) as a host float64 array: PAR_ORDER's values, then the
    interfaces' theta* from the top (`theta_stars`, or theta_star2/theta_star), zero-padded to LMAX."""
    p = dict(DEFAULTS, **(params or {}))
    stars = list(p.get('theta_stars') or ([p['theta_star2'], p['theta_star']] if layers > 2 else [p['theta_star']]))
    if len(stars) != layers - 1:
        raise ValueError('one theta* per interface')
    p['theta_star'] = stars[-1]
    return np.array([float(p[k]) for k in PAR_ORDER] + [float(x) for x in stars] + [0.0] * (LMAX - len(stars)))


class AirPhysics:
    """The radiation of the air physics (reference: AirPhysics with longwave='ktables', atmosphere.py:56-123):
    `lw`, `sw` (insilica.radiation.Longwave/Shortwave), `cloudo` ((4L + 3), N: per layer cover, total cover, the
    clouds' effect at the top, per layer longwave and shortwave in-cloud optical depths and liquid share), `pos`
    (unit cell positions), `swalb` (the surface albedo the shortwave sees)."""

    def __init__(self, N, layers, params=None, mean_sun=False, lw_tables=None, sw_tables=None, pos=None,
                 precision='mixed', device=None):
        from .radiation import Longwave, Shortwave
        self.N, self.L = N, layers
        self.p = dict(DEFAULTS, **(params or {}))
        self.par = par_array(layers, params)
        self.mode = 1 if mean_sun else 0
        dev = device or torch.device('cuda', torch.cuda.current_device())
        self.pos = None if pos is None else torch.as_tensor(np.ascontiguousarray(pos, dtype=np.float64), device=dev)
        self.cloudo = torch.zeros(4 * layers + 3, N, dtype=torch.float64, device=dev)
        self.lw = Longwave(N, layers, self.par, tables=lw_tables, precision=precision, device=dev,
                           **{k: self.p[k] for k in ('co2', 'n2o', 'ch4', 'tmode') if k in self.p})
        ice_re = self.p['re_ice'] if self.p.get('ice_sw', 0) > 0 else None
        self.sw = Shortwave(N, layers, self.par, self.lw, tables=sw_tables, ice_re=ice_re, precision=precision,
                            device=dev, **({'inhom': self.p['cloud_inhom']} if 'cloud_inhom' in self.p else {}))
        self.swalb = torch.full((N,), float(self.p['alb']), dtype=torch.float64, device=dev)

    def radiate(self, mu, theta, Ts, t, interval, q=None, cw=None, gx=None):
        """The k-tables' longwave and shortwave from the present state, the sun over the coming `interval` (s) from
        the device clock t (1 double); the clouds' cloudo used when cw (the cloud water) is given."""
        co = self.cloudo if cw is not None else None
        self.lw(mu, theta, q, co, Ts, gx)
        self.sw(self.mode, t, interval, self.pos, self.swalb, co)
