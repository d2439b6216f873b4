"""Longwave and shortwave radiation by k-tables (reference: src/insilica/radiation.py; kernels/air.cu:2111-2511).

`Longwave(N, layers, par, ...)` and `Shortwave(N, layers, par, lw, ...)` keep the reference's constructor arguments
and call signatures (radiation.py:26-58, 88-115), with the number of columns N (or a DeviceGrid) in place of the
reference's dg. Their tables come from an npz in the reference's format (`lw_ktables.npz`: logk (6, NP, NT, NG),
planck (NTP, NG), P, T, T_pl, o3; `sw_ktables.npz`: logk (7, NP, NT, NG), P, T, sun, ray, haze, ng, haze_par) or from
a dict of the same arrays (insilica.ktables.synthetic_lw / synthetic_sw). State lives in torch tensors on the HIP
device: `lw.sub` (4L*15, N) float32, `lw.radf` (3(L+1)+3+6L, N) and `sw.swf` (3L+2, N) float64. Each call is one
C-ABI entry point on the current stream (isw_lw_setup, isw_lw_column, isw_sw_column); no CPU path.
"""
import ctypes
from pathlib import Path

import numpy as np
import torch

from . import _hip

HAZE_SCALE = 1500.0                # m, the haze's scale height (reference: radiation.py:24)
NGMAX = 256                        # g-points per table: one block per column, one thread per g-point
# The clouds' single-scattering albedo per shortwave band and the inhomogeneity factor on their optical depth
# (reference: radiation.py:61-67).
CLOUD_SSA = (0.60, 0.75, 0.98, 0.994, 0.9995, 0.9998, 0.99999, 0.99999, 0.99999, 0.99999)
CLOUD_INHOM = 0.7
# Ice clouds' optics by Ebert & Curry (1992) in five intervals chosen by each band's upper edge (reference:
# radiation.py:68-85).
EC92_C = (1.000e-05, 1.100e-04, 1.240e-02, 3.779e-02, 4.666e-01)
EC92_D = (0.0, 1.405e-05, 6.867e-04, 1.284e-03, 2.050e-05)
EC92_E = (7.661e-01, 7.730e-01, 7.865e-01, 8.172e-01, 9.595e-01)
EC92_F = (5.851e-04, 5.665e-04, 7.204e-04, 7.463e-04, 1.076e-04)
BAND_TOP = (2600, 4000, 5250, 7700, 9350, 12850, 16000, 29000, 38000, 50000)

# The air's constants the column kernels read (isw_air): Earth air's (reference: kernels/air.cu:75-116 defaults).
EARTH_AIR = dict(m_air=28.9647, m_air_kg=28.9647e-3, o2=0.2095, ray=1.0, broad=1.0, o3=1.0, eps=0.622, rd=287.0,
                 vf=0.61)


def ice_optics(re_um, band_top=BAND_TOP):
    """Per band (ssa, g) of ice clouds of effective radius re_um (Ebert & Curry 1992)."""
    out = []
    for top in band_top:
        j = 0 if top > 14300 else 1 if top > 7700 else 2 if top > 5300 else 3 if top > 4000 else 4
        out.append((1 - EC92_C[j] - EC92_D[j] * re_um, min(EC92_E[j] + EC92_F[j] * re_um, 1 - 1e-6)))
    return out


def _load(tables):
    if isinstance(tables, dict):
        return tables
    if tables is None:
        raise ValueError('radiation tables: pass an npz path in the reference format or a dict '
                         '(insilica.ktables.synthetic_lw / synthetic_sw)')
    with np.load(Path(tables)) as z:
        return {k: z[k] for k in z.files}


def _dev(device):
    return device or torch.device('cuda', torch.cuda.current_device())


def _f32(a, dev):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev)


def _air(air):
    a = _hip.IswAir()
    for k, v in dict(EARTH_AIR, **(air or {})).items():
        if k != 'cia':
            setattr(a, k, float(v))
    a.cia = 1 if (air or {}).get('cia') else 0
    return a


def _prec(precision):
    if precision not in ('mixed', 'double'):
        raise ValueError(precision)
    return _hip.ISW_MIXED if precision == 'mixed' else _hip.ISW_FP64


def _n(dg):
    return int(dg if isinstance(dg, (int, np.integer)) else dg.N)


class Longwave:
    def __init__(self, dg, layers, par, co2=400e-6, n2o=0.32e-6, ch4=1.8e-6, tmode=2, tables=None, air=None,
                 precision='mixed', device=None):
        """dg: a DeviceGrid or the number of columns; par: the column parameters (atmosphere.par_array, host);
        tables: the k-tables (npz path or dict); air: another air's constants (isw_air fields; cia needs the tables'
        'cia' array, (NT, NG), weights x 1e40)."""
        t = _load(tables)
        dev = _dev(device)
        self.N, self.L, self.tmode = _n(dg), int(layers), int(tmode)
        if not 1 <= self.L <= _hip.LMAX:
            raise ValueError(f'1 <= layers <= {_hip.LMAX}')
        self.par = np.ascontiguousarray(par, dtype=np.float64)
        self.prec = _prec(precision)
        self.logk, self.planck = _f32(t['logk'], dev), _f32(t['planck'], dev)
        self.o3 = _f32(t['o3'], dev)
        self.cia = _f32(t['cia'], dev) if (air or {}).get('cia') else None
        P, T, Tp = np.asarray(t['P'], float), np.asarray(t['T'], float), np.asarray(t['T_pl'], float)
        self.NC, self.NP, self.NT, self.NG = self.logk.shape
        if self.NC != 6 or self.NG > NGMAX:
            raise ValueError(f'longwave tables: 6 components and at most {NGMAX} g-points')
        self.lp0, self.dlp = float(np.log(P[0])), float((np.log(P[0]) - np.log(P[-1])) / (len(P) - 1))
        self.t0, self.dtt, self.tp0 = float(T[0]), float(T[1] - T[0]), float(Tp[0])
        self.gases = (co2, n2o, ch4)
        self.air = _air(air)
        self.radf = torch.zeros(3 * (self.L + 1) + 3 + 6 * self.L, self.N, dtype=torch.float64, device=dev)
        self.sub = torch.zeros(4 * self.L * 15, self.N, dtype=torch.float32, device=dev)
        tb = _hip.IswLwTables()
        tb.NC, tb.NP, tb.NT, tb.NG, tb.NTP, tb.NO3 = self.NC, self.NP, self.NT, self.NG, self.planck.shape[0], len(self.o3)
        tb.lp0, tb.dlp, tb.t0, tb.dtt, tb.tp0 = self.lp0, self.dlp, self.t0, self.dtt, self.tp0
        tb.o3lp0, tb.o3dlp = self.lp0, self.dlp          # the ozone profile on the tables' pressure axis
        for k in ('logk', 'planck', 'o3', 'cia'):
            x = getattr(self, k)
            setattr(tb, k, ctypes.cast(x.data_ptr(), _hip._f32p) if x is not None else None)
        self.tables = tb
        self._par = self.par.ctypes.data_as(ctypes.c_void_p)

    def setup(self, mu, theta, q, cloudo, gx=None, hs=HAZE_SCALE):
        """lw_setup: the sub-levels into self.sub."""
        _hip.call('isw_lw_setup', self.N, self.L, self._par, ctypes.byref(self.tables), ctypes.byref(self.air),
                  *self.gases, _hip.ptr(gx), _hip.ptr(mu), _hip.ptr(theta), _hip.ptr(q), _hip.ptr(cloudo), hs,
                  self.tmode, _hip.ptr(self.sub), self.prec, _hip.stream())

    def column(self, Ts):
        """lw_column: the fluxes into self.radf from self.sub and the surface temperature Ts (N)."""
        _hip.call('isw_lw_column', self.N, self.L, self._par, ctypes.byref(self.tables), _hip.ptr(self.sub),
                  _hip.ptr(Ts), _hip.ptr(self.radf), self.prec, _hip.stream())

    def __call__(self, mu, theta, q, cloudo, Ts, gx=None):
        """cloudo: the clouds' cover and optical depths ((4L + 3), N) or None; gx: the carried CO2, N2O and CH4
        (3, L, N) or None for the planet's own."""
        self.setup(mu, theta, q, cloudo, gx)
        self.column(Ts)


class Shortwave:
    def __init__(self, dg, layers, par, lw, inhom=CLOUD_INHOM, ice_re=None, tables=None, air=None,
                 cloud_ssa=None, band_top=None, precision='mixed', device=None):
        """lw: the Longwave whose sub-levels this reads; ice_re: ice clouds' effective radius (m) for their own optics
        (None: the drops'); cloud_ssa, band_top: per band, the drops' albedo and the band's upper edge (cm-1)
        (default: the tables' own, else the reference's ten bands)."""
        t = _load(tables)
        dev = _dev(device)
        self.N, self.L, self.lwm, self.inhom = _n(dg), int(layers), lw, float(inhom)
        self.par = np.ascontiguousarray(par, dtype=np.float64)
        self.prec = _prec(precision)
        self.logk = _f32(t['logk'], dev)
        self.NC, self.NP, self.NT, self.NG = self.logk.shape
        if self.NC != 7 or self.NG > NGMAX:
            raise ValueError(f'shortwave tables: 7 components and at most {NGMAX} g-points')
        sun = np.asarray(t['sun'], float)
        self.sun, self.ray, self.hzf = _f32(sun / sun.sum(), dev), _f32(t['ray'], dev), _f32(t['haze'], dev)
        ng = np.asarray(t['ng'], int)
        if cloud_ssa is None:                            # the tables' own bands, else the reference's ten
            cloud_ssa = t['cloud_ssa'] if 'cloud_ssa' in t else CLOUD_SSA
        if band_top is None:
            band_top = t['band_top'] if 'band_top' in t else BAND_TOP
        self.cssa = _f32(np.repeat(cloud_ssa, ng), dev)
        self.cssa_i = self.g_i = None
        if ice_re is not None:
            io = ice_optics(ice_re * 1e6, band_top)
            self.cssa_i = _f32(np.repeat([x[0] for x in io], ng), dev)
            self.g_i = _f32(np.repeat([x[1] for x in io], ng), dev)
        self.cia = _f32(t['cia'], dev) if (air or {}).get('cia') else None
        self.haze = np.asarray(t['haze_par'], float)         # tau550, ssa, g, scale height
        self.yscale = float(10.0 / (t['T'][1] - t['T'][0]))   # lw_setup's temperature positions are in 10 K steps
        self.air = _air(air)
        self.swf = torch.zeros(3 * self.L + 2, self.N, dtype=torch.float64, device=dev)
        tb = _hip.IswSwTables()
        tb.NC, tb.NP, tb.NT, tb.NG = self.NC, self.NP, self.NT, self.NG
        tb.yscale, tb.tau550, tb.hssa, tb.hg, tb.inhom = self.yscale, *map(float, self.haze[:3]), self.inhom
        for k in ('logk', 'sun', 'ray', 'hzf', 'cssa', 'cssa_i', 'g_i', 'cia'):
            x = getattr(self, k)
            setattr(tb, k, ctypes.cast(x.data_ptr(), _hip._f32p) if x is not None else None)
        self.tables = tb
        self._par = self.par.ctypes.data_as(ctypes.c_void_p)

    def __call__(self, mode, t, interval, pos, alb, cloudo):
        """mode 0: the sun from the device clock t (1 double, s) over `interval` s at the cells' unit positions pos
        (N, 3); mode 1: a fixed mean cosine 0.5. alb (N) the surface albedo; cloudo as the longwave's or None."""
        _hip.call('isw_sw_column', self.N, self.L, self._par, ctypes.byref(self.tables), ctypes.byref(self.air),
                  _hip.ptr(self.lwm.sub), int(mode), _hip.ptr(t), float(interval), _hip.ptr(pos), _hip.ptr(alb),
                  _hip.ptr(cloudo), _hip.ptr(self.swf), self.prec, _hip.stream())
