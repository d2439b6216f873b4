"""The CPU oracle of the k-table radiation -- TEST INFRASTRUCTURE ONLY.

ctypes wrappers of radiation.c (og_lw_setup, og_lw_column, og_sw_column, og_sw_layer): the reference's double path
(src/insilica/kernels/air.cu:2111-2511) on host numpy arrays, with the tables in the reference's npz format or as the
dicts insilica.ktables makes. Used by the tests as the checker of the HIP kernels and by bench.py's CPU baseline.
"""
import ctypes

import numpy as np

from . import _p, lib

_f32p = ctypes.POINTER(ctypes.c_float)
EARTH_AIR = dict(m_air=28.9647, m_air_kg=28.9647e-3, o2=0.2095, ray=1.0, broad=1.0, o3=1.0, eps=0.622, rd=287.0,
                 vf=0.61)
CLOUD_SSA = (0.60, 0.75, 0.98, 0.994, 0.9995, 0.9998, 0.99999, 0.99999, 0.99999, 0.99999)


class OGAir(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_double) for k in ('m_air', 'm_air_kg', 'o2', 'ray', 'broad', 'o3', 'eps', 'rd',
                                                    'vf')]
                + [('cia', ctypes.c_int32), ('pad_', ctypes.c_int32)])


class OGLwt(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_int32) for k in ('NC', 'NP', 'NT', 'NG', 'NTP', 'NO3')]
                + [(k, ctypes.c_double) for k in ('lp0', 'dlp', 't0', 'dtt', 'tp0', 'o3lp0', 'o3dlp')]
                + [(k, _f32p) for k in ('logk', 'planck', 'o3', 'cia')])


class OGSwt(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_int32) for k in ('NC', 'NP', 'NT', 'NG')]
                + [(k, ctypes.c_double) for k in ('yscale', 'tau550', 'hssa', 'hg', 'inhom')]
                + [(k, _f32p) for k in ('logk', 'sun', 'ray', 'hzf', 'cssa', 'cssa_i', 'g_i', 'cia')])


def air_struct(air=None):
    a = OGAir()
    for k, v in dict(EARTH_AIR, **(air or {})).items():
        if k != 'cia':
            setattr(a, k, float(v))
    a.cia = 1 if (air or {}).get('cia') else 0
    return a


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _fp(a):
    return None if a is None else ctypes.cast(a.ctypes.data, _f32p)


class LwTables:
    """The longwave tables as the oracle reads them (kept alive with their struct)."""

    def __init__(self, t, cia=None):
        self.logk, self.planck, self.o3 = _f32(t['logk']), _f32(t['planck']), _f32(t['o3'])
        self.cia = None if cia is None else _f32(cia)
        P, T, Tp = np.asarray(t['P'], float), np.asarray(t['T'], float), np.asarray(t['T_pl'], float)
        s = OGLwt()
        s.NC, s.NP, s.NT, s.NG = self.logk.shape
        s.NTP, s.NO3 = self.planck.shape[0], len(self.o3)
        s.lp0, s.dlp = float(np.log(P[0])), float((np.log(P[0]) - np.log(P[-1])) / (len(P) - 1))
        s.t0, s.dtt, s.tp0 = float(T[0]), float(T[1] - T[0]), float(Tp[0])
        s.o3lp0, s.o3dlp = s.lp0, s.dlp
        s.logk, s.planck, s.o3, s.cia = _fp(self.logk), _fp(self.planck), _fp(self.o3), _fp(self.cia)
        self.s = s


class SwTables:
    def __init__(self, t, inhom=0.7, cssa=None, cssa_i=None, g_i=None, cia=None):
        self.logk = _f32(t['logk'])
        sun = np.asarray(t['sun'], float)
        self.sun, self.ray, self.hzf = _f32(sun / sun.sum()), _f32(t['ray']), _f32(t['haze'])
        ng = np.asarray(t['ng'], int)
        self.cssa = _f32(np.repeat(t.get('cloud_ssa', CLOUD_SSA), ng) if cssa is None else cssa)
        self.cssa_i = None if cssa_i is None else _f32(cssa_i)
        self.g_i = None if g_i is None else _f32(g_i)
        self.cia = None if cia is None else _f32(cia)
        hz = np.asarray(t['haze_par'], float)
        s = OGSwt()
        s.NC, s.NP, s.NT, s.NG = self.logk.shape
        s.yscale = float(10.0 / (t['T'][1] - t['T'][0]))
        s.tau550, s.hssa, s.hg, s.inhom = float(hz[0]), float(hz[1]), float(hz[2]), float(inhom)
        for k in ('logk', 'sun', 'ray', 'hzf', 'cssa', 'cssa_i', 'g_i', 'cia'):
            setattr(s, k, _fp(getattr(self, k)))
        self.s = s


def _d(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def lw_setup(N, L, par, lwt, mu, theta, q=None, cloudo=None, gx=None, co2=400e-6, n2o=0.32e-6, ch4=1.8e-6,
             hs=1500.0, tmode=2, air=None):
    """air.cu lw_setup: the sub-levels (4L * 15, N) float32."""
    sub = np.zeros((4 * L * 15, N), np.float32)
    par, mu, theta, q, cloudo, gx = (_d(x) for x in (par, mu, theta, q, cloudo, gx))
    lib().og_lw_setup(N, L, _p(par), ctypes.byref(lwt.s), ctypes.byref(air_struct(air)), co2, n2o, ch4, _p(gx),
                      _p(mu), _p(theta), _p(q), _p(cloudo), hs, tmode, _p(sub))
    return sub


def lw_column(N, L, par, lwt, sub, Ts, nsub=4):
    """air.cu lw_column: radf (3(L+1)+3+6L, N)."""
    radf = np.zeros((3 * (L + 1) + 3 + 6 * L, N))
    sub, par, Ts = np.ascontiguousarray(sub, dtype=np.float32), _d(par), _d(Ts)
    lib().og_lw_column(N, L, nsub, _p(par), ctypes.byref(lwt.s), _p(sub), _p(Ts), _p(radf))
    return radf


def sw_column(N, L, par, swt, sub, alb, mode=1, tnow=0.0, interval=3600.0, pos=None, cloudo=None, air=None,
              nsub=4):
    """air.cu sw_column: swf (3L+2, N)."""
    swf = np.zeros((3 * L + 2, N))
    sub, par, alb, pos, cloudo = np.ascontiguousarray(sub, dtype=np.float32), _d(par), _d(alb), _d(pos), _d(cloudo)
    lib().og_sw_column(N, L, nsub, _p(par), ctypes.byref(swt.s), ctypes.byref(air_struct(air)), _p(sub), mode,
                       tnow, interval, _p(pos), _p(alb), _p(cloudo), _p(swf))
    return swf


def sw_layer(tau, ssa, g, mu0):
    """air.cu sw_layer: (Rd, Td, Rs, Ts, Tdd)."""
    out = np.zeros(5)
    lib().og_sw_layer(float(tau), float(ssa), float(g), float(mu0), _p(out))
    return out
