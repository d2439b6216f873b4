"""The CPU oracle of the Held-Suarez column and the surface exchange -- TEST INFRASTRUCTURE ONLY.

ctypes wrappers of column.c (the reference's double path of air.cu hs_column with its crossing, surface_layer,
cell_speed, cell_speed_perot, gas_src) on host numpy arrays. Used by the tests as the checker of the HIP kernels.
"""
import ctypes

import numpy as np

from . import _p, lib
from .radiation import air_struct


def _d(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def hs_column(par, hs, lat, mu, theta, dt):
    """air.cu hs_column: returns (theta, wx, qnet, stored); mu is read only."""
    par, hs, lat, mu = _d(par), _d(hs), _d(lat), _d(mu)
    theta = _d(theta).copy()
    L, N = mu.shape
    wx, qnet, stored = np.zeros((max(L - 1, 1), N)), np.zeros(N), np.zeros(N)
    lib().og_hs_column(N, L, _p(par), _p(hs), _p(lat), _p(mu), _p(theta), float(dt), _p(wx), _p(qnet), _p(stored))
    return theta, wx, qnet, stored


def surface_layer(par, uabs, mu, theta, Ts, qv=None, ice=None, land=None, lc=None, air=None):
    """air.cu surface_layer: returns (cd, ch)."""
    par, uabs, mu, theta, Ts, qv, ice, land, lc = (_d(x) for x in (par, uabs, mu, theta, Ts, qv, ice, land, lc))
    L, N = mu.shape
    cd, ch = np.zeros(N), np.zeros(N)
    lib().og_surface_layer(N, L, _p(par), ctypes.byref(air_struct(air)), _p(uabs), _p(mu), _p(theta), _p(qv), _p(Ts),
                           _p(ice), _p(land), _p(cd), _p(ch), _p(lc))
    return cd, ch


def cell_speed(g, u):
    out = np.zeros(g.N)
    ce = np.ascontiguousarray(g.cellEdge, dtype=np.int32)
    lib().og_cell_speed(g.N, _p(_d(u)), _p(ce), _p(_d(g.le)), _p(_d(g.de)), _p(_d(g.areaCell)), _p(out))
    return out


def cell_speed_perot(g, u, op):
    out = np.zeros(g.N)
    ce = np.ascontiguousarray(g.cellEdge, dtype=np.int32)
    lib().og_cell_speed_perot(g.N, _p(_d(u)), _p(ce), _p(_d(op)), _p(out))
    return out


def gas_src(src, h, dt, x, L):
    src, h = _d(src), _d(h)
    x = _d(x).copy()
    G, N = src.shape
    lib().og_gas_src(N, L, G, _p(src), _p(h), float(dt), _p(x))
    return x


def air_column(par, mode, t, pos, uabs, mu, theta, Ts, dt, insol=None, cap=None, elev=None, radf=None, swf=None,
               ch=None, firn=None, air=None, mix=False, ocean=False, q=None, ice=None, cw=None, cloudo=False,
               land=None, soilcap=None, soil=None, pond=None, snow=None, lc=None, xfer=None, dgo=False):
    """air.cu air_column (the dry column; the moist one, par's moist > 0, with q (L, N) and ice (N); clouds with cloud
    > 0 and the cloud water cw (L, N)): returns a dict of theta, Ts, wx, olr, sw_abs, stored, rain, evap, conv, tsurf,
    rainp, flash, swalb (and qocean, qsw with ocean=True; firn when given; q and ice when moist; cw, and cloudo
    ((4L + 3), N) with cloudo=True, with clouds; over land (moist, the mask land (N) > 0.5, soilcap (N)) soil, pond and
    snow (N), and cap when given (c_wet rewrites it); lc (8, N) the land cover's rows; xfer (3 L, N) the gas
    tracers' mole fractions per layer, carried by the column's moves of air; dgo=True: the diagnostics rows (5 L - 2,
    N): each layer's heating Q and latent heating (W/m^2), each interface's gross up and down crossing, each layer's
    convective mass). mu is read only."""
    par, t, pos, uabs, mu, insol, elev, radf, swf, ch, land, soilcap, lc = (
        _d(x) for x in (par, t, pos, uabs, mu, insol, elev, radf, swf, ch, land, soilcap, lc))
    cap, soil, pond, snow = (None if x is None else _d(x).copy() for x in (cap, soil, pond, snow))
    theta, Ts = _d(theta).copy(), _d(Ts).copy()
    q = None if q is None else _d(q).copy()
    ice = None if ice is None else _d(ice).copy()
    cw = None if cw is None else _d(cw).copy()
    firn = None if firn is None else _d(firn).copy()
    xfer = None if xfer is None else _d(xfer).copy()
    L, N = mu.shape
    z = lambda *s: np.zeros(s)
    o = dict(theta=theta, Ts=Ts, wx=z(max(L - 1, 1), N), olr=z(N), sw_abs=z(N), stored=z(N), rain=z(N), evap=z(N),
             conv=z(N), tsurf=z(N), rainp=z(4, N), flash=np.full(N, -1.0), swalb=z(N))
    if ocean:
        o['qocean'], o['qsw'] = z(N), z(N)
    if cw is not None and cloudo:
        o['cloudo'] = z(4 * L + 3, N)
    if dgo:
        o['dgo'] = z(5 * L - 2, N)
    lib().og_air_column(N, L, _p(par), ctypes.byref(air_struct(air)), int(mode), _p(t), _p(insol), _p(pos), _p(uabs),
                        _p(mu), _p(theta), _p(q), _p(Ts), _p(ice), float(dt), _p(o['wx']), _p(o['olr']), _p(o['sw_abs']),
                        _p(o['stored']), _p(o['rain']), _p(o['evap']), _p(o['conv']), _p(o['tsurf']), _p(o['rainp']),
                        _p(cap), _p(o.get('qocean')), _p(o.get('qsw')), _p(elev), _p(o['flash']), _p(radf), _p(swf),
                        _p(o['swalb']), _p(ch), _p(firn), _p(cw), _p(o.get('cloudo')), _p(land), _p(soilcap),
                        _p(soil), _p(pond), _p(snow), _p(lc), _p(xfer), _p(o.get('dgo')), int(bool(mix)))
    if xfer is not None:
        o['xfer'] = xfer
    if firn is not None:
        o['firn'] = firn
    if q is not None:
        o['q'], o['ice'] = q, ice
    if cw is not None:
        o['cw'] = cw
    if land is not None:
        o['soil'], o['pond'], o['snow'] = soil, pond, snow
        if cap is not None:
            o['cap'] = cap
    return o


def plume_ascent(par, mu, theta, q, th0, q0, pi_, shal=False, mix=False, air=None):
    """air.cu plume_ascent alone for one column (mu, theta, q: (L,)): a dict of A (buoyancy by mass), kd (the layer it
    detrains into, -1 if none), thd, qd (its theta and vapor there), G (its mass per kg), FZ (frozen condensate), lt
    (the lightning inputs), ptop (the pressure of its last buoyant sub-level), qx (L) what its mixing gave each layer
    per kg and Gl (L) what it took in from each (ent_ifs)."""
    par, mu, theta, q = _d(par), _d(mu), _d(theta), _d(q)
    L = mu.size
    out, qx, Gl = np.zeros(11), np.zeros(L), np.zeros(L)
    lib().og_plume_ascent(L, _p(par), ctypes.byref(air_struct(air)), _p(mu), _p(theta), _p(q), float(th0), float(q0),
                          float(pi_), int(bool(shal)), int(bool(mix)), _p(out), _p(qx), _p(Gl))
    return dict(A=out[0], kd=int(out[1]), thd=out[2], qd=out[3], G=out[4], FZ=out[5], lt=out[6:10], ptop=out[10], qx=qx,
                Gl=Gl)
