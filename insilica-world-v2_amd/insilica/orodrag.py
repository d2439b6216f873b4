"""The gravity waves' drag on the air (reference: src/insilica/orodrag.py, kernels/drag.cu): the mountains' drag of
Lott & Miller (1997) from each cell's sub-grid orography (OroDrag) and the spectrum of gravity waves not tied to the
ground of Alexander & Dunkerton (1999) (WaveDrag), with the reference's classes, arguments and behaviour, over the HIP
library's isw_oro_drag and isw_wave_drag. Both act on the air layers' winds after the wind stress and book the kinetic
energy they remove as heat in the ocean's dissipation accumulator (the reference's World._tick, world.py:311-314).
"""
import math

import numpy as np
import torch

from . import _hip
from .atmosphere import perot_operators
from .constants import GRAV, KAPPA, P0

RD = 287.0


def edge_normals(g):
    """Each edge's normal (the direction of positive u) as (east, north) components at its first and second cell,
    (E, 4): the chord from eI to eJ projected onto each cell's tangent plane and normalised, so that a uniform wind
    W at a cell gives u_e = W . n_e there (reference: orodrag.operators, orodrag.py:29-36)."""
    pos = np.asarray(g.pos, float)
    chord = pos[g.eJ] - pos[g.eI]
    out = np.empty((g.E, 4))
    for col, cell in ((0, np.asarray(g.eI)), (2, np.asarray(g.eJ))):
        r = pos[cell]
        tang = chord - np.einsum('ij,ij->i', chord, r)[:, None] * r
        tang /= np.linalg.norm(tang, axis=1)[:, None]
        lon = np.arctan2(r[:, 1], r[:, 0])
        lat = np.arcsin(np.clip(r[:, 2], -1.0, 1.0))
        out[:, col] = -np.sin(lon) * tang[:, 0] + np.cos(lon) * tang[:, 1]
        out[:, col + 1] = (-np.sin(lat) * np.cos(lon) * tang[:, 0] - np.sin(lat) * np.sin(lon) * tang[:, 1]
                           + np.cos(lat) * tang[:, 2])
    return out


def operators(g):
    """Perot's reconstruction of a cell's wind from its edges (N, 6, 2) and the edges' normals (E, 4)."""
    return perot_operators(g), edge_normals(g)


def _dev(dg, a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64).reshape(-1), device=dg.dev)


class OroDrag:
    """The mountains' drag. sso: (mu, sigma, gamma, theta) per cell -- the sub-grid heights' standard deviation (m),
    slope, anisotropy and the principal axis's angle from east (radians); p: the air physics' settings (oro_g,
    oro_cd, oro_hnc, oro_ric, and grav, kappa, air_rd when given). (reference: orodrag.py:38-68)"""

    def __init__(self, dg, L, sso, p, dt):
        g = dg.host
        op, nrm = operators(g)
        self.dg, self.L = dg, L
        self.op, self.nrm = _dev(dg, op), _dev(dg, nrm)
        self.sso = _dev(dg, np.stack([np.asarray(x, dtype=float) for x in sso]))
        self.par = _hip.IswOro(grav=p.get('grav', GRAV), p0=P0, kappa=p.get('kappa', KAPPA), rd=p.get('air_rd', RD),
                               G=p['oro_g'], cd=p['oro_cd'], hnc=p['oro_hnc'], ric=p['oro_ric'], dt=float(dt),
                               mu_min=1.0)
        self.beta = torch.zeros((L, g.N), dtype=torch.float64, device=dg.dev)
        self.F = torch.zeros((2 * L, g.N), dtype=torch.float64, device=dg.dev)

    def apply(self, h, theta, u, dt, diss):
        """One step's drag on the layers' winds u (in place); the kinetic energy removed is added to diss (J/m^2 per
        edge)."""
        assert dt == self.par.dt
        _hip.call('isw_oro_drag', self.dg.ref, self.L, _hip.ctypes.byref(self.par), _hip.ptr(h), _hip.ptr(theta),
                  _hip.ptr(u), _hip.ptr(self.op), _hip.ptr(self.nrm), _hip.ptr(self.sso), _hip.ptr(self.beta),
                  _hip.ptr(self.F), float(dt), _hip.ptr(diss), _hip.stream())


class WaveDrag:
    """The gravity waves not tied to the ground. p: the air physics' settings (gw_bt, gw_cw, gw_lambda, gw_psrc,
    gw_cmax, gw_dc, gw_conv). conv: the column's convective rain (N), set when gw_conv > 0. (reference:
    orodrag.py:71-96)"""

    def __init__(self, dg, L, p, dt):
        g = dg.host
        op, nrm = operators(g)
        self.dg, self.L, self.dt = dg, L, dt
        self.op, self.nrm = _dev(dg, op), _dev(dg, nrm)
        self.par = _hip.IswGw(grav=p.get('grav', GRAV), p0=P0, kappa=p.get('kappa', KAPPA), rd=p.get('air_rd', RD),
                              bt=p['gw_bt'], cw=p['gw_cw'], kw=2 * math.pi / p['gw_lambda'], psrc=p['gw_psrc'],
                              cmax=p['gw_cmax'], dc=p['gw_dc'], dt=float(dt), fc=p.get('gw_conv', 0.0))
        self.F = torch.zeros((2 * L, g.N), dtype=torch.float64, device=dg.dev)
        self.zero = torch.zeros((L, g.N), dtype=torch.float64, device=dg.dev)
        self.conv = None
        self.cmean = torch.zeros(1, dtype=torch.float64, device=dg.dev) if p.get('gw_conv', 0.0) > 0 else None

    def apply(self, h, theta, u, dt, diss):
        """One step's wave forcing on u (in place); the kinetic energy it removed (negative: gave) is added to diss."""
        assert dt == self.dt
        conv = self.conv if self.cmean is not None else None
        _hip.call('isw_wave_drag', self.dg.ref, self.L, _hip.ctypes.byref(self.par), _hip.ptr(h), _hip.ptr(theta),
                  _hip.ptr(u), _hip.ptr(self.op), _hip.ptr(self.nrm), _hip.ptr(conv), _hip.ptr(self.cmean),
                  _hip.ptr(self.zero), _hip.ptr(self.F), float(dt), _hip.ptr(diss), _hip.stream())
