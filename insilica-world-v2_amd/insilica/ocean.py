"""The ocean under the air: Boussinesq layers on the air's grid with temperature T and salinity S carried by FCT
transport, the exact air-sea wind stress, and the land column under land cells. The dynamical part of the reference's
`Ocean` (src/insilica/ocean.py:49-339): construction, `step` without the column mixing, `couple`, `ground`,
`update_cap`, `heat`, `salt`, `momentum`. The mixed layer's daily column processes (convection, Kraus-Turner
entrainment, penetrating sunlight, fresh-water fluxes) are not part of this build.
"""
import math

import numpy as np
import torch

from . import _hip
from .constants import GRAV
from .layers import Layers
from .transport import Transport

OCEAN = dict(depth=4000.0, h_mixed=100.0, h_thermo=600.0, T_thermo=293.15, T_deep=273.15, S_ref=35.0, rho=1025.0,
             cw=3990.0, dt=3600.0, gprime=(0.016, 0.016), hmin=10.0, visc4_U=0.5, bottom_stress=2.5e-4,
             interface_stress=5.0e-4, thin=2.0, walls=(), land_drag=1.0, h_layers=None, interface_cd=0.0)


def visc4(g, U):
    """nu4 = U d^3 / pi^4 with d the mean cell spacing (reference: src/insilica/world.py:48-52)."""
    d = float(np.mean(g.de))
    return U * d ** 3 / math.pi ** 4


def wall_mask(g, walls):
    """(E,) mask, 0 on edges whose two cells lie on opposite sides of a wall at one of the longitudes (degrees)."""
    m = np.ones(g.E)
    for lon in walls:
        d = np.angle(np.exp(1j * (g.lon - math.radians(lon))))
        a, b = d[g.eI], d[g.eJ]
        m[(np.abs(a) < math.pi / 2) & (np.abs(b) < math.pi / 2) & ((a >= 0) != (b >= 0))] = 0.0
    return m


class Ocean:
    def __init__(self, dg, cfg=None, Tsurf=None, state=None, land=None, c_land=2.0e6, grav=GRAV, soil=None):
        """cfg overrides OCEAN; Tsurf (N,) starts the mixed layer's temperature; state = (h, u, T[, S]) restores one;
        land (N, 0/1): cells with no water, walled off, their T[0] the ground's skin with heat capacity c_land; soil
        = dict(dz, c, k): the land column's layers under the skin."""
        self.dg = dg
        self.cfg = c = dict(OCEAN, **(cfg or {}))
        g = dg.host
        N, E = dg.N, dg.E
        nu4 = visc4(g, c['visc4_U'])
        depth = np.full(N, float(c['depth']))
        self.b = -depth
        L = self.L = len(c['gprime']) + 1
        upper = list(c['h_layers'] or (c['h_mixed'], c['h_thermo']))
        assert len(upper) == L - 1, 'one starting thickness per layer above the abyss'
        wall = wall_mask(g, c['walls']) if len(c['walls']) else None
        self.c_land, self.grav = c_land, grav
        self.soil = dict(soil) if soil and len(soil.get('dz') or ()) else None
        if self.soil is not None:
            assert len(self.soil['dz']) <= _hip.KSOIL, 'at most 16 soil layers'
        self.land = None
        if land is not None:
            m = wall if wall is not None else np.ones(g.E)
            m[(land[g.eI] > 0.5) | (land[g.eJ] > 0.5)] = 0.0
            wall = m
            self.land = dg.asarray(land)
        self.wall_host = wall
        self.core = Layers(dg, L, b=self.b, grav=grav, gprime=list(c['gprime']),
                           stress=[0.0] * (L - 1) + [c['bottom_stress']], nu4=nu4, floor=c['hmin'],
                           vmix=c['interface_stress'], vmix_cd=c['interface_cd'], thin=c['thin'], wall=wall)
        self.tr = Transport(dg, layers=L, tracers=2)
        if state is None:
            Ts = np.asarray(Tsurf, dtype=float)
            h = [None] * L
            h[L - 1] = np.maximum(depth - sum(upper), c['hmin'])
            below = h[L - 1]
            for k in range(L - 2, 0, -1):
                h[k] = np.maximum(np.minimum(depth - sum(upper[:k]) - below, upper[k]), c['hmin'])
                below = below + h[k]
            h[0] = depth - below
            h = np.stack(h)
            u = np.zeros((L, E))
            Tk = [c['T_thermo'] + (c['T_deep'] - c['T_thermo']) * (k - 1) / (L - 2) for k in range(1, L - 1)] if L > 2 else []
            T = np.stack([Ts] + [np.clip(Ts, c['T_deep'], t) for t in Tk] + [np.full(N, c['T_deep'])])
            S = np.full((L, N), c['S_ref'])
        else:
            h, u, T = state[:3]
            S = state[3] if len(state) > 3 else np.full((L, N), c['S_ref'])
        self.h, self.u = dg.asarray(h), dg.asarray(u)
        self.rows = dg.asarray(np.concatenate([np.asarray(T, dtype=float), np.asarray(S, dtype=float)]))
        self.T, self.S = self.rows[:L], self.rows[L:]
        if self.soil is not None:
            self.soil_dz = dg.asarray([float(d) for d in self.soil['dz']])
            self.Tg = self.T[0][None, :].repeat(len(self.soil['dz']), 1).contiguous()
        self.core.mask(self.u, L)
        self.M = self.h * dg.areaCell[None, :]
        self.nfast = int(math.ceil(1.2 * self.core.substeps(self.h, c['dt'])))
        z = dg.zeros
        self._Fm, self._M1 = z(L, E), z(L, N)
        self.cap = z(N)
        self.tau_acc, self.diss = z(E), z(E)
        self.update_cap()

    def _land(self):
        return self.land

    def update_cap(self):
        """The mixed layer's heat capacity rho_w c_w h_0 per cell (land's capacity stays once set)."""
        c = self.cfg
        _hip.call('isw_ocn_cap', self.dg.ref, _hip.ptr(self.h), c['rho'] * c['cw'], _hip.ptr(self.cap),
                  _hip.ptr(self.land), float(self.c_land), _hip.stream())

    def couple(self, air_h, air_u, c, dt, cd=None):
        """The wind stress over one air step, exactly; the kinetic energy it removed is left in self.diss (J/m^2)."""
        _hip.call('isw_couple_stress', self.dg.ref, int(air_h.shape[0]), _hip.ptr(air_h), _hip.ptr(air_u),
                  _hip.ptr(self.h), _hip.ptr(self.u), float(c), float(self.cfg['rho']), float(dt), _hip.ptr(self.diss),
                  _hip.ptr(self.tau_acc), _hip.ptr(self.core.wall), float(self.cfg['land_drag']), _hip.ptr(cd),
                  _hip.stream())

    def ground(self, dt):
        """The land column's conduction under the skin over dt (nothing without land and soil)."""
        if self.soil is not None and self.land is not None:
            _hip.call('isw_land_column', self.dg.ref, len(self.soil['dz']), _hip.ptr(self.land), _hip.ptr(self.T),
                      _hip.ptr(self.cap), _hip.ptr(self.Tg), _hip.ptr(self.soil_dz), float(self.soil['c']),
                      float(self.soil['k']), float(dt), _hip.stream())

    def step(self, mix=False):
        """One ocean step: the layers move and T, S are carried (the daily column mixing is not part of this build)."""
        if mix:
            raise NotImplementedError('the ocean column mixing is not part of this build')
        c = self.cfg
        self.core.step(self.h, self.u, c['dt'], self.nfast, self._Fm)
        self.tr.mass_after(self.M, self._Fm, self._M1)
        self.tr.step_rows(self.M, self._M1, self._Fm, self.rows, order=3)
        self.M.copy_(self._M1)
        self.update_cap()

    def heat(self):
        """Heat content (J) relative to 273.15 K: rho_w c_w sum A h (T - 273.15), plus the ground's where land."""
        A = self.dg.areaCell
        rc = self.cfg['rho'] * self.cfg['cw']
        if self.land is None:
            return float(torch.sum(A * torch.sum(self.h * (self.T - 273.15), dim=0))) * rc
        ground = self.cap * (self.T[0] - 273.15)
        if self.soil is not None:
            ground = ground + self.soil['c'] * torch.sum(self.soil_dz[:, None] * (self.Tg - 273.15), dim=0)
        return float(torch.sum(A * ((1 - self.land) * rc * torch.sum(self.h * (self.T - 273.15), dim=0)
                                    + self.land * ground)))

    def salt(self):
        A = self.dg.areaCell
        if self.land is not None:
            A = A * (1 - self.land)
        return float(torch.sum(A * torch.sum(self.h * self.S, dim=0))) * self.cfg['rho'] / 1000.0

    def momentum(self):
        g = self.dg
        he = 0.5 * (self.h[:, g.eI.long()] + self.h[:, g.eJ.long()])
        return float(torch.sum(torch.sum(he * self.u, dim=0) * g.le * g.de)) * self.cfg['rho']
