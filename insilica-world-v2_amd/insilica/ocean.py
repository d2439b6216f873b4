"""The ocean under the air: Boussinesq layers on the air's grid with temperature T and salinity S carried by FCT
transport, the exact air-sea wind stress, the land column under land cells, and the mixed layer's column processes:
the reference's `Ocean` (src/insilica/ocean.py:49-339): construction, `step` (with `mix`, the daily mixing:
convective adjustment, Kraus-Turner entrainment or shoaling, the interior interfaces' crossings, salt with the water
and the exchange between layers), `accumulate` (the air step's sunlight below the mixed layer and its fresh water),
`couple`, `ground`, `update_cap`, `heat`, `salt`, `momentum`. Density classes (T_classes), the currents' own density
(eos_currents), squeezed-layer flags (shelf_bed), sea-ice drift and the Ekman spreading (ek_depth) are not part of
this build.
"""
import ctypes
import math

import numpy as np
import torch

from . import _hip
from .constants import GRAV
from .layers import Layers
from .transport import Transport

ALPHA_REF = 2.0e-4          # 1/K: the buoyancy of dTmin, below which a mixed layer's base counts as unstratified
OCEAN = dict(depth=4000.0, h_mixed=100.0, h_thermo=600.0, T_thermo=293.15, T_deep=273.15, alpha=0.52e-4, alpha2=1.0e-5,
             beta=7.6e-4, S_ref=35.0, rho=1025.0, cw=3990.0, dt=3600.0, mix_dt=86400.0, gprime=(0.016, 0.016),
             hmin=10.0, hmax=1000.0, dTmin=0.1, m_kt=0.5, n_kt=0.2, visc4_U=0.5, bottom_stress=2.5e-4,
             interface_stress=5.0e-4, thin=2.0, walls=(), sw_R=0.58, sw_z1=0.35, sw_z2=23.0, land_drag=1.0,
             h_layers=None, kappa_d=0.0, dwf=0.0, kappa_h=0.0, fw_local=0.0, interface_cd=0.0, ri_b=0.0, reach=0.0,
             h_ent=0.0, sw_up=0.0, det_class=0.0)
UNSUPPORTED = ('T_classes', 'S_classes', 'eos_currents', 'shelf_bed', 'ice_drift', 'ek_depth', 'bed')


def ocean_struct(c, grav):
    """The settings the column kernels read (isw_ocean) from an Ocean's cfg."""
    o = _hip.IswOcean()
    o.grav, o.alpha, o.alpha2, o.beta = grav, c['alpha'], c['alpha2'], c['beta']
    o.S_ref, o.alpha_ref, o.rho_w, o.cw = c['S_ref'], ALPHA_REF, c['rho'], c['cw']
    for k in ('m_kt', 'n_kt', 'hmin', 'hmax', 'dTmin', 'sw_R', 'sw_z1', 'sw_z2', 'kappa_d', 'dwf', 'ri_b', 'reach',
              'h_ent', 'det_class', 'fw_local'):
        setattr(o, k, float(c[k]))
    return o


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
        for k in UNSUPPORTED:
            if c.get(k):
                raise NotImplementedError(f'ocean option {k!r} is not part of this build')
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
        self.cap = z(N)                                   # rho_w c_w h_0, for the air's surface physics
        self.qocean, self.qsw = z(N), z(N)                # W/m^2 into the mixed layer, and its sunlight
        self.q_acc, self.sun_acc = z(N), z(N)             # their non-solar part and the sunlight summed (J/m^2)
        self.fw_acc, self.ice_prev = z(N), z(N)           # fresh water in (kg/m^2), summed; the sea ice at the last step
        self.tau_acc, self.taum, self.diss = z(E), z(N), z(E)
        self.wx = z(L - 1, N)
        self._Tn, self._Sn = z(L, N), z(L, N)
        self._du, self.shear = (z(E), z(N)) if c['ri_b'] > 0 else (None, None)
        self.ostruct = ocean_struct(c, grav)
        ratio = c['mix_dt'] / c['dt']
        assert abs(ratio - round(ratio)) < 1e-9 and round(ratio) >= 1, 'mix_dt must be a whole number of ocean steps'
        self.mix_every = int(round(ratio))
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

    def accumulate(self, dt, evap=None, rain=None, ice=None, mouth=None):
        """The air step's sunlight below the mixed layer's base to the layers below, the fluxes the mixed layer keeps
        into the mixing's sums, and rain, evaporation and the sea ice's change (kg/m^2/s, kg/m^2) freshening or
        salting the mixed layer (reference: ocean.py:214-224)."""
        c, N, s = self.cfg, self.dg.N, _hip.stream()
        o = ctypes.byref(self.ostruct)
        if evap is not None:
            _hip.call('isw_ocn_fresh', N, o, _hip.ptr(self.h), _hip.ptr(self.S), _hip.ptr(evap), _hip.ptr(rain),
                      _hip.ptr(ice), _hip.ptr(self.ice_prev), float(dt), _hip.ptr(self.fw_acc), _hip.ptr(self.land),
                      _hip.ptr(mouth), _hip.ptr(self.h), _hip.ptr(self.M), _hip.ptr(self.dg.areaCell), s)
        _hip.call('isw_ocn_pen', N, self.L, o, _hip.ptr(self.h), _hip.ptr(self.T), _hip.ptr(self.qocean),
                  _hip.ptr(self.qsw), float(dt), _hip.ptr(self.q_acc), _hip.ptr(self.sun_acc), _hip.ptr(self.land),
                  _hip.ptr(self.wx if c['sw_up'] > 0 else None), s)

    def step(self, mix=True):
        """One ocean step: convective adjustment, the layers move and T, S are carried; with mix, the mixed layer then
        overturns, entrains or shoals on the forcing summed since the last mixing (reference: ocean.py:226-295)."""
        c, dg, N, s = self.cfg, self.dg, self.dg.N, _hip.stream()
        o = ctypes.byref(self.ostruct)
        _hip.call('isw_ocn_overturn', N, self.L, o, _hip.ptr(self.h), _hip.ptr(self.T), _hip.ptr(self.S),
                  _hip.ptr(self.land), s)
        self.core.step(self.h, self.u, c['dt'], self.nfast, self._Fm)
        self.tr.mass_after(self.M, self._Fm, self._M1)
        self.tr.step_rows(self.M, self._M1, self._Fm, self.rows, order=3)
        self.M.copy_(self._M1)
        self.update_cap()
        if not mix:
            return
        dt = float(c['mix_dt'])
        _hip.call('isw_ocn_forcing', dg.ref, dt, _hip.ptr(self.tau_acc), _hip.ptr(self.taum), _hip.ptr(self.q_acc),
                  _hip.ptr(self.sun_acc), _hip.ptr(self.fw_acc), s)
        if c['kappa_h'] > 0:
            _hip.call('isw_ocn_diffuse', dg.ref, self.L, _hip.ptr(self.h), _hip.ptr(self.T), _hip.ptr(self.S),
                      _hip.ptr(self._Tn), _hip.ptr(self._Sn), _hip.ptr(self.core.wall), float(c['kappa_h']), dt, s)
            self.T.copy_(self._Tn)
            self.S.copy_(self._Sn)
        if self.shear is not None:
            torch.sub(self.u[0], self.u[1], out=self._du)
            _hip.call('isw_cell_speed', dg.ref, _hip.ptr(self._du), _hip.ptr(self.shear), s)
        _hip.call('isw_ocn_column', N, self.L, o, _hip.ptr(self.h), _hip.ptr(self.T), _hip.ptr(self.S),
                  _hip.ptr(self.taum), _hip.ptr(self.q_acc), _hip.ptr(self.fw_acc), _hip.ptr(self.sun_acc), dt,
                  _hip.ptr(self.wx), _hip.ptr(self.land), _hip.ptr(self.shear), None, None, s)
        seq = 1 if c['reach'] > 0 else 0
        _hip.call('isw_ocn_xch_salt', N, self.L, _hip.ptr(self.wx), _hip.ptr(self.h), _hip.ptr(self.S), seq, s)
        self.core.exchange(self.wx, self.h, self.u, self.M, self.T, seq=bool(seq))
        for a in (self.tau_acc, self.q_acc, self.sun_acc, self.fw_acc):
            _hip.call('isw_fill', _hip.ptr(a), a.numel(), 0.0, s)
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
