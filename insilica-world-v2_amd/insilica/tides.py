"""The Moon's and the Sun's tides on the ocean (reference: src/insilica/tides.py, kernels/tides.cu): the
tide-generating force of both bodies on every ocean layer (through the layers' Bernoulli function), the bottom's
quadratic drag and the internal tide's drag, and the mixing the bottom stress's dissipation drives, with the
reference's settings, arguments and behaviour, over the HIP library's isw_tide_pot, isw_tide_itd and isw_tide_mix.
A world with params['tides'] (True / {} for Earth's Moon, or a dict of changes; moon None takes the Moon away) and the
coupled ocean under the air calls force() before each ocean step and mix() after it (reference: world.py:349-354).
"""
import math

import numpy as np
import torch

from . import _hip
from .ocean import ocean_struct

DAY = 86400.0
# Earth's Moon and Sun (reference: tides.py:23-30; the constants are the published ones: the Moon's GM, mean distance,
# eccentricity, inclination, sidereal month, 18.61-year node regression and 8.85-year perigee advance; the Sun's GM
# and the astronomical unit; Love factor 1 + k2 - h2 = 0.69; bottom drag 2.5e-3; the internal tide's wavenumber
# 2 pi / 10 km; the floor mixing's decay height 500 m)
TIDES = dict(moon=dict(gm=4.9028e12, a=3.84748e8, e=0.0549, i=5.145, T=27.321662 * DAY, node_T=-6798.38 * DAY,
                       aps_T=3232.6 * DAY, L0=0.0, node0=0.0, aps0=0.0),
             sun_gm=1.32712440018e20, sun_a=1.495978707e11, love=0.69, gamma=0.2,
             cd=2.5e-3, kappa_it=2 * math.pi / 10.0e3, zeta=500.0)


def settings(tides):
    """TIDES with a world's changes (tides: a dict, or True / {} for Earth's)."""
    c = {k: (dict(v) if isinstance(v, dict) else v) for k, v in TIDES.items()}
    if isinstance(tides, dict):
        for k, v in tides.items():
            if k == 'moon' and isinstance(v, dict):
                c['moon'].update(v)
            else:
                c[k] = v
    return c


def sky(p, radius, tides=True):
    """The bodies' places as isw_sky from the air physics' settings p (tilt, year_s, day_s, ecc, lam_peri), the
    planet's radius and the tide settings (reference: tides.sky_par)."""
    c = settings(tides)
    m = c['moon']
    s = _hip.IswSky(tilt=float(p['tilt']), year_s=float(p['year_s']), day_s=float(p['day_s']),
                    ecc=float(p.get('ecc', 0.0)), lam_peri=float(p.get('lam_peri', 0.0)), sun_gm=c['sun_gm'],
                    sun_a=c['sun_a'], love=c['love'], radius=float(radius))
    if m is None:                                  # (no Moon: its GM 0, a placeholder orbit)
        s.moon_gm, s.moon_a, s.moon_e, s.moon_i, s.moon_T, s.node_T, s.aps_T = 0.0, 1.0, 0.0, 0.0, 1.0, 1.0, 1.0
    else:
        s.moon_gm, s.moon_a, s.moon_e, s.moon_i = m['gm'], m['a'], m['e'], math.radians(m['i'])
        s.moon_T, s.node_T, s.aps_T = m['T'], m['node_T'], m['aps_T']
        s.moon_L0, s.node0, s.aps0 = math.radians(m['L0']), math.radians(m['node0']), math.radians(m['aps0'])
    return s


def sky_array(s):
    """An isw_sky's values in order (the reference's sky_par vector)."""
    return np.array([getattr(s, k) for k, _ in s._fields_])


class Tides:
    def __init__(self, dg, sea, p, radius, tides=True, rough=None, nbot=None):
        """sea: the world's Ocean; p: the air physics' settings; radius: the planet's (m); rough: the sea floor's
        roughness h^2 per cell (m^2) for the internal tide's drag, or None; nbot: the buoyancy frequency above the
        floor (1/s), or None (from the two lowest layers). (reference: tides.py:55-68)"""
        self.dg, self.sea, self.cfg = dg, sea, settings(tides)
        c = self.cfg
        self.sky = sky(p, radius, tides)
        N = dg.N
        self.V = dg.zeros(N)                       # the potential (m^2/s^2), read by the layers' Bernoulli function
        self.dsum = dg.zeros(N)                    # the bottom stress's dissipation, summed (J/m^2)
        self.pos = torch.as_tensor(np.ascontiguousarray(dg.host.pos, dtype=float), device=dg.dev)
        sea.core.tide = self.V
        sea.core.cd_bot = float(c['cd'])
        self.h2 = None if rough is None else torch.as_tensor(np.asarray(rough, dtype=float), device=dg.dev)
        self.nbot = None if nbot is None else torch.as_tensor(np.asarray(nbot, dtype=float), device=dg.dev)
        if self.h2 is not None:
            sea.core.cbot = dg.zeros(dg.E)         # the internal tide's drag per edge, set each ocean step
        self.ostruct = ocean_struct(sea.cfg, sea.grav)

    def force(self, time, back):
        """The potential at time[0] - back (s; the ocean step's middle), and the internal tide's drag."""
        s = self.sea
        _hip.call('isw_tide_pot', self.dg.ref, _hip.ctypes.byref(self.sky), _hip.ptr(time), float(back),
                  _hip.ptr(self.pos), _hip.ptr(self.V), _hip.stream())
        if self.h2 is not None:
            _hip.call('isw_tide_itd', self.dg.ref, s.L, _hip.ctypes.byref(self.ostruct), _hip.ptr(s.h), _hip.ptr(s.T),
                      _hip.ptr(s.S), _hip.ptr(self.h2), float(self.cfg['kappa_it']), _hip.ptr(self.nbot),
                      _hip.ptr(s.core.cbot), _hip.stream())

    def mix(self, dt):
        """The tidal mixing after an ocean step of dt seconds."""
        s, c = self.sea, self.sea.cfg
        _hip.call('isw_tide_mix', self.dg.ref, s.L, _hip.ctypes.byref(self.ostruct), _hip.ptr(s.h), _hip.ptr(s.u),
                  _hip.ptr(s.T), _hip.ptr(s.S), _hip.ptr(s.land), float(c['bottom_stress']), float(c['rho']),
                  float(self.cfg['gamma']), float(dt), _hip.ptr(self.dsum), float(self.cfg['cd']),
                  _hip.ptr(s.core.cbot), float(self.cfg['zeta']), _hip.stream())
