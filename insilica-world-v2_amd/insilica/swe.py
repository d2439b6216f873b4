"""Shallow water on the TRiSK C-grid: the reference's `SWE` (src/insilica/swe.py:90-132) over libinsilica_hip.so,
L independent layers stacked as (L, N) and (L, E).

    dh/dt = -div(h_e u)
    du/dt = sum_e' w_ee' F_e' (q_e + q_e')/2 - d(g(h + b) + K)/dn - r u - nu4 del^4 u

Wicker-Skamarock RK3 with the positivity limiter in every stage; del^4 from the step's start wind held through the
stages. One isw_swe_step call per step; nothing is allocated per step.
"""
from .constants import GRAV
from .layers import Layers
from . import _hip


class SWE:
    def __init__(self, dg, layers=1, grav=GRAV, b=None, drag=0.0, nu4=0.0):
        self.dg, self.L, self.grav, self.nu4 = dg, layers, grav, nu4
        # the layer stack's scratch and settings serve the SWE entry point (it reads F, B, q, scale, h1/h2, u1/u2,
        # the del^4 buffers, grav, drag, nu4, b and fV from them)
        self.core = Layers(dg, layers, b=b, grav=grav, gprime=[0.0] * (layers - 1), drag=drag, nu4=nu4)
        self.b = self.core.b

    def step(self, h, u, dt, Fm=None):
        """One RK3 step in place on h (L, N) and u (L, E). Fm, if given, receives the final stage's l dt F."""
        _hip.call('isw_swe_step', self.dg.ref, self.core.ref, _hip.ptr(h), _hip.ptr(u), _hip.ptr(Fm), float(dt),
                  _hip.stream())
