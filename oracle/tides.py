"""The CPU oracle of the Moon's and the Sun's tides -- TEST INFRASTRUCTURE ONLY.

ctypes wrappers of tides.c (the reference's kernels/tides.cu as Tides.force and Tides.mix sequence them,
tides.py:70-87) on host numpy arrays, in place; og is an oracle.Grid. Used by the tests as the checker of the HIP
library.
"""
import numpy as np

from . import _p, lib


def _d(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def tide_pot(og, tp, t, back, pos, V):
    """The potential at time t - back into V (N) in place; tp the 19 sky values (insilica.tides.sky_array)."""
    lib().og_tide_pot(og.N, _p(_d(tp)), float(t), float(back), _p(_d(pos)), _p(V))


def tide_itd(og, L, h, T, S, rough, kappa, eos, cbot, nbot=None):
    """The internal tide's drag per edge into cbot (E) in place; eos = (grav, alpha, alpha2, beta)."""
    a = og.arr
    lib().og_tide_itd(og.N, og.E, L, _p(h), _p(T), _p(S), _p(_d(rough)), _p(a['eI']), _p(a['eJ']), float(kappa),
                      *[float(x) for x in eos], _p(cbot), _p(_d(nbot)))


def tide_mix(og, L, h, u, T, S, land, c_bot, rho, gamma, dt, eos, dsum, cd, cbot, zeta):
    """The tidal mixing of T, S (L, N) in place, its dissipation added to dsum (N)."""
    a = og.arr
    lib().og_tide_mix(og.N, og.E, L, _p(h), _p(u), _p(T), _p(S), _p(a['cellEdge']), _p(a['le']), _p(a['de']),
                      _p(a['areaCell']), _p(_d(land)), float(c_bot), float(rho), float(gamma), float(dt),
                      *[float(x) for x in eos], _p(dsum), float(cd), _p(cbot), float(zeta))
