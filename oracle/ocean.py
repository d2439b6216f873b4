"""The oracle's ocean column processes -- TEST INFRASTRUCTURE ONLY.

ctypes wrappers of oracle/ocean.c (the reference's kernels/ocean.cu mixing kernels restated in C), on NumPy arrays
(layer-major, C-contiguous, float64). Settings come as a dict with the keys of the reference's OCEAN cfg
(ocean.py:29-33) plus the equation of state's grav, alpha, alpha2, beta and ALPHA_REF as alpha_ref.
"""
import ctypes

import numpy as np

from . import _p, lib

_BOUND = False


def _bind():
    global _BOUND
    if _BOUND:
        return lib()
    L = lib()
    vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    sig = {
        'og_ocn_overturn': [i, i, vp, vp, vp, d, d, d, d, vp],
        'og_ocn_column': [i, i] + [vp] * 7 + [d] * 17 + [vp, vp, d, d, vp, d, d, d, vp, d, vp],
        'og_ocn_xch_salt': [i, i, vp, vp, vp, i],
        'og_ocn_pen': [i, i, vp, vp, vp, vp, d, d, d, d, d, vp, vp, vp, vp],
        'og_ocn_fresh': [i, vp, vp, vp, vp, vp, vp, d, d, d, vp, vp, vp, d, vp, vp, vp],
        'og_ocn_diffuse': [vp, i, vp, vp, vp, vp, vp, vp, d, d],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.restype, f.argtypes = None, args
    _BOUND = True
    return L


def _c(a):
    return None if a is None else np.ascontiguousarray(a, dtype=float)


def overturn(N, L, h, T, S, grav, alpha, alpha2, beta, land=None):
    _bind().og_ocn_overturn(N, L, _p(h), _p(T), _p(S), grav, alpha, alpha2, beta, _p(_c(land)))


def column(N, L, p, h, T, S, taum, qns, qfw, qsun, dt, wx, land=None, shear=None, Tc=None, Sc=None):
    """ocn_column over one mixing interval dt; T, S adjusted in place; wx (L - 1, N) written."""
    _bind().og_ocn_column(N, L, _p(h), _p(T), _p(S), _p(_c(taum)), _p(_c(qns)), _p(_c(qfw)), _p(_c(qsun)),
                          p['grav'], p['alpha'], p['alpha2'], p['beta'], p['S_ref'], p['alpha_ref'], p['rho_w'],
                          p['cw'], p['m_kt'], p['n_kt'], p['hmin'], p['hmax'], p['dTmin'], p['sw_R'], p['sw_z1'],
                          p['sw_z2'], dt, _p(wx), _p(_c(land)), p['kappa_d'], p['dwf'], _p(_c(shear)), p['ri_b'],
                          p['reach'], p['h_ent'], _p(_c(Tc)), p['det_class'], _p(_c(Sc)))


def xch_salt(N, L, wx, h, S, seq=False):
    _bind().og_ocn_xch_salt(N, L, _p(_c(wx)), _p(h), _p(S), 1 if seq else 0)


def pen(N, L, p, h, T, qocean, qsw, dt, qns_acc, sun_acc, land=None, up=None):
    _bind().og_ocn_pen(N, L, _p(h), _p(T), _p(_c(qocean)), _p(_c(qsw)), p['rho_w'] * p['cw'], p['sw_R'], p['sw_z1'],
                       p['sw_z2'], dt, _p(qns_acc), _p(sun_acc), _p(_c(land)), _p(_c(up)))


def fresh(N, p, h, S, evap, rain, ice, ice_prev, dt, fw_acc, land=None, mouth=None, hw=None, M=None, area=None):
    _bind().og_ocn_fresh(N, _p(h), _p(S), _p(_c(evap)), _p(_c(rain)), _p(_c(ice)), _p(ice_prev), p['S_ref'],
                         p['rho_w'], dt, _p(fw_acc), _p(_c(land)), _p(mouth), p['fw_local'], _p(hw), _p(M),
                         _p(_c(area)))


def diffuse(og, L, h, T, S, Tn, Sn, wall, kappa_h, dt):
    _bind().og_ocn_diffuse(og.ref, L, _p(_c(h)), _p(_c(T)), _p(_c(S)), _p(Tn), _p(Sn), _p(_c(wall)), kappa_h, dt)
