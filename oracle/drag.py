"""The CPU oracle of the gravity waves' drag on the air -- TEST INFRASTRUCTURE ONLY.

ctypes wrappers of drag.c (the reference's kernels/drag.cu as OroDrag.apply and WaveDrag.apply sequence them,
orodrag.py:62-96) on host numpy arrays; og is an oracle.Grid. Used by the tests as the checker of the HIP library.
"""
import numpy as np

from . import _p, lib


def _d(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def oro_drag(og, mu, theta, u, op, nrm, sso, par, dt, diss):
    """One step's mountain drag: returns (u, diss, beta, F) -- updated copies of u and diss, and the cell pass's
    implicit factors (L, N) and forces (2L, N)."""
    L, N, E = mu.shape[0], og.N, og.E
    mu, theta, op, nrm, sso, par = _d(mu), _d(theta), _d(op), _d(nrm), _d(sso), _d(par)
    u, diss = _d(u).copy(), _d(diss).copy()
    beta, F = np.zeros((L, N)), np.zeros((2 * L, N))
    a = og.arr
    lib().og_oro_drag(N, E, L, _p(a['cellEdge']), _p(a['eI']), _p(a['eJ']), _p(mu), _p(theta), _p(u), _p(op), _p(nrm),
                      _p(sso), _p(par), _p(beta), _p(F), float(dt), _p(diss))
    return u, diss, beta, F


def wave_drag(og, mu, theta, u, op, nrm, par, dt, diss, conv=None):
    """One step's non-orographic wave drag: returns (u, diss, F) -- updated copies of u and diss, and the forces."""
    L, N, E = mu.shape[0], og.N, og.E
    mu, theta, op, nrm, par, conv = _d(mu), _d(theta), _d(op), _d(nrm), _d(par), _d(conv)
    u, diss = _d(u).copy(), _d(diss).copy()
    F, zero = np.zeros((2 * L, N)), np.zeros((L, N))
    a = og.arr
    lib().og_wave_drag(N, E, L, _p(a['cellEdge']), _p(a['eI']), _p(a['eJ']), _p(a['areaCell']), _p(mu), _p(theta),
                       _p(u), _p(op), _p(nrm), _p(par), _p(conv), _p(zero), _p(F), float(dt), _p(diss))
    return u, diss, F


def gw_mean(og, conv):
    return float(lib().og_gw_mean(og.N, _p(_d(conv)), _p(og.arr['areaCell'])))


def blk_int(t, A):
    return float(lib().og_blk_int(float(t), float(A)))
