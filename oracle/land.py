"""The CPU oracle of the land's water routing and ice flow -- TEST INFRASTRUCTURE ONLY.

ctypes wrappers of land.c (the reference's kernels/land.cu as LandWater.step and ice_step sequence them, land.py:33-60)
on host numpy arrays; og is an oracle.Grid. Used by the tests as the checker of the HIP library.
"""
import numpy as np

from . import _p, lib


def _d(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def route(og, b, land, soilcap, pond, soil, ice, rho_w, rho_ice, u_dt, kseep_dt, mouth, qout=None):
    """One routing step: returns (pond, soil, mouth, qout) updated copies."""
    b, land, soilcap, ice = _d(b), _d(land), _d(soilcap), _d(ice)
    pond, soil, mouth = _d(pond).copy(), _d(soil).copy(), _d(mouth).copy()
    qout = None if qout is None else _d(qout).copy()
    lib().og_land_route(og.ref, _p(b), _p(land), _p(soilcap), _p(pond), _p(soil), _p(ice), float(rho_w), float(rho_ice),
                        float(u_dt), float(kseep_dt), _p(mouth), _p(qout))
    return pond, soil, mouth, qout


def ice_flow(og, b, land, ice, rho_ice, gam_dt, calved, ice_prev=None, iout=None):
    """One ice-flow step: returns (ice, calved, ice_prev, iout) updated copies."""
    b, land = _d(b), _d(land)
    ice, calved = _d(ice).copy(), _d(calved).copy()
    ice_prev = None if ice_prev is None else _d(ice_prev).copy()
    iout = None if iout is None else _d(iout).copy()
    lib().og_ice_flow(og.ref, _p(b), _p(land), _p(ice), float(rho_ice), float(gam_dt), _p(ice_prev), _p(calved),
                      _p(iout))
    return ice, calved, ice_prev, iout
