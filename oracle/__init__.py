"""The CPU oracle -- TEST INFRASTRUCTURE ONLY.

ctypes bindings of liboracle.so (oracle.c: a scalar C restatement of the reference's double-precision kernels and of
the Python drivers that sequence them: layers.py, transport.py, swe.py; radiation.c: the k-table radiation, wrapped
in oracle/radiation.py). Only the tests, __graft_entry__.smoke() and
bench.py's CPU-baseline leg use it, as the checker of the HIP library and as the timed CPU path; the product never
imports it. `build()` compiles it with gcc (-O2 -ffp-contract=off -fopenmp).
"""
import ctypes
import math
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / 'liboracle.so'
LMAX = 8

_i32p, _i8p, _f64p = (ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int8), ctypes.POINTER(ctypes.c_double))


def build():
    """Compile liboracle.so from oracle.c, radiation.c, column.c, ocean.c, land.c, drag.c and tides.c (seconds)."""
    cmd = ['gcc', '-O2', '-ffp-contract=off', '-fopenmp', '-fPIC', '-shared', '-o', str(LIB) + '.tmp',
           str(HERE / 'oracle.c'), str(HERE / 'radiation.c'), str(HERE / 'column.c'), str(HERE / 'ocean.c'),
           str(HERE / 'land.c'), str(HERE / 'drag.c'), str(HERE / 'tides.c'), '-lm']
    subprocess.run(cmd, check=True)
    os.replace(str(LIB) + '.tmp', LIB)


class OGGrid(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_int32) for k in ('N', 'E', 'NV', 'n_eoe')]
                + [(k, _i32p) for k in ('eI', 'eJ', 'nbr', 'cellEdge', 'cellEdgeAsc', 'ev1', 'ev2', 'vEdge', 'vCell',
                                        'eoeStart', 'eoe', 'd2idx')]
                + [(k, _i8p) for k in ('cellSign', 'vEdgeSign', 'vSlot')]
                + [(k, _f64p) for k in ('le', 'de', 'areaCell', 'areaV', 'fV', 'kite', 'w', 'd2wt')])


class OGLayersCfg(ctypes.Structure):
    _fields_ = ([('L', ctypes.c_int32), ('air', ctypes.c_int32)]
                + [(k, ctypes.c_double) for k in ('grav', 'cp', 'kappa', 'p0', 'rd', 'nu4', 'floor_', 'div_damp', 'thin',
                                                  'shelf', 'vmix', 'vmix_cd', 'thin_wind')]
                + [(k, ctypes.c_double * LMAX) for k in ('gp', 'drag', 'stress')]
                + [(k, _f64p) for k in ('fV', 'b', 'wall', 'tide', 'cbot')]
                + [('cd_bot', ctypes.c_double)])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        sig = {
            'og_layers_new': (vp, [vp, vp]), 'og_layers_free': (None, [vp]), 'og_layers_Gc': (vp, [vp]),
            'og_layers_tides': (None, [vp, vp, vp, d]),
            'og_layers_L2': (vp, [vp]), 'og_layers_diss': (vp, [vp]),
            'og_layers_gfast': (None, [vp, vp, vp]), 'og_layers_substeps': (i, [vp, vp, d]),
            'og_layers_step': (None, [vp, vp, vp, d, i, vp, vp]), 'og_layers_lap': (None, [vp, vp, vp]),
            'og_layers_exchange': (None, [vp, vp, vp, vp, vp, vp, vp, i, d]),
            'og_mass_update': (None, [vp, i, vp, vp, vp]),
            'og_transport': (None, [vp, i, i, d, vp, vp, vp, vp, i]),
            'og_swe_step': (None, [vp, i, d, vp, vp, d, vp, vp, d, vp]),
            'og_couple_stress': (None, [vp, i, vp, vp, vp, vp, d, d, d, vp, vp, vp, d, vp]),
            'og_land_column': (None, [vp, i, vp, vp, vp, vp, vp, d, d, d]),
            'og_ocn_cap': (None, [vp, vp, d, vp, vp, d]),
            'og_vmix_heat': (None, [vp, i, vp, vp, d, d, d, d, vp]),
            'og_lay_vmix': (None, [vp, i, vp, d, d, vp, vp, d]),
            'og_dsum': (d, [vp, vp, ctypes.c_int64]),
            'og_lw_setup': (None, [i, i, vp, vp, vp, d, d, d, vp, vp, vp, vp, vp, d, i, vp]),
            'og_lw_column': (None, [i, i, i, vp, vp, vp, vp, vp]),
            'og_sw_column': (None, [i, i, i, vp, vp, vp, vp, i, d, d, vp, vp, vp, vp]),
            'og_sw_layer': (None, [d, d, d, d, vp]),
            'og_hs_column': (None, [i, i, vp, vp, vp, vp, vp, d, vp, vp, vp]),
            'og_surface_layer': (None, [i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
            'og_cell_speed': (None, [i, vp, vp, vp, vp, vp, vp]),
            'og_cell_speed_perot': (None, [i, vp, vp, vp, vp]),
            'og_gas_src': (None, [i, i, i, vp, vp, d, vp]),
            'og_air_column': (None, [i, i, vp, vp, i] + [vp] * 9 + [d] + [vp] * 29 + [i]),
            'og_plume_ascent': (None, [i, vp, vp, vp, vp, vp, d, d, d, i, i, vp, vp, vp]),
            'og_land_route': (None, [vp, vp, vp, vp, vp, vp, vp, d, d, d, d, vp, vp]),
            'og_ice_flow': (None, [vp, vp, vp, vp, d, d, vp, vp, vp]),
            'og_sun_lon': (d, [d, vp, vp]),
            'og_oro_drag': (None, [i, i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, d, vp]),
            'og_wave_drag': (None, [i, i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, d, vp]),
            'og_gw_mean': (d, [i, vp, vp]),
            'og_blk_int': (d, [d, d]),
            'og_tide_pot': (None, [i, vp, d, d, vp, vp]),
            'og_tide_mix': (None, [i, i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, d, d, d, d, d, d, d, d, vp, d, vp, d]),
            'og_tide_itd': (None, [i, i, i, vp, vp, vp, vp, vp, vp, d, d, d, d, d, vp, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def _p(a):
    """Pointer of a C-contiguous array (None passes NULL)."""
    if a is None:
        return None
    assert a.flags['C_CONTIGUOUS'], 'oracle arrays must be C-contiguous'
    return a.ctypes.data


def d2_stencil(g):
    """The 3rd-order transport flux's curvature stencil per edge end (reference: src/insilica/gpu.py:86-108)."""
    E = g.E
    nbr, deg, dist = g.nbr.astype(np.int64), g.deg.astype(np.int64), g.dist
    idx = np.zeros((E, 2, 4), dtype=np.int32)
    wt = np.zeros((E, 2, 4))
    for side, (c, k) in enumerate(((g.eI.astype(np.int64), g.eKi.astype(np.int64)),
                                   (g.eJ.astype(np.int64), g.eKj.astype(np.int64)))):
        m = deg[c]
        j = nbr[c, k]
        d1 = dist[c, k]
        hexa = m == 6
        ka = np.where(hexa, (k + 3) % 6, (k + 2) % m)
        kb = np.where(hexa, (k + 3) % 6, (k + 3) % m)
        d2v = np.where(hexa, dist[c, (k + 3) % 6], 0.5 * (dist[c, ka] + dist[c, kb]))
        s = 2 / (d1 + d2v) * g.de * g.de / 12
        idx[:, side] = np.stack([c, j, nbr[c, ka], nbr[c, kb]], axis=1)
        wt[:, side, 0] = -s * (1 / d1 + 1 / d2v)
        wt[:, side, 1] = s / d1
        wt[:, side, 2] = 0.5 * s / d2v
        wt[:, side, 3] = 0.5 * s / d2v
    return idx.reshape(-1), wt.reshape(-1)


class Grid:
    """A mesh's tables as the oracle reads them (int32/int8/float64, C-contiguous)."""

    INT32 = ('eI', 'eJ', 'nbr', 'cellEdge', 'cellEdgeAsc', 'ev1', 'ev2', 'vEdge', 'vCell', 'eoeStart', 'eoe')
    INT8 = ('cellSign', 'vEdgeSign', 'vSlot')
    REALS = ('le', 'de', 'areaCell', 'areaV', 'fV', 'kite', 'w')

    def __init__(self, grid):
        self.host = grid
        self.N, self.E, self.NV = grid.N, grid.E, grid.NV
        self.arr = {}
        for k in self.INT32:
            self.arr[k] = np.ascontiguousarray(getattr(grid, k), dtype=np.int32)
        for k in self.INT8:
            self.arr[k] = np.ascontiguousarray(getattr(grid, k), dtype=np.int8)
        for k in self.REALS:
            self.arr[k] = np.ascontiguousarray(getattr(grid, k), dtype=np.float64)
        idx, wt = d2_stencil(grid)
        self.arr['d2idx'], self.arr['d2wt'] = np.ascontiguousarray(idx), np.ascontiguousarray(wt)
        s = OGGrid()
        s.N, s.E, s.NV, s.n_eoe = grid.N, grid.E, grid.NV, len(self.arr['eoe'])
        for k, v in self.arr.items():
            setattr(s, k, ctypes.cast(v.ctypes.data, {np.int32: _i32p, np.int8: _i8p, np.float64: _f64p}[v.dtype.type]))
        self.s = s
        self.ref = ctypes.byref(s)


class Layers:
    """The reference's Layers (src/insilica/layers.py:161-338) on the CPU oracle."""

    def __init__(self, og, layers=1, b=None, grav=9.80616, gprime=(), drag=0.0, stress=0.0, nu4=0.0, floor=1.0,
                 rotation=1.0, div_damp=0.1, air=None, vmix=0.0, thin=2.0, wall=None, thin_wind=0.0, shelf=0.0,
                 vmix_cd=0.0):
        self.og, self.L, self.air = og, layers, air
        c = OGLayersCfg()
        c.L, c.air = layers, 1 if air else 0
        c.grav, c.nu4, c.floor_, c.div_damp, c.thin, c.shelf = grav, nu4, floor, div_damp, thin, shelf
        c.vmix, c.vmix_cd, c.thin_wind = vmix, vmix_cd, thin_wind
        c.cp, c.kappa, c.rd, c.p0 = ((air or {}).get('cp', 1004.6), (air or {}).get('kappa', 287.04 / 1004.6),
                                     (air or {}).get('rd', 287.04), 1.0e5)
        gp = list(np.atleast_1d(np.asarray(gprime, dtype=float))) if layers > 1 else [0.0]
        for k, v in enumerate(gp):
            c.gp[k] = v
        for k, v in enumerate(np.broadcast_to(np.asarray(drag, dtype=float), (layers,))):
            c.drag[k] = v
        for k, v in enumerate(np.broadcast_to(np.asarray(stress, dtype=float), (layers,))):
            c.stress[k] = v
        self.fV = np.ascontiguousarray(og.arr['fV'] * rotation)
        self.b = np.zeros(og.N) if b is None else np.ascontiguousarray(b, dtype=float)
        self.wall = None if wall is None else np.ascontiguousarray(wall, dtype=float)
        c.fV = ctypes.cast(self.fV.ctypes.data, _f64p)
        c.b = ctypes.cast(self.b.ctypes.data, _f64p)
        c.wall = ctypes.cast(self.wall.ctypes.data, _f64p) if self.wall is not None else None
        self.cfg = c
        self.h_ = lib().og_layers_new(og.ref, ctypes.byref(c))
        self.nu4, self.nu4dt = nu4, 0.0
        self.cpa, self.kap, self.rda = c.cp, c.kappa, c.rd
        self.grav = grav

    def __del__(self):
        if getattr(self, 'h_', None):
            lib().og_layers_free(self.h_)
            self.h_ = None

    def _view(self, fn, n):
        return np.ctypeslib.as_array(ctypes.cast(fn(self.h_), _f64p), shape=(n,))

    @property
    def Gc(self):
        return self._view(lib().og_layers_Gc, self.og.N)

    @property
    def L2(self):
        return self._view(lib().og_layers_L2, self.L * self.og.E).reshape(self.L, self.og.E)

    @property
    def diss(self):
        return self._view(lib().og_layers_diss, self.og.E)

    def set_tides(self, tide, cbot, cd_bot):
        """The tides' hooks (reference: layers.py:57-58, set by tides.Tides): the potential per cell (N, read at each
        stage), the internal tide's drag per edge (E or None) and the bottom's quadratic drag coefficient. The arrays
        are held, not copied: the caller updates them in place."""
        self._tide, self._cbot = tide, cbot
        lib().og_layers_tides(self.h_, _p(tide), _p(cbot), float(cd_bot))

    def update_gfast(self, h, theta):
        lib().og_layers_gfast(self.h_, _p(h), _p(theta))

    def substeps(self, h, dt):
        return lib().og_layers_substeps(self.h_, _p(h), dt)

    def lap(self, u, out):
        lib().og_layers_lap(self.h_, _p(u), _p(out))

    def step(self, h, u, dt, nfast, Fm=None, theta=None):
        lib().og_layers_step(self.h_, _p(h), _p(u), dt, nfast, _p(Fm), _p(theta))
        if self.nu4:
            self.nu4dt = self.nu4 * dt

    def exchange(self, wx, h, u, M, psi, theta=None, seq=False):
        lib().og_layers_exchange(self.h_, _p(np.ascontiguousarray(wx, dtype=float)), _p(h), _p(u), _p(M), _p(psi),
                                 _p(theta), 1 if seq else 0, self.nu4dt if (self.air and theta is not None) else 0.0)


def mass_update(og, L, M0, Fm, M1):
    lib().og_mass_update(og.ref, L, _p(M0), _p(Fm), _p(M1))


def transport(og, L, M0, M1, Fm, psi, beta=0.25, order=3):
    """Transport.step_many on rows = psi.shape[0] stacked (tracer, layer) rows, in place."""
    rows = psi.shape[0] if psi.ndim == 2 else 1
    assert rows % L == 0
    lib().og_transport(og.ref, L, rows, beta, _p(M0), _p(M1), _p(Fm), _p(psi), order)


def swe_step(og, L, h, u, dt, grav=9.80616, b=None, drag=0.0, nu4=0.0, Fm=None):
    b = np.zeros(og.N) if b is None else np.ascontiguousarray(b, dtype=float)
    dr = np.ascontiguousarray(np.broadcast_to(np.asarray(drag, dtype=float), (L,)))
    lib().og_swe_step(og.ref, L, grav, _p(b), _p(dr), nu4, _p(h), _p(u), dt, _p(Fm))


def couple_stress(og, La, ha, ua, ho, uo, c, rho_w, dt, diss, tau_acc, wall=None, land_drag=1.0, cd=None):
    lib().og_couple_stress(og.ref, La, _p(ha), _p(ua), _p(ho), _p(uo), c, rho_w, dt, _p(diss), _p(tau_acc), _p(wall),
                           land_drag, _p(cd))


def land_column(og, K, land, T, cap, Tg, dz, c_soil, k_soil, dt):
    lib().og_land_column(og.ref, K, _p(land), _p(T), _p(cap), _p(Tg), _p(np.ascontiguousarray(dz, dtype=float)),
                         c_soil, k_soil, dt)


def ocn_cap(og, h, rho_cw, cap, land=None, c_land=2.0e6):
    lib().og_ocn_cap(og.ref, _p(h), rho_cw, _p(cap), _p(land), c_land)


def vmix_heat(og, L, h, diss, theta, grav=9.80616, cp=1004.6, kappa=287.04 / 1004.6, p0=1.0e5):
    lib().og_vmix_heat(og.ref, L, _p(h), _p(diss), grav, cp, kappa, p0, _p(theta))


def lay_vmix(og, L, h, c, dt, u, diss, cd=0.0):
    """The interfacial stress on its own (reference: kernels/layers.cu:575-593), in place on u (L, E); diss (E)."""
    lib().og_lay_vmix(og.ref, L, _p(h), c, dt, _p(u), _p(diss), cd)


def dsum(x, y=None):
    x = np.ascontiguousarray(x, dtype=float).ravel()
    y = None if y is None else np.ascontiguousarray(y, dtype=float).ravel()
    return lib().og_dsum(_p(x), _p(y), x.size)


def nfast_for(layers, h, dt):
    """World's rule (reference: src/insilica/world.py:228): ceil(1.2 * substeps)."""
    return int(math.ceil(1.2 * layers.substeps(h, dt)))
