"""Stacked-layer shallow water with a split-explicit fast mode on the TRiSK C-grid: the reference's `Layers`
(src/insilica/layers.py:161-379) over libinsilica_hip.so.

L layers (k = 0 on top), thickness h (L, N) and normal wind u (L, E). Boussinesq layers feel the Montgomery potential
g eta + M'_k; air layers (h the mass per area, a potential temperature theta per layer) the Montgomery potential of
isentropic layers. Time-split RK3 (Wicker & Skamarock 2002): per stage the slow tendencies are held while the fast mode
(the ocean's free surface and barotropic wind, or every air layer's mass and wind) is sub-stepped forward-backward
from t^n with divergence damping; a thickness floor limits each layer's outflow. `step` is one isw_layers_step call:
all of it runs on the GPU and allocates nothing, so a caller may capture it in a HIP graph.
"""
import ctypes
import math

import numpy as np
import torch

from . import _hip
from .constants import CP_DRY, GRAV, KAPPA, P0, R_DRY

LMAX = 8
FAST_COURANT = 0.5          # fast sub-step: tau sqrt(gH) / d_min <= this
DIV_DAMP = 0.1              # the fast mode's divergence damping


class Layers:
    def __init__(self, dg, layers=1, b=None, grav=GRAV, gprime=(), drag=0.0, stress=0.0, nu4=0.0, floor=1.0,
                 rotation=1.0, div_damp=DIV_DAMP, air=None, vmix=0.0, thin=2.0, wall=None, thin_wind=0.0, shelf=0.0,
                 vmix_cd=0.0, shelf_bed=False, batch=None):
        """Same arguments as the reference's Layers (layers.py:161-218). air: None for Boussinesq layers, or a dict
        (cp, kappa, rd) for air layers. wall: None or an (E,) mask, 0 on edges no water crosses. shelf_bed is not
        supported (the reference's squeezed-layer flags, off by default)."""
        assert 1 <= layers <= LMAX and np.atleast_1d(np.asarray(gprime, dtype=float)).size == layers - 1
        if shelf_bed:
            raise NotImplementedError('shelf_bed (squeezed-layer flags) is not part of this build')
        assert wall is None or not air, 'walls are for Boussinesq layers'
        self.dg, self.L = dg, layers
        self.grav, self.nu4, self.floor, self.div_damp, self.thin = grav, nu4, floor, div_damp, thin
        self.thin_wind, self.shelf, self.vmix, self.vmix_cd = thin_wind, shelf, vmix, vmix_cd
        self.air = air
        self.cpa = (air or {}).get('cp', CP_DRY)
        self.kap = (air or {}).get('kappa', KAPPA)
        self.rda = (air or {}).get('rd', R_DRY)
        N, E, NV, L = dg.N, dg.E, dg.NV, layers
        z = dg.zeros
        self.b = z(N) if b is None else torch.as_tensor(b, dtype=torch.float64, device=dg.dev).clone()
        self.wall = None if wall is None else torch.as_tensor(np.asarray(wall, dtype=float), device=dg.dev).clone()
        self.fV = dg.fV * rotation
        self.Gc = torch.full((N,), float(grav), dtype=torch.float64, device=dg.dev)
        shapes = dict(B=(L, N), q=(L, NV), Tp=(L, E), F=(L, E), scale=(L, N), G=(E,), He=(E,), ubs=(E,), ub0=(E,),
                      ub=(E,), acc=(E,), eta0=(N,), eta=(N,), eta_prev=(N,), h1=(L, N), h2=(L, N), u1=(L, E),
                      u2=(L, E), D1=(L, N), Z1=(L, NV), L1=(L, E), L2=(L, E), hold=(L, N), hedge=(L, E), diss=(E,),
                      pibar=(L, N), S=(L, N), muf=(L, N), mup=(L, N), phi=(L, N), uf=(L, E), accL=(L, E), ufb=(L, E), ubb=(E,),
                      ubc=(E,))
        for k, shp in shapes.items():
            setattr(self, k, z(*shp))
        self.pk = z(3 * L * E) if not dg.mixed else z(1)
        self.pkf = z(3 * L * E, dtype=torch.float32) if dg.mixed else None
        s = _hip.IswLayers()
        s.L, s.air, s.prec = L, 1 if air else 0, dg.prec
        s.grav, s.cp, s.kappa, s.p0, s.rd = grav, self.cpa, self.kap, P0, self.rda
        s.nu4, s.floor_, s.div_damp, s.thin, s.shelf = nu4, floor, div_damp, thin, shelf
        s.vmix, s.vmix_cd, s.thin_wind = vmix, vmix_cd, thin_wind
        gp = list(np.atleast_1d(np.asarray(gprime, dtype=float))) if layers > 1 else []
        self.gprime = gp
        for k, v in enumerate(gp):
            s.gp[k] = v
        self.drag = np.broadcast_to(np.asarray(drag, dtype=float), (L,)).copy()
        self.stress = np.broadcast_to(np.asarray(stress, dtype=float), (L,)).copy()
        for k in range(L):
            s.drag[k], s.stress[k] = self.drag[k], self.stress[k]
        f64 = lambda t: ctypes.cast(t.data_ptr(), _hip._f64p) if t is not None else None
        s.fV, s.b, s.wall, s.Gc = f64(self.fV), f64(self.b), f64(self.wall), f64(self.Gc)
        for k in _hip.LAYER_SCRATCH:
            setattr(s, k, f64(getattr(self, k)))
        s.pkf = ctypes.cast(self.pkf.data_ptr(), _hip._f32p) if self.pkf is not None else None
        # (the air's fast sub-steps' coefficients sign * l_e * h_e per cell slot, written with each stage's h_e:
        # FP64 in the plain form, FP32 in the mixed form)
        self.cfc = z(L * N * 6) if air and not dg.mixed else None
        self.cfcf = z(L * N * 6, dtype=torch.float32) if air and dg.mixed else None
        s.cfc = f64(self.cfc)
        s.cfcf = ctypes.cast(self.cfcf.data_ptr(), _hip._f32p) if self.cfcf is not None else None
        self.s = s
        self.ref = ctypes.byref(s)
        self.batch = None
        if batch is not None:
            self._set_batch(batch)
        self.dmin = float(np.min(dg.host.de))
        self.nu4dt = 0.0
        self._tide = self._cbot = None
        self._cd_bot = 0.0

    # tides (reference: layers.py:57-58, set by tides.Tides): the potential per cell, the internal tide's drag per
    # edge and the bottom's quadratic drag coefficient, read by the step's kernels through the struct
    @property
    def tide(self):
        return self._tide

    @tide.setter
    def tide(self, t):
        self._tide = t
        self.s.tide = ctypes.cast(t.data_ptr(), _hip._f64p) if t is not None else None

    @property
    def cbot(self):
        return self._cbot

    @cbot.setter
    def cbot(self, t):
        self._cbot = t
        self.s.cbot = ctypes.cast(t.data_ptr(), _hip._f64p) if t is not None else None

    @property
    def cd_bot(self):
        return self._cd_bot

    @cd_bot.setter
    def cd_bot(self, v):
        self._cd_bot = float(v)
        self.s.cd_bot = float(v)

    def _set_batch(self, b):
        """A batch of worlds on a union mesh (insilica.batch): per-world gravity, del^4 coefficient and fast
        sub-steps through the isw_batch table; each cell's fast-mode coefficient starts at its world's gravity."""
        self.batch = b
        t = _hip.IswBatch()
        t.nworlds = b['nworlds']
        for k in range(3):
            t.nmax[k] = int(b['nmax'][k])
        for k in ('wc', 'we', 'nsub'):
            setattr(t, k, ctypes.cast(b[k].data_ptr(), _hip._i32p))
        for k in ('grav', 'nu4', 'tau', 'inv_n'):
            setattr(t, k, ctypes.cast(b[k].data_ptr(), _hip._f64p))
        self._batch_struct = t
        self.s.batch = ctypes.pointer(t)
        self.s.nu4 = float(torch.max(b['nu4']))
        self.nu4 = self.s.nu4
        self.Gc.copy_(b['grav'][b['wc'].long()])

    def mask(self, a, rows):
        """Zero the wall edges of `rows` rows of the edge field a (no-op without walls)."""
        if self.wall is not None:
            a.view(rows, -1).mul_((self.wall != 0).to(a.dtype))

    def substeps(self, h, dt):
        """Fast sub-steps per full step for the current deepest column (reference: layers.py:225-228)."""
        c2 = float(torch.max(torch.sum(h, dim=0) * self.Gc))
        return max(1, math.ceil(dt * math.sqrt(c2) / (FAST_COURANT * self.dmin)))

    def update_gfast(self, h, theta):
        """For air: each cell's fast-mode coefficient R T_s / mu plus the interfaces' stiffness."""
        if self.air:
            _hip.call('isw_layers_gfast', self.dg.ref, self.ref, _hip.ptr(h), _hip.ptr(theta), _hip.stream())

    def lap_max(self, iters=30):
        """Largest |eigenvalue| of the edge vector Laplacian by power iteration from a fixed seed (1/m^2)."""
        v = torch.as_tensor(np.random.default_rng(0).standard_normal((self.L, self.dg.E)), device=self.dg.dev)
        lam = 0.0
        for _ in range(iters):
            self._lap(v, self.L1)
            lam = float(torch.linalg.norm(self.L1[0]) / torch.linalg.norm(v[0]))
            v = self.L1 / torch.linalg.norm(self.L1[0])
        return lam

    def _lap(self, u, out):
        _hip.call('isw_layers_lap', self.dg.ref, self.ref, _hip.ptr(u), _hip.ptr(out), _hip.stream())

    def step(self, h, u, dt, nfast, Fm=None, theta=None):
        """One time-split RK3 step in place on h (L, N), u (L, E) with nfast fast sub-steps per full step (stage s
        takes ceil(c_s nfast)). Fm, if given, receives each layer's final-stage volume flux l dt F."""
        _hip.call('isw_layers_step', self.dg.ref, self.ref, _hip.ptr(h), _hip.ptr(u), _hip.ptr(theta), _hip.ptr(Fm),
                  float(dt), int(nfast), _hip.stream())
        if self.nu4:
            self.nu4dt = self.nu4 * dt

    def exchange(self, wx, h, u, M, psi, theta=None, seq=False):
        """Move thickness wx (L-1, N) between neighbouring layers with its tracer and momentum, conservatively."""
        heat = bool(self.air) and theta is not None
        _hip.call('isw_layers_exchange', self.dg.ref, self.ref, _hip.ptr(wx.contiguous()), _hip.ptr(h), _hip.ptr(u),
                  _hip.ptr(M), _hip.ptr(psi), _hip.ptr(theta if heat else None), 1 if seq else 0,
                  float(self.nu4dt if heat else 0.0), _hip.stream())

    def vmix_heat(self, h, diss, theta):
        _hip.call('isw_vmix_heat', self.dg.ref, self.ref, _hip.ptr(h), _hip.ptr(diss), _hip.ptr(theta), _hip.stream())

    def energy(self, h, u, theta=None):
        """Total energy per unit density: kinetic + potential (air: + enthalpy and the column's potential energy),
        as the reference's Layers.energy (layers.py:360-379), evaluated on the host in float64."""
        gh = self.dg.host
        hh, uu, b = h.cpu().numpy(), u.cpu().numpy(), self.b.cpu().numpy()
        ce = np.where(gh.cellEdge >= 0, gh.cellEdge, 0)
        ok = gh.cellEdge >= 0
        kin = sum(np.sum(hh[k] * np.sum(np.where(ok, gh.le[ce] * gh.de[ce] * 0.25 * uu[k][ce] ** 2, 0.0), axis=1))
                  for k in range(self.L))
        if self.air:
            th = theta.cpu().numpy()
            p = np.cumsum(self.grav * hh, axis=0)
            pt = np.vstack([np.zeros((1, hh.shape[1])), p[:-1]])
            pib = self.cpa * (p ** (self.kap + 1) - pt ** (self.kap + 1)) / ((self.kap + 1) * (p - pt) * P0 ** self.kap)
            return float(kin + np.sum(gh.areaCell * (np.sum(hh * pib * th, axis=0) + self.grav * b * hh.sum(0))))
        z = b + np.cumsum(hh[::-1], axis=0)[::-1]
        pot = np.sum(gh.areaCell * self.grav * (z[0] ** 2 - b ** 2) / 2)
        for k in range(1, self.L):
            pot += np.sum(gh.areaCell * self.gprime[k - 1] * (z[k] ** 2 - b ** 2) / 2)
        return float(kin + pot)
