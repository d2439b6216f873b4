"""Tracer transport: the reference's `Transport` (src/insilica/transport.py:18-71) over libinsilica_hip.so.

Moves the mass-weighted tracer M psi with each edge's mass flux Fm (per step, eI -> eJ positive). order 1: upwind.
order 3: 3rd-order upwind-biased fluxes (beta = 0.25) in RK3 stages with Zalesak flux-corrected transport against
the upwind solution, so tracers stay within their local bounds. Every tracer of a world goes through the same launches
as rows (tracer, layer). A world that keeps its tracers stacked in one (rows, N) tensor passes it without copies.
"""
import ctypes

import torch

from . import _hip

BETA = 0.25


class Transport:
    def __init__(self, dg, layers=1, beta=BETA, tracers=1):
        self.dg, self.L, self.beta, self.K = dg, layers, beta, tracers
        N, E, R = dg.N, dg.E, tracers * layers
        z = dg.zeros
        self.psi = z(R, N)                           # the staging rows for tracers passed as separate arrays
        self.td, self.p1, self.p2, self.Rp, self.Rm = (z(R, N) for _ in range(5))
        self.Hf, self.A = z(R, E), z(R, E)
        self.pf = z(R, N, dtype=torch.float32) if dg.mixed else None
        self.pf2 = z(R, N, dtype=torch.float32) if dg.mixed else None
        s = _hip.IswTransport()
        s.L, s.rows_max, s.prec, s.beta = layers, R, dg.prec, beta
        for k in ('td', 'p1', 'p2', 'Rp', 'Rm', 'Hf', 'A'):
            setattr(s, k, ctypes.cast(getattr(self, k).data_ptr(), _hip._f64p))
        s.pf = ctypes.cast(self.pf.data_ptr(), _hip._f32p) if self.pf is not None else None
        s.pf2 = ctypes.cast(self.pf2.data_ptr(), _hip._f32p) if self.pf2 is not None else None
        self.s = s
        self.ref = ctypes.byref(s)

    def mass_after(self, M0, Fm, M1):
        """M1 = M0 moved by Fm (what the cells hold after the step)."""
        _hip.call('isw_mass_update', self.dg.ref, self.L, _hip.ptr(M0), _hip.ptr(Fm), _hip.ptr(M1), _hip.stream())

    def step(self, M0, M1, Fm, psi, order=3):
        return self.step_many(M0, M1, Fm, [psi], order)[0]

    def step_rows(self, M0, M1, Fm, rows, order=3):
        """Move the stacked rows (R, N) (row y by layer y % L's fluxes) in place, in one call."""
        assert rows.is_contiguous() and rows.shape[0] % self.L == 0
        _hip.call('isw_transport_step', self.dg.ref, self.ref, _hip.ptr(M0), _hip.ptr(M1), _hip.ptr(Fm),
                  _hip.ptr(rows), int(rows.shape[0]), int(order), _hip.stream())
        return rows

    def step_many(self, M0, M1, Fm, tracers, order=3):
        """Move each (L, N) tracer in `tracers` by the same mass fluxes, in place, in one set of launches."""
        n, L = len(tracers), self.L
        assert n <= self.K
        rows = self.psi[:n * L]
        for j, t in enumerate(tracers):
            rows[j * L:(j + 1) * L].copy_(t.reshape(L, -1))
        self.step_rows(M0, M1, Fm, rows, order)
        for j, t in enumerate(tracers):
            t.copy_(rows[j * L:(j + 1) * L].reshape(t.shape))
        return tracers
