"""The gravity waves' drag on the GPU (isw_oro_drag, isw_wave_drag through `OroDrag` and `WaveDrag`) against the
oracle, on seeded coupled-stack states with synthetic sub-grid orography, and the reference's ledgers on the device.

Bar: FP64 parity <= 1e-12 relative per field (the cell passes' pow, log, atan2, asin and exp come from the device
math library on one side and glibc on the other, which may differ by an ulp); the dissipation, a difference of
squares of winds, to 1e-13 of the largest edge kinetic energy; the ledgers of the reference's
tests/test_world.py:231-244 -- the kinetic energy removed is the heat booked, the waves keep each column's
momentum -- to 1e-13 at level 6.
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def rel(a, b):
    b = np.asarray(b)
    return float(np.max(np.abs(np.asarray(a) - b)) / max(float(np.max(np.abs(b))), 1e-300))


def _setup(grids, n, seed):
    import oracle
    from insilica import cases
    from insilica.atmosphere import DEFAULTS
    from insilica.gpu import DeviceGrid
    g = grids(n)
    dg, og = DeviceGrid(g, 'double'), oracle.Grid(g)
    built = cases.coupled_stack(g, seed=seed)
    h, th, u = built['h'], built['theta'], 4 * built['u']
    sso = cases.synthetic_sso(g, seed=seed, land=built['land'])
    p = dict(DEFAULTS, oro=1.0, gw=1.0)
    return g, dg, og, h, th, u, sso, p


def ke_scale(g, h, u):
    """The largest edge kinetic energy per area, sum_k h_e u^2 / 2: the scale of the dissipation's rounding (it is
    a difference of squares of winds that match to an ulp)."""
    he = 0.5 * (h[:, g.eI] + h[:, g.eJ])
    return float(np.max(np.sum(0.5 * he * u * u, 0)))


def _dev(dg, a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dg.dev)


@pytest.mark.parametrize('n', [4, 6])
def test_oro_drag_matches_oracle(grids, n):
    from oracle import drag as od
    from insilica.orodrag import OroDrag
    print(f'case: oro drag n={n}', flush=True)
    g, dg, og, h, th, u, sso, p = _setup(grids, n, seed=5)
    dt = 900.0
    oro = OroDrag(dg, 3, sso, p, dt)
    H, TH, U, D = _dev(dg, h), _dev(dg, th), _dev(dg, u), _dev(dg, np.zeros(g.E))
    oro.apply(H, TH, U, dt, D)
    torch.cuda.synchronize()
    par = np.array([getattr(oro.par, k) for k, _ in oro.par._fields_])
    u1, d1, beta, F = od.oro_drag(og, h, th, u, oro.op.cpu().numpy(), oro.nrm.cpu().numpy(),
                                  oro.sso.cpu().numpy(), par, dt, np.zeros(g.E))
    assert np.count_nonzero(beta) > 0.01 * beta.size and np.count_nonzero(F) > 0.01 * F.size
    assert rel(oro.beta.cpu().numpy(), beta) <= 1e-12
    assert rel(oro.F.cpu().numpy(), F) <= 1e-12
    assert rel(U.cpu().numpy(), u1) <= 1e-12 and rel(U.cpu().numpy() - u, u1 - u) <= 1e-9
    assert np.max(np.abs(D.cpu().numpy() - d1)) <= 1e-13 * ke_scale(g, h, u)


@pytest.mark.parametrize('n,conv', [(4, False), (4, True), (6, False)])
def test_wave_drag_matches_oracle(grids, n, conv):
    from oracle import drag as od
    from insilica.orodrag import WaveDrag
    print(f'case: wave drag n={n} conv={conv}', flush=True)
    g, dg, og, h, th, u, sso, p = _setup(grids, n, seed=6)
    dt = 900.0
    if conv:
        p['gw_conv'] = 0.6
    gw = WaveDrag(dg, 3, p, dt)
    cv = np.random.default_rng(7).uniform(0, 1e-4, g.N) if conv else None
    if conv:
        gw.conv = _dev(dg, cv)
    H, TH, U, D = _dev(dg, h), _dev(dg, th), _dev(dg, u), _dev(dg, np.zeros(g.E))
    gw.apply(H, TH, U, dt, D)
    torch.cuda.synchronize()
    par = np.array([getattr(gw.par, k) for k, _ in gw.par._fields_])
    u1, d1, F = od.wave_drag(og, h, th, u, gw.op.cpu().numpy(), gw.nrm.cpu().numpy(), par, dt, np.zeros(g.E),
                             conv=cv)
    assert np.count_nonzero(F) > 0.1 * F.size // 2
    assert rel(gw.F.cpu().numpy(), F) <= 1e-12
    # (the winds to 1e-12 of the field; their small increments, u1 - u0, to 1e-9 of the largest: an ulp of the
    # forces moves the rounding of u0 + dt fn / m by an ulp of u, which is large against the increment)
    assert rel(U.cpu().numpy(), u1) <= 1e-12 and rel(U.cpu().numpy() - u, u1 - u) <= 1e-9
    assert np.max(np.abs(D.cpu().numpy() - d1)) <= 1e-13 * ke_scale(g, h, u)
    if conv:
        assert float(gw.cmean[0]) == od.gw_mean(og, cv)             # (the same fixed-order sum)


def test_ledgers_level6(grids):
    """At level 6: the kinetic energy each drag removes is the heat it books (<= 1e-13 of the kinetic energy), the
    waves keep each edge's column momentum (<= 1e-13 of the gross)."""
    from insilica.orodrag import OroDrag, WaveDrag
    g, dg, og, h, th, u, sso, p = _setup(grids, 6, seed=8)
    dt = 900.0
    H, TH = _dev(dg, h), _dev(dg, th)
    lede = _dev(dg, g.le * g.de)
    eI, eJ = dg.eI.long(), dg.eJ.long()
    he = 0.5 * (H[:, eI] + H[:, eJ])
    ke = lambda U: float(torch.sum(lede * torch.sum(0.5 * he * U * U, 0)))
    for drag in (OroDrag(dg, 3, sso, p, dt), WaveDrag(dg, 3, p, dt)):
        U, D = _dev(dg, u), _dev(dg, np.zeros(g.E))
        k0, m0 = ke(U), torch.sum(he * U, 0)
        drag.apply(H, TH, U, dt, D)
        removed, booked = k0 - ke(U), float(torch.sum(lede * D))
        assert abs(removed) > 1e-9 * k0
        assert abs(removed - booked) <= 1e-13 * k0, (type(drag).__name__, removed, booked)
        if isinstance(drag, WaveDrag):
            dm = torch.sum(he * U, 0) - m0
            gross = float(torch.sum(lede * torch.sum(he * torch.abs(U), 0)))
            assert float(torch.sum(lede * torch.abs(dm))) <= 1e-13 * gross


def test_world_with_drags_matches_oracle_world():
    """A coupled-stack world with both drags (the synthetic orography over its land) in the FP64 form against the
    oracle's world, tick for tick (the drags after the wind stress, before its heat, as the reference's tick)."""
    from insilica.world import World
    from oracle.world import OracleWorld
    print('case: world with drags n=4', flush=True)
    w = World(4, 'coupled_stack', precision='double', params={'physics': {'oro': 1.0, 'gw': 1.0}, 'sso': 'synthetic'})
    assert w.oro is not None and w.gw is not None
    o = OracleWorld(w)
    w.step(4, graph=False)
    o.step(4)
    worst = 0.0
    st = w.state_arrays()
    for k, ref in o.state().items():
        if k == 'sea.cap' or k not in st:
            continue
        got = st[k].detach().cpu().numpy()
        worst = max(worst, rel(got, ref))
    assert worst <= 1e-10, worst
    # the graph replays the same ticks as direct launches
    w2 = World(4, 'coupled_stack', precision='double', params={'physics': {'oro': 1.0, 'gw': 1.0}, 'sso': 'synthetic'})
    w2.step(4, graph=True)
    torch.cuda.synchronize()
    for k, v in w.state_arrays().items():
        assert torch.equal(v, w2.state_arrays()[k]), k
