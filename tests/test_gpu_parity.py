"""Parity of the HIP library (plain FP64 form) with the CPU oracle, through the C ABI, on seeded inputs.

Bar: the FP64 form matches the oracle to <= 1e-12 relative (max |diff| / max |oracle|) after one step; longer runs
are held to <= 1e-10 (differences of one rounding in `pow`/`exp` between the device math library and glibc grow slowly
with the flow). Sums built on the reference's fixed-order tree are bit-identical.
"""
import math

import numpy as np
import pytest
import torch

from conftest import relmax

pytestmark = pytest.mark.gpu

TOL1, TOLN = 1e-12, 1e-10


def _np(t):
    return t.detach().cpu().numpy().copy()


def _solid_body(g, alpha=math.pi / 2 - 0.05, period_days=12):
    R = 6.371e6
    u0 = 2 * math.pi * R / (period_days * 86400)
    om = np.array([-math.sin(alpha) * u0 / R, 0, math.cos(alpha) * u0 / R])
    return np.sum(np.cross(om, g.xm * R) * g.ne, axis=1)


@pytest.fixture(scope='module')
def g4(grids):
    return grids(4)


def _pair_layers(g, dg, og, **kw):
    from insilica.layers import Layers
    import oracle
    lay = Layers(dg, **kw)
    okw = dict(kw)
    if okw.get('wall') is not None:
        okw['wall'] = np.asarray(okw['wall'], dtype=float)
    olay = oracle.Layers(og, **okw)
    return lay, olay


def test_layers_ocean_fp64(g4):
    """Two Boussinesq layers (Phillips jet, perturbed) with walls, bottom stress, interfacial stress and del^4."""
    import oracle
    from insilica import cases
    from insilica.gpu import DeviceGrid
    from insilica.ocean import visc4
    print('case: ocean layers n=4', flush=True)
    g = g4
    dg, og = DeviceGrid(g, 'double'), oracle.Grid(g)
    h, b, u, gp = cases.phillips(g, 3.0)
    rng = np.random.default_rng(1)
    h = h + rng.uniform(-20, 20, h.shape)
    land = (rng.random(g.N) < 0.1).astype(float)
    wall = np.ones(g.E)
    wall[(land[g.eI] > 0.5) | (land[g.eJ] > 0.5)] = 0.0
    kw = dict(layers=2, b=b, gprime=[gp], stress=[0.0, 2.5e-4], nu4=visc4(g, 5.0), floor=1.0, vmix=5e-4, wall=wall)
    lay, olay = _pair_layers(g, dg, og, **kw)
    dt = 1200.0
    nf = int(math.ceil(1.2 * lay.substeps(dg.asarray(h), dt)))
    assert nf == int(math.ceil(1.2 * olay.substeps(np.ascontiguousarray(h), dt)))
    hd, ud = dg.asarray(h), dg.asarray(u * wall[None])
    ho, uo = np.ascontiguousarray(h), np.ascontiguousarray(u * wall[None])
    Fd, Fo = dg.zeros(2, g.E), np.zeros((2, g.E))
    for it in range(12):
        lay.step(hd, ud, dt, nf, Fd)
        olay.step(ho, uo, dt, nf, Fo)
        torch.cuda.synchronize()
        tol = TOL1 if it == 0 else TOLN
        for name, a, r in (('h', hd, ho), ('u', ud, uo), ('Fm', Fd, Fo), ('diss', lay.diss, olay.diss)):
            d = relmax(_np(a), r)
            assert d <= tol, f'step {it}: {name} differs by {d:.2e}'
    assert np.all(_np(ud)[:, wall == 0] == 0)


def test_layers_air_fp64(g4):
    """Three air layers (the perturbed aquaplanet start), del^4, fast sub-steps from the column's stiffness."""
    import oracle
    from insilica import cases
    from insilica.constants import CP_DRY, KAPPA, R_DRY
    from insilica.gpu import DeviceGrid
    from insilica.ocean import visc4
    print('case: air layers n=4', flush=True)
    g = g4
    dg, og = DeviceGrid(g, 'double'), oracle.Grid(g)
    st = cases.coupled_stack(g, seed=3)
    air = dict(cp=CP_DRY, kappa=KAPPA, rd=R_DRY)
    kw = dict(layers=3, gprime=[0.0, 0.0], nu4=visc4(g, 15.0), floor=10.0, air=air, thin_wind=100.0,
              stress=[0.0, 0.0, 1e-2])
    lay, olay = _pair_layers(g, dg, og, **kw)
    hd, ud, td = dg.asarray(st['h']), dg.asarray(st['u']), dg.asarray(st['theta'])
    ho, uo, to = (np.ascontiguousarray(st[k]) for k in ('h', 'u', 'theta'))
    lay.update_gfast(hd, td)
    olay.update_gfast(ho, to)
    torch.cuda.synchronize()
    assert relmax(_np(lay.Gc), olay.Gc) <= TOL1
    dt = 1800.0
    nf = 15
    Fd, Fo = dg.zeros(3, g.E), np.zeros((3, g.E))
    for it in range(8):
        lay.step(hd, ud, dt, nf, Fd, theta=td)
        olay.step(ho, uo, dt, nf, Fo, theta=to)
        torch.cuda.synchronize()
        tol = TOL1 if it == 0 else TOLN
        for name, a, r in (('h', hd, ho), ('u', ud, uo), ('Fm', Fd, Fo), ('theta', td, to)):
            d = relmax(_np(a), r)
            assert d <= tol, f'step {it}: {name} differs by {d:.2e}'


@pytest.mark.parametrize('air', [True, False])
def test_layers_one_substep_fp64(g4, air):
    """nfast = 1: the last stage's single fast sub-step writes its output over its own input wind (u0 == un)."""
    import oracle
    from insilica import cases
    from insilica.constants import CP_DRY, KAPPA, R_DRY
    from insilica.gpu import DeviceGrid
    print(f'case: one sub-step, air={air}', flush=True)
    g = g4
    dg, og = DeviceGrid(g, 'double'), oracle.Grid(g)
    if air:
        st = cases.coupled_stack(g, seed=6)
        kw = dict(layers=3, gprime=[0.0, 0.0], floor=10.0, air=dict(cp=CP_DRY, kappa=KAPPA, rd=R_DRY))
        h, u, th = st['h'], st['u'], st['theta']
    else:
        h, b, u, gp = cases.phillips(g, 1.0)
        kw = dict(layers=2, b=b, gprime=[gp], floor=1.0)
        th = None
    lay, olay = _pair_layers(g, dg, og, **kw)
    hd, ud = dg.asarray(h), dg.asarray(u)
    ho, uo = np.ascontiguousarray(h), np.ascontiguousarray(u)
    td, to = (dg.asarray(th), np.ascontiguousarray(th)) if air else (None, None)
    if air:
        lay.update_gfast(hd, td)
        olay.update_gfast(ho, to)
    for it in range(3):
        lay.step(hd, ud, 60.0, 1, theta=td)
        olay.step(ho, uo, 60.0, 1, theta=to)
    torch.cuda.synchronize()
    assert relmax(_np(hd), ho) <= TOLN and relmax(_np(ud), uo) <= TOLN


def test_exchange_fp64(g4):
    """Layers.exchange (both orders) on three air layers with theta: h, M, psi, u, theta."""
    import oracle
    from insilica import cases
    from insilica.constants import CP_DRY, KAPPA, R_DRY
    from insilica.gpu import DeviceGrid
    print('case: exchange n=4', flush=True)
    g = g4
    dg, og = DeviceGrid(g, 'double'), oracle.Grid(g)
    st = cases.coupled_stack(g, seed=4)
    air = dict(cp=CP_DRY, kappa=KAPPA, rd=R_DRY)
    lay, olay = _pair_layers(g, dg, og, layers=3, gprime=[0.0, 0.0], floor=10.0, air=air, thin_wind=100.0)
    rng = np.random.default_rng(5)
    h = st['h']
    wx = np.stack([rng.uniform(-0.3, 0.3, g.N) * np.minimum(h[k], h[k + 1]) for k in range(2)])
    M = h * g.areaCell[None]
    psi = st['q'].copy()
    for seq in (False, True):
        d = dict(h=dg.asarray(h), u=dg.asarray(st['u']), M=dg.asarray(M), psi=dg.asarray(psi), th=dg.asarray(st['theta']))
        o = dict(h=h.copy(), u=st['u'].copy(), M=M.copy(), psi=psi.copy(), th=st['theta'].copy())
        lay.exchange(dg.asarray(wx), d['h'], d['u'], d['M'], d['psi'], theta=d['th'], seq=seq)
        olay.exchange(wx, o['h'], o['u'], o['M'], o['psi'], theta=o['th'], seq=seq)
        torch.cuda.synchronize()
        for k in d:
            assert relmax(_np(d[k]), o[k]) <= TOL1, f'seq={seq}: {k}'


def test_transport_fp64(g4):
    """FCT transport of 2 tracers x 2 layers by solid-body fluxes, and upwind, against the oracle."""
    import oracle
    from insilica import cases
    from insilica.gpu import DeviceGrid
    from insilica.transport import Transport
    print('case: transport n=4', flush=True)
    g = g4
    dg, og = DeviceGrid(g, 'double'), oracle.Grid(g)
    steps, dt = 40, 12 * 86400 / 576
    Fm1 = _solid_body(g) * g.le * dt
    Fm = np.stack([Fm1, 0.5 * Fm1])
    bell = cases.cosine_bell(g)
    rows = np.stack([bell, bell[::-1].copy(), 1.0 + 0.0 * bell, bell ** 0.5])
    M = np.stack([g.areaCell, 2 * g.areaCell])
    for order in (3, 1):
        tr = Transport(dg, layers=2, tracers=2)
        Md, M1d, Fd, rd = dg.asarray(M), dg.zeros(2, g.N), dg.asarray(Fm), dg.asarray(rows)
        Mo, M1o, ro = M.copy(), np.zeros((2, g.N)), rows.copy()
        for it in range(steps):
            tr.mass_after(Md, Fd, M1d)
            tr.step_rows(Md, M1d, Fd, rd, order)
            Md, M1d = M1d, Md
            oracle.mass_update(og, 2, Mo, Fm, M1o)
            oracle.transport(og, 2, Mo, M1o, Fm, ro, order=order)
            Mo, M1o = M1o, Mo
        torch.cuda.synchronize()
        assert relmax(_np(Md), Mo) <= TOL1
        assert relmax(_np(rd), ro) <= TOLN, f'order {order}'


def test_swe_fp64(g4):
    """The one-layer SWE solver on Williamson 5 (mountain, drag, del^4), 20 steps."""
    import oracle
    from insilica import cases
    from insilica.gpu import DeviceGrid
    from insilica.swe import SWE
    print('case: swe n=4', flush=True)
    g = g4
    dg, og = DeviceGrid(g, 'double'), oracle.Grid(g)
    h, b, u = cases.williamson(g, 'w5')
    nu4 = 1e14
    sw = SWE(dg, 1, b=b, drag=1e-6, nu4=nu4)
    hd, ud, Fd = dg.asarray(h[None]), dg.asarray(u[None]), dg.zeros(1, g.E)
    ho, uo, Fo = h[None].copy(), u[None].copy(), np.zeros((1, g.E))
    for it in range(20):
        sw.step(hd, ud, 450.0, Fd)
        oracle.swe_step(og, 1, ho, uo, 450.0, b=b, drag=1e-6, nu4=nu4, Fm=Fo)
        torch.cuda.synchronize()
        tol = TOL1 if it == 0 else TOLN
        for name, a, r in (('h', hd, ho), ('u', ud, uo), ('Fm', Fd, Fo)):
            assert relmax(_np(a), r) <= tol, f'step {it}: {name}'


def test_couple_land_cap_fp64(g4):
    """couple_stress (with walls), land_column (16 soil layers), ocn_cap on seeded inputs."""
    import oracle
    from insilica import _hip
    from insilica.gpu import DeviceGrid
    print('case: couple/land/cap n=4', flush=True)
    g = g4
    dg, og = DeviceGrid(g, 'double'), oracle.Grid(g)
    rng = np.random.default_rng(7)
    ha = rng.uniform(200, 8000, (3, g.N))
    ua = rng.uniform(-20, 20, (3, g.E))
    ho = rng.uniform(10, 200, g.N)
    uo = rng.uniform(-1, 1, g.E)
    wall = (rng.random(g.E) > 0.1).astype(float)
    cd = rng.uniform(1e-3, 3e-2, g.N)
    # the device inputs are held in named tensors until the kernels have run: a temporary's memory may be reused
    had, hod, walld, cdd = dg.asarray(ha), dg.asarray(ho), dg.asarray(wall), dg.asarray(cd)
    for use_cd in (False, True):
        d = [dg.asarray(x) for x in (ua, uo, np.zeros(g.E), np.ones(g.E))]
        o = [x.copy() for x in (ua, uo, np.zeros(g.E), np.ones(g.E))]
        _hip.call('isw_couple_stress', dg.ref, 3, _hip.ptr(had), _hip.ptr(d[0]), _hip.ptr(hod), _hip.ptr(d[1]),
                  0.017, 1025.0, 450.0, _hip.ptr(d[2]), _hip.ptr(d[3]), _hip.ptr(walld), 1.0,
                  _hip.ptr(cdd if use_cd else None), _hip.stream())
        oracle.couple_stress(og, 3, ha, o[0], ho, o[1], 0.017, 1025.0, 450.0, o[2], o[3], wall=wall, land_drag=1.0,
                             cd=cd if use_cd else None)
        torch.cuda.synchronize()
        for a, r in zip(d, o):
            assert relmax(_np(a), r) <= TOL1
    K = 16
    land = (rng.random(g.N) < 0.4).astype(float)
    T = rng.uniform(250, 310, g.N)
    cap = rng.uniform(1e5, 4e6, g.N)
    Tg = rng.uniform(260, 300, (K, g.N))
    dz = 0.05 * 1.5 ** np.arange(K)
    Td, Tgd, landd, capd0, dzd = dg.asarray(T), dg.asarray(Tg), dg.asarray(land), dg.asarray(cap), dg.asarray(dz)
    _hip.call('isw_land_column', dg.ref, K, _hip.ptr(landd), _hip.ptr(Td), _hip.ptr(capd0), _hip.ptr(Tgd),
              _hip.ptr(dzd), 2.0e6, 1.0, 450.0, _hip.stream())
    To, Tgo = T.copy(), Tg.copy()
    oracle.land_column(og, K, land, To, cap, Tgo, dz, 2.0e6, 1.0, 450.0)
    torch.cuda.synchronize()
    assert relmax(_np(Td), To) <= TOL1 and relmax(_np(Tgd), Tgo) <= TOL1
    capd, capo = dg.asarray(np.where(land > 0.5, 0.0, 1.0)), np.where(land > 0.5, 0.0, 1.0)
    _hip.call('isw_ocn_cap', dg.ref, _hip.ptr(hod), 1025.0 * 3990.0, _hip.ptr(capd), _hip.ptr(landd), 2.0e6,
              _hip.stream())
    oracle.ocn_cap(og, ho, 1025.0 * 3990.0, capo, land, 2.0e6)
    torch.cuda.synchronize()
    assert np.array_equal(_np(capd), capo)


@pytest.mark.parametrize('n', [1, 255, 2048, 2049, 5_000_000])
def test_sums(n, g4):
    """dsum is bit-identical to the oracle's restatement of the reference's tree; csum is within 2 ulp of the
    exactly rounded sum (math.fsum) on ill-conditioned data."""
    import oracle
    from insilica.gpu import DeviceGrid, csum, dsum
    dg = DeviceGrid(g4, 'double')
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 8, n)
    y = rng.uniform(0.5, 2.0, n)
    xd, yd = dg.asarray(x), dg.asarray(y)
    assert dsum(dg, xd) == oracle.dsum(x)
    assert dsum(dg, xd, yd) == oracle.dsum(x, y)
    exact = math.fsum(x)
    assert abs(csum(dg, xd) - exact) <= 2 * np.spacing(abs(exact)) + 1e-300
    xs = np.concatenate([x, -x[::-1]]) + 1.0 if n > 1 else x
    assert abs(csum(dg, dg.asarray(xs)) - math.fsum(xs)) <= 2 * np.spacing(abs(math.fsum(xs))) + 1e-300


def test_world_coupled_fp64(grids):
    """One tick, then 15 more (two ocean steps), of the coupled air-ocean world against the oracle's tick."""
    from insilica.world import World
    from oracle.world import OracleWorld
    print('case: world coupled_stack n=4', flush=True)
    w = World(4, 'coupled_stack', precision='double')
    o = OracleWorld(w)
    w.step(1, graph=False)
    o.step(1)
    st = w.state_arrays()
    for k, r in o.state().items():
        if k in st:
            assert relmax(_np(st[k]), r) <= TOL1, f'tick 1: {k}'
    w.step(15, graph=True)
    o.step(15)
    st = w.state_arrays()
    for k, r in o.state().items():
        if k in st:
            assert relmax(_np(st[k]), r) <= TOLN, f'tick 16: {k}'


def test_graph_equals_direct(grids):
    """Ticks replayed from captured HIP graphs equal directly launched ticks bit for bit (several replays, with
    eager work and synchronisations in between)."""
    from insilica.world import World
    print('case: graph vs direct n=4', flush=True)
    a = World(4, 'coupled_stack', precision='mixed')
    b = World(4, 'coupled_stack', precision='mixed')
    for _ in range(3):
        a.step(4, graph=True)
        b.step(4, graph=False)
        _ = float(torch.sum(a.h))
        assert a.state_hash() == b.state_hash()


def test_level6_tick_fp64():
    """The level-6 coupled world (40,962 cells, 3 air + 2 ocean layers): one tick against the oracle."""
    from insilica.world import World
    from oracle.world import OracleWorld
    print('case: world coupled_stack n=6', flush=True)
    w = World(6, 'coupled_stack', precision='double', params=dict(nfast=15))
    o = OracleWorld(w)
    w.step(1, graph=False)
    o.step(1)
    st = w.state_arrays()
    for k, r in o.state().items():
        if k in st:
            assert relmax(_np(st[k]), r) <= TOL1, f'{k}'
