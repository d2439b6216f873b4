"""The air column (dry, and moist over the sea with and without clouds) on the GPU against the CPU oracle, through the C
ABI (isw_air_column), and the aqua_dry and aqua worlds against the oracle's tick.

Bars: FP64 <= 1e-12 relative (max |diff| / max |oracle|) on theta, Ts, the crossing's masses, olr, sw_abs, the
surface outputs and (moist) vapor, ice, rain, evaporation and the plumes' mass (the device's pow/exp/sin/cos/asin/fmod
against glibc's are the only difference); the stored energy, a difference of two ~1e11 J/m^2 totals, to 1e-9 and to
its closure sw_abs - olr within 2e-3 J/m^2; the plumes' mass to 1e-10 (it divides by a finite difference of two
lifted parcels' theta). Mixed form (float32 1 - e^-x, edge Exner powers and saturation arithmetic, the reference's
mixed build) against the oracle's mixed build: <= 1e-6 dry, <= 1e-5 moist (float32 saturation adjustment one ulp
apart, device expf against glibc's, amplified through the crossing); the plumes' mass 1e-3, the masses moved and
the rain 1e-4 (float32 parcels through the same finite difference).
"""
import ctypes

import numpy as np
import pytest
import torch

from conftest import relmax

pytestmark = pytest.mark.gpu

GRAV = 9.80616


def _dev(a):
    return None if a is None else torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device='cuda')


def _np(t):
    return t.detach().cpu().numpy().copy()


def par(L, **kw):
    from insilica.atmosphere import par_array
    stars = list(np.linspace(380.0, 305.0, L - 1)) if L > 2 else [305.0]
    return par_array(L, dict(kw, theta_stars=stars))


def columns(N, L, seed):
    rng = np.random.default_rng(seed)
    pos = rng.normal(size=(N, 3))
    pos /= np.linalg.norm(pos, axis=1)[:, None]
    share = np.concatenate([[0.0], np.linspace(0.2, 1.0, L)])
    edges = share[:, None] * 1.0e5 * rng.uniform(0.97, 1.03, (1, N))
    mu = np.diff(edges, axis=0) / GRAV
    th = np.linspace(420.0, 285.0, L)[:, None] + rng.normal(0, 2.0, (L, N))
    Ts = 288.0 - 40.0 * (pos[:, 2] ** 2 - 1.0 / 3.0) + rng.normal(0, 2.0, N)
    return dict(pos=pos, mu=mu, theta=th, Ts=Ts, uabs=rng.uniform(0, 15, N))


def gpu_air_column(p, mode, t, c, dt, insol=None, cap=None, elev=None, radf=None, swf=None, ch=None, firn=None,
                   mixed=False, ocean=False, q=None, ice=None, cw=None, cloudo=False, land=None, soilcap=None,
                   soil=None, pond=None, snow=None, lc=None, xfer=None, dgo=False):
    from insilica import _hip
    from insilica.radiation import _air
    L, N = c['mu'].shape
    z = lambda *s: torch.zeros(*s, dtype=torch.float64, device='cuda')
    o = dict(theta=_dev(c['theta']), Ts=_dev(c['Ts']), wx=z(L - 1, N), olr=z(N), sw_abs=z(N), stored=z(N),
             rain=z(N), evap=z(N), conv=z(N), tsurf=z(N), rainp=z(4, N), flash=torch.full((N,), -1.0,
             dtype=torch.float64, device='cuda'), swalb=z(N))
    if ocean:
        o['qocean'], o['qsw'] = z(N), z(N)
    if firn is not None:
        o['firn'] = _dev(firn)
    if q is not None:
        o['q'], o['ice'] = _dev(q), _dev(ice)
    if cw is not None:
        o['cw'] = _dev(cw)
        if cloudo:
            o['cloudo'] = z(4 * L + 3, N)
    if land is not None:
        o['soil'], o['pond'], o['snow'] = _dev(soil), _dev(pond), _dev(snow)
        if cap is not None:
            o['cap'] = _dev(cap)
    if xfer is not None:
        o['xfer'] = _dev(xfer)
    if dgo:
        o['dgo'] = torch.full((5 * L - 2, N), np.nan, dtype=torch.float64, device='cuda')
    tl, tsc, tlc = _dev(land), _dev(soilcap), _dev(lc)
    ins = [_dev(x) for x in (np.array([t]), insol, c['pos'], c['uabs'], c['mu'], cap, elev, radf, swf, ch)]
    tt, ti, tp, tu, tm, tc, te, tr, ts, tch = ins
    P = _hip.ptr
    air = _air(None)
    _hip.call('isw_air_column', N, L, p.ctypes.data_as(ctypes.c_void_p), ctypes.byref(air), mode, P(tt), P(ti),
              P(tp), P(tu), P(tm), P(o['theta']), P(o.get('q')), P(o['Ts']), P(o.get('ice')), float(dt), P(o['wx']),
              P(o['olr']),
              P(o['sw_abs']), P(o['stored']), P(o['rain']), P(o['evap']), P(o['conv']), P(o['tsurf']),
              P(o['rainp']), P(o.get('dgo')), P(o.get('cap', tc)), P(o.get('qocean')), P(o.get('qsw')), P(tl), P(tsc),
              P(o.get('soil')), P(o.get('pond')), P(o.get('snow')), P(o.get('cw')), P(o.get('cloudo')), P(te),
              P(o['flash']), P(tr), P(ts), P(o['swalb']), P(tch), P(tlc), P(o.get('firn')), P(o.get('xfer')),
              1 if mixed else 0, _hip.stream())
    torch.cuda.synchronize()
    return {k: _np(v) for k, v in o.items()}


def cell_errors(g, o):
    """Per column: the largest |difference| over the field's rows, relative to the field's largest |value|."""
    d = np.abs(g - o).reshape(-1, g.shape[-1]).max(0)
    return d / max(np.abs(o).max(), 1e-300)


def compare(g, o, tol, dt, label, bars=None, outliers=0.0):
    """outliers: the share of columns allowed past the bars (columns on a knife edge of the reference's own
    design, named by the caller), every other column within them."""
    keys = [k for k in ('theta', 'Ts', 'wx', 'olr', 'sw_abs', 'tsurf', 'qocean', 'qsw', 'swalb', 'firn', 'q', 'ice',
                        'cw', 'cloudo', 'soil', 'pond', 'snow', 'cap', 'xfer', 'dgo') if k in o] + \
        (['rain', 'evap', 'conv', 'rainp'] if 'q' in o else [])
    keys.append('stored')
    ce = {k: cell_errors(g[k], o[k]) for k in keys}
    bar = {'conv': 1e-10, } if tol < 1e-9 else {'conv': 1e-3, 'wx': 1e-4, 'rain': 1e-4, 'rainp': 1e-4}
    bar['stored'] = 1e-9
    bar.update(bars or {})
    past = np.zeros(o['stored'].size, bool)
    for k in keys:
        past |= ce[k] > max(tol, bar.get(k, 0.0))
    keep = ~past if outliers > 0 else np.ones_like(past)
    e = {k: float(ce[k][keep].max()) for k in keys}
    print(label, {k: float(f'{v:.3g}') for k, v in e.items()}, 'moved', int(np.count_nonzero(o['wx'])),
          'columns past the bars', int(past.sum()), flush=True)
    # the plume's mass divides by the parcel's buoyancy slope, a difference over 500 Pa of two lifted parcels' theta:
    # a ~600x cancellation of one-ulp pow/exp differences (device against glibc): 1e-13..1e-12 in FP64, and in the
    # mixed form's float32 parcels ~1e-4; the masses it moves and its rain carry a share of that
    assert past.sum() <= outliers * past.size
    for k in keys:
        assert e[k] <= max(tol, bar.get(k, 0.0)), k
    np.testing.assert_allclose(g['stored'], (g['sw_abs'] - g['olr']) * dt, rtol=0, atol=2e-3)
    zero = ('flash',) if 'q' in o else ('rain', 'evap', 'conv', 'rainp', 'flash')
    for k in zero:
        assert np.all(g[k] == 0), k
    assert np.count_nonzero(o['wx']) > 0


@pytest.mark.parametrize('L,mixed', [(2, False), (3, False), (5, False), (3, True), (7, True)])
def test_gray_sun(L, mixed):
    from oracle import column as ocol
    N, dt, t = 40962, 900.0, 0.37 * 86400.0 + 40 * 86400.0
    print(f'case: air_column gray mode 0 L={L} mixed={mixed} N={N}', flush=True)
    c = columns(N, L, seed=L)
    p = par(L, ecc=0.0167, lam_peri=4.94)
    g = gpu_air_column(p, 0, t, c, dt, mixed=mixed)
    o = ocol.air_column(p, 0, np.array([t]), c['pos'], c['uabs'], c['mu'], c['theta'], c['Ts'], dt, mix=mixed)
    compare(g, o, 1e-6 if mixed else 1e-12, dt, f'gray L={L} mixed={mixed}')


@pytest.mark.parametrize('opts', [dict(sfc=1.0, sw_tau=0.5), dict(bulk_c=1.2e-3, bulk_z=50.0, bulk_w=0.05, sw_tau=0.0),
                                  dict(bl_bot=0.5, mu_bl=4000.0, theta_rise=0.01, firn=10.0)])
def test_options(opts):
    from oracle import column as ocol
    L, N, dt = 3, 40962, 600.0
    print(f'case: air_column options {opts} mode 1 N={N}', flush=True)
    c = columns(N, L, seed=11)
    rng = np.random.default_rng(12)
    insol = rng.uniform(0, 450, N)
    cap = rng.uniform(1e7, 4e8, N)
    ch = rng.uniform(1e-3, 3e-2, N)
    elev = rng.uniform(-100, 3000, N)
    firn = rng.uniform(240, 273.15, N)
    p = par(L, **opts)
    kw = dict(insol=insol, cap=cap, ch=ch, elev=elev, firn=firn if 'firn' in opts else None, ocean=True)
    g = gpu_air_column(p, 1, 0.0, c, dt, **kw)
    o = ocol.air_column(p, 1, np.array([0.0]), c['pos'], c['uabs'], c['mu'], c['theta'], c['Ts'], dt, **kw)
    compare(g, o, 1e-12, dt, f'options {opts}')


@pytest.mark.parametrize('edge_heat,mixed', [(0.0, False), (1.0, False), (1.0, True)])
def test_ktables(edge_heat, mixed):
    from oracle import column as ocol
    from oracle import radiation as orad
    from insilica.ktables import synthetic_lw, synthetic_sw
    L, N, dt = 5, 8192, 900.0
    print(f'case: air_column k-tables edge_heat={edge_heat} mixed={mixed} N={N}', flush=True)
    c = columns(N, L, seed=21)
    p = par(L, edge_k=2.0, edge_heat=edge_heat)
    lwt, swt = orad.LwTables(synthetic_lw()), orad.SwTables(synthetic_sw())
    q = np.full((L, N), 1e-3)
    sub = orad.lw_setup(N, L, p, lwt, c['mu'], c['theta'], q)
    radf = orad.lw_column(N, L, p, lwt, sub, c['Ts'])
    swf = orad.sw_column(N, L, p, swt, sub, np.full(N, 0.06), mode=1)
    t = 0.61 * 86400.0
    g = gpu_air_column(p, 0, t, c, dt, radf=radf, swf=swf, mixed=mixed)
    o = ocol.air_column(p, 0, np.array([t]), c['pos'], c['uabs'], c['mu'], c['theta'], c['Ts'], dt, radf=radf,
                        swf=swf, mix=mixed)
    compare(g, o, 1e-6 if mixed else 1e-12, dt, f'k-tables edge_heat={edge_heat} mixed={mixed}')


def moist_columns(N, L, seed, polar=0.0):
    c = columns(N, L, seed)
    rng = np.random.default_rng(seed + 50)
    c['theta'][-1] += 15.0                                  # a warm bottom layer: the plumes go
    c['theta'][-1] -= polar * c['pos'][:, 2] ** 2            # (polar: that much colder at the poles)
    pe = np.concatenate([np.zeros((1, N)), np.cumsum(GRAV * c['mu'], 0)])
    pm = 0.5 * (pe[1:] + pe[:-1])
    T = c['theta'] * (pm / 1e5) ** (287.04 / 1004.6)
    qs = 0.622 * 611.0 * np.exp(17.27 * (T - 273.15) / (T - 35.85)) / pm
    c['q'] = np.linspace(0.1, 1.6, L)[:, None] * qs * rng.uniform(0.8, 1.2, (L, N))
    c['ice'] = np.where(rng.uniform(size=N) < 0.3, rng.uniform(0, 300, N), 0.0)
    return c


@pytest.mark.parametrize('L,mixed,opts', [(3, False, {}), (5, False, dict(lw_w0=5.0, sat_layer=0.0)),
                                          (3, True, {}), (7, True, dict(q_sea=0.98, ice_air=1.0))])
def test_moist(L, mixed, opts):
    from oracle import column as ocol
    N, dt, t = 40962, 900.0, 0.52 * 86400.0 + 100 * 86400.0
    print(f'case: air_column moist L={L} mixed={mixed} {opts} N={N}', flush=True)
    c = moist_columns(N, L, seed=30 + L)
    p = par(L, moist=1.0, **opts)
    g = gpu_air_column(p, 0, t, c, dt, mixed=mixed, q=c['q'], ice=c['ice'])
    o = ocol.air_column(p, 0, np.array([t]), c['pos'], c['uabs'], c['mu'], c['theta'], c['Ts'], dt, mix=mixed,
                        q=c['q'], ice=c['ice'])
    assert o['rain'].max() > 0 and o['ice'].max() > 0
    assert L > 5 or o['conv'].max() > 0                     # (seven thin layers: the parcel never reaches theta*)
    compare(g, o, 1e-5 if mixed else 1e-12, dt, f'moist L={L} mixed={mixed}')


def test_moist_ktables_dry_edge():
    from oracle import column as ocol
    from oracle import radiation as orad
    from insilica.ktables import synthetic_lw, synthetic_sw
    L, N, dt = 5, 8192, 900.0
    print(f'case: air_column moist k-tables dry_edge N={N}', flush=True)
    c = moist_columns(N, L, seed=41)
    p = par(L, moist=1.0, edge_k=2.0, dry_edge=1.0)
    lwt, swt = orad.LwTables(synthetic_lw()), orad.SwTables(synthetic_sw())
    sub = orad.lw_setup(N, L, p, lwt, c['mu'], c['theta'], c['q'])
    radf = orad.lw_column(N, L, p, lwt, sub, c['Ts'])
    swf = orad.sw_column(N, L, p, swt, sub, np.full(N, 0.06), mode=1)
    g = gpu_air_column(p, 1, 0.0, c, dt, insol=np.full(N, 340.0), radf=radf, swf=swf, q=c['q'], ice=c['ice'])
    o = ocol.air_column(p, 1, np.array([0.0]), c['pos'], c['uabs'], c['mu'], c['theta'], c['Ts'], dt,
                        insol=np.full(N, 340.0), radf=radf, swf=swf, q=c['q'], ice=c['ice'])
    compare(g, o, 1e-12, dt, 'moist k-tables dry_edge')


def cloudy_columns(N, L, seed):
    c = moist_columns(N, L, seed, polar=45.0)                # (cold polar columns: ice clouds, snow)
    rng = np.random.default_rng(seed + 70)
    # cloud water in most cells, up to a few times the autoconversion threshold; none in the rest
    c['cw'] = np.where(rng.uniform(size=(L, N)) < 0.7, rng.uniform(0, 1.2e-3, (L, N)), 0.0)
    return c


@pytest.mark.parametrize('L,mixed,opts', [(3, False, {}), (5, False, dict(ice_air=1.0, revap=1.0, lw_w0=5.0)),
                                          (4, False, dict(conv_revap=0.05, cw_qs=1.0, sat_layer=0.0)),
                                          (3, True, {}), (7, True, dict(ice_air=1.0, conv_revap=0.05))])
def test_clouds(L, mixed, opts):
    """The moist column with clouds: the cloud water, the cover and the clouds' effects at the top (cloudo's first
    L + 3 rows) to the moist column's bars. The mixed form in seven layers: the masses crossing an interface to 3e-4,
    vapor and theta to 3e-5 (measured 1.7e-4, 1.3e-5 and 5e-6 in the worst of 40962 cells, 22, 2 and 0 cells above
    1e-5; with land beside it theta 1.3e-5; no
    branch flips: a crossing's mass is the layer's net heating, radiative cooling less float32 latent heating, times
    its edge density, so one-ulp device-against-glibc expm1f/powf differences in the autoconversion and condensation
    reach it amplified)."""
    from oracle import column as ocol
    N, dt, t = 40962, 900.0, 0.45 * 86400.0 + 30 * 86400.0
    print(f'case: air_column clouds L={L} mixed={mixed} {opts} N={N}', flush=True)
    c = cloudy_columns(N, L, seed=60 + L)
    p = par(L, moist=1.0, cloud=1.0, **opts)
    g = gpu_air_column(p, 0, t, c, dt, mixed=mixed, q=c['q'], ice=c['ice'], cw=c['cw'], cloudo=True)
    o = ocol.air_column(p, 0, np.array([t]), c['pos'], c['uabs'], c['mu'], c['theta'], c['Ts'], dt, mix=mixed,
                        q=c['q'], ice=c['ice'], cw=c['cw'], cloudo=True)
    assert o['rain'].max() > 0 and o['cw'].max() > 0 and o['cloudo'][L].max() > 0.5
    if 'ice_air' in opts:
        assert o['rainp'][3].max() > 0                      # snow reached the ground
    compare(g, o, 1e-5 if mixed else 1e-12, dt, f'clouds L={L} mixed={mixed}',
            bars=dict(wx=3e-4, q=3e-5, theta=3e-5) if mixed else None)


def test_clouds_ktables():
    """Clouds under the k-tables' radiation: every one of cloudo's 4L + 3 rows (cover, the effects at the top, the
    in-cloud longwave and shortwave optical depths and liquid share)."""
    from oracle import column as ocol
    from oracle import radiation as orad
    from insilica.ktables import synthetic_lw, synthetic_sw
    L, N, dt = 5, 8192, 900.0
    print(f'case: air_column clouds k-tables N={N}', flush=True)
    c = cloudy_columns(N, L, seed=81)
    p = par(L, moist=1.0, cloud=1.0, edge_k=2.0, edge_heat=1.0, ice_air=1.0)
    lwt, swt = orad.LwTables(synthetic_lw()), orad.SwTables(synthetic_sw())
    sub = orad.lw_setup(N, L, p, lwt, c['mu'], c['theta'], c['q'])
    radf = orad.lw_column(N, L, p, lwt, sub, c['Ts'])
    swf = orad.sw_column(N, L, p, swt, sub, np.full(N, 0.06), mode=1)
    kw = dict(insol=np.full(N, 340.0), radf=radf, swf=swf, q=c['q'], ice=c['ice'], cw=c['cw'], cloudo=True)
    g = gpu_air_column(p, 1, 0.0, c, dt, **kw)
    o = ocol.air_column(p, 1, np.array([0.0]), c['pos'], c['uabs'], c['mu'], c['theta'], c['Ts'], dt, **kw)
    assert np.abs(o['cloudo'][2 * L + 3:3 * L + 3]).max() > 0
    compare(g, o, 1e-12, dt, 'clouds k-tables')


def land_fields(c, seed, firn=False):
    """Two thirds of the cells land: a soil bucket of 50-300 kg/m^2 a fifth to fully wet, ponds on 40%, snow on the
    land's ice (with firn: an ice sheet of at least 1.5 t/m^2 on the land's polar cells), the ground's capacity."""
    N = c['Ts'].size
    rng = np.random.default_rng(seed)
    land = (rng.uniform(size=N) < 2 / 3).astype(float)
    f = dict(land=land, soilcap=rng.uniform(50.0, 300.0, N))
    f['soil'] = f['soilcap'] * rng.uniform(0.2, 1.0, N)
    f['pond'] = np.where(rng.uniform(size=N) < 0.4, rng.uniform(0.0, 60.0, N), 0.0)
    if firn:
        polar = (land > 0.5) & (np.abs(c['pos'][:, 2]) > 0.8)
        c['ice'] = np.where(polar, np.maximum(c['ice'], 1500.0), c['ice'])
    f['snow'] = np.minimum(c['ice'], rng.uniform(0.0, 80.0, N))
    f['cap'] = np.where(land > 0.5, 2.0e6, 4.186e8)
    return f


def land_cover(N, seed):
    rng = np.random.default_rng(seed)
    return np.stack([rng.uniform(0.01, 1.0, N), rng.uniform(0.25, 0.35, N), rng.uniform(0.1, 0.2, N),
                     rng.uniform(0.0, 1.0, N), rng.uniform(0.1, 0.2, N), rng.uniform(0.3, 1.0, N),
                     rng.uniform(50.0, 150.0, N), rng.uniform(100.0, 250.0, N)])


@pytest.mark.parametrize('L,mixed,opts', [
    (3, False, {}),
    (5, False, dict(land_rs=100.0, rs_max=5000.0, c_deep=1.0, c_wet=3.0e6, c_dry=1.3e6, land_exch=2.0)),
    (4, False, dict(firn=2.0, ice_air=1.0, cloud=1.0, revap=1.0, land_rs=100.0, bare_fx=2.0, lc=1)),
    (3, True, {}),
    (7, True, dict(c_deep=1.0, land_rs=100.0, cloud=1.0, ice_air=1.0))])
def test_land(L, mixed, opts):
    """The moist column over land and sea: the soil's, ponds' and snow's water and the ground's capacity to the moist
    column's bars (the mixed form with clouds as the cloudy column's)."""
    from oracle import column as ocol
    opts = dict(opts)
    N, dt, t = 40962, 900.0, 0.45 * 86400.0 + 30 * 86400.0
    print(f'case: air_column land L={L} mixed={mixed} {opts} N={N}', flush=True)
    cloudy = opts.get('cloud', 0) > 0
    c = cloudy_columns(N, L, seed=90 + L) if cloudy else moist_columns(N, L, seed=90 + L, polar=45.0)
    kw = land_fields(c, seed=95 + L, firn='firn' in opts)
    if opts.pop('lc', 0):
        kw['lc'] = land_cover(N, 7)
    if 'firn' in opts:
        kw['firn'] = np.random.default_rng(3).uniform(240.0, 273.15, N)
    if cloudy:
        kw.update(cw=c['cw'], cloudo=True)
    p = par(L, moist=1.0, **opts)
    g = gpu_air_column(p, 0, t, c, dt, mixed=mixed, q=c['q'], ice=c['ice'], ocean=True, **kw)
    o = ocol.air_column(p, 0, np.array([t]), c['pos'], c['uabs'], c['mu'], c['theta'], c['Ts'], dt, mix=mixed,
                        q=c['q'], ice=c['ice'], ocean=True, **kw)
    land = kw['land'] > 0.5
    assert o['evap'][land].max() > 0 and o['rain'][land].max() > 0
    assert np.abs(o['soil'] - kw['soil']).max() > 0 and np.abs(o['snow'] - kw['snow']).max() > 0
    compare(g, o, 1e-5 if mixed else 1e-12, dt, f'land L={L} mixed={mixed}',
            bars=dict(wx=3e-4, q=3e-5, theta=3e-5) if mixed and cloudy else None)


@pytest.mark.parametrize('L,mixed,opts', [
    (3, False, dict(moist=0.0, bl_reach=1.0)),
    (3, False, dict(bl_lcl=1.0, bl_reach=1.0)),
    (5, False, dict(bl_lcl=1.0, ent_lcl=1.0, mu_night=2000.0, cloud=1.0, land=1)),
    (4, False, dict(bl_reach=1.0, ktables=1)),
    (3, True, dict(bl_lcl=1.0)),
    (7, True, dict(bl_lcl=1.0, bl_reach=1.0, cloud=1.0, ice_air=1.0))])
def test_mixed_layer(L, mixed, opts):
    """The mixed boundary layer (bl_mix) in the dry, moist, cloudy and land columns and under the k-tables' edge
    heating, to the moist column's bars (the mixed form with clouds as the cloudy column's). A third of the columns
    have the layer above the mixed one colder than it (dry convection, bl_reach)."""
    from oracle import column as ocol
    opts = dict(opts)
    N, dt, t = 40962, 900.0, 0.45 * 86400.0 + 30 * 86400.0
    print(f'case: air_column mixed layer L={L} mixed={mixed} {opts} N={N}', flush=True)
    moist = opts.pop('moist', 1.0) > 0
    land, kt = opts.pop('land', 0), opts.pop('ktables', 0)
    cloudy = opts.get('cloud', 0) > 0
    c = cloudy_columns(N, L, seed=110 + L) if cloudy else moist_columns(N, L, seed=110 + L, polar=30.0)
    rng = np.random.default_rng(120 + L)
    c['theta'][-2] -= np.where(rng.uniform(size=N) < 1 / 3, 25.0, 0.0)
    kw = dict(ocean=True)
    if moist:
        kw.update(q=c['q'], ice=c['ice'])
    if cloudy:
        kw.update(cw=c['cw'], cloudo=True)
    if land:
        kw.update(land_fields(c, seed=125 + L))
    p = par(L, moist=1.0 if moist else 0.0, bl_mix=1.0, **opts)
    if kt:
        from oracle import radiation as orad
        from insilica.ktables import synthetic_lw, synthetic_sw
        p = par(L, moist=1.0, bl_mix=1.0, edge_k=2.0, edge_heat=1.0, **opts)
        lwt, swt = orad.LwTables(synthetic_lw()), orad.SwTables(synthetic_sw())
        sub = orad.lw_setup(N, L, p, lwt, c['mu'], c['theta'], c['q'])
        kw['radf'] = orad.lw_column(N, L, p, lwt, sub, c['Ts'])
        kw['swf'] = orad.sw_column(N, L, p, swt, sub, np.full(N, 0.06), mode=1)
    g = gpu_air_column(p, 0, t, c, dt, mixed=mixed, **kw)
    o = ocol.air_column(p, 0, np.array([t]), c['pos'], c['uabs'], c['mu'], c['theta'], c['Ts'], dt, mix=mixed, **kw)
    assert np.abs(o['wx'][-1]).max() > 0
    # bl_lcl ends the mixed layer where its air, lifted, saturates: the parcel the plume lifts to the layer's top is
    # then saturated to the last bits, and one-ulp device-against-glibc exp/pow differences decide whether it
    # condenses there, which switches the plume's closure (measured: 23-34 of 40962 columns in FP64, 400 in the mixed
    # form, whose float32 saturation test is that much wider a knife edge); every other column is held to the bars
    compare(g, o, (1e-5 if moist else 1e-6) if mixed else 1e-12, dt, f'mixed layer L={L} mixed={mixed}',
            bars=dict(wx=3e-4, q=3e-5, theta=3e-5) if mixed and cloudy else None,
            outliers=(2e-2 if mixed else 2e-3) if opts.get('bl_lcl', 0) > 0 and moist else 0.0)


@pytest.mark.parametrize('L,mixed,opts', [
    (3, False, dict(sat_prof=1.0, sat_base=1.0, q_shape=1.0)),
    (3, False, dict(edge_vap=1.0, freeze_dry=1.0, plume_root=1.0)),
    (5, False, dict(sat_prof=1.0, edge_vap=1.0, freeze_dry=1.0, edge_k=2.0, plume_root=1.0, root_q=1.0, sfc=1.0)),
    (4, False, dict(sat_prof=1.0, edge_vap=1.0, q_shape=1.0, plume_root=1.0, cloud=1.0, bl_mix=1.0, bl_reach=1.0)),
    (3, True, dict(sat_prof=1.0, edge_vap=1.0, freeze_dry=1.0, q_shape=1.0, plume_root=1.0)),
    (7, True, dict(sat_prof=1.0, edge_vap=1.0, freeze_dry=1.0, edge_k=2.0, plume_root=1.0, cloud=1.0)),
    (5, False, dict(sat_prof=1.0, ice_air=1.0, cloud=1.0, edge_k=2.0))])
def test_earth_options(L, mixed, opts):
    """The earth preset's moist options (the spread layers' capacity, the third layer's base condensation, edge vapor,
    freeze-drying, the surface air's vapor, surface-rooted plumes with the thermals' excess) alone and together, with
    clouds and the mixed layer, to the moist column's bars."""
    from oracle import column as ocol
    opts = dict(opts)
    N, dt, t = 40962, 900.0, 0.45 * 86400.0 + 30 * 86400.0
    print(f'case: air_column earth options L={L} mixed={mixed} {opts} N={N}', flush=True)
    cloudy = opts.get('cloud', 0) > 0
    c = cloudy_columns(N, L, seed=140 + L) if cloudy else moist_columns(N, L, seed=140 + L, polar=30.0)
    kw = dict(ocean=True, q=c['q'], ice=c['ice'])
    if cloudy:
        kw.update(cw=c['cw'], cloudo=True)
    if opts.get('sfc'):
        kw['ch'] = np.random.default_rng(150 + L).uniform(1e-3, 3e-2, N)
    p = par(L, moist=1.0, **opts)
    g = gpu_air_column(p, 0, t, c, dt, mixed=mixed, **kw)
    o = ocol.air_column(p, 0, np.array([t]), c['pos'], c['uabs'], c['mu'], c['theta'], c['Ts'], dt, mix=mixed, **kw)
    compare(g, o, 1e-5 if mixed else 1e-12, dt, f'earth options L={L} mixed={mixed}',
            bars=dict(wx=3e-4, q=3e-5, theta=3e-5) if mixed and cloudy else None)


@pytest.mark.parametrize('L,mixed,opts', [
    (3, False, dict()),
    (4, False, dict(plume_ent=1e-5, ent_cloud=1e-4)),
    (4, False, dict(ent_cloud=1e-4, ent_det=1.0, cloud=1.0, ice_air=1.0, downdraft=0.3, graupel_v=1.0)),
    (5, False, dict(plume_root=1.0, plume_ent=1e-5, downdraft=0.3, revap=1.0, cloud=1.0, sat_prof=1.0, edge_k=2.0)),
    (3, True, dict(plume_ent=1e-5, ent_cloud=1e-4)),
    (7, True, dict(plume_ent=1e-5, ent_cloud=1e-4, cloud=1.0, ice_air=1.0, downdraft=0.3, graupel_v=1.0)),
    (4, False, dict(ent_cloud=1e-4, ent_ifs=1.0, buoy_v=1.0, plume_ent=1e-5)),
    (5, False, dict(ent_cloud=1e-4, ent_ifs=1.0, det_all=1.0, ent_loc=1.0, cloud=1.0, ice_air=1.0, graupel_v=1.0)),
    (4, True, dict(ent_cloud=1e-4, ent_ifs=1.0, det_all=1.0, ent_loc=1.0, buoy_v=1.0)),
    (4, False, dict(plume_root=1.0, root_q=1.0, sfc=1.0, shallow=0.03, ent_cloud=1e-4, ent_ifs=1.0, cloud=1.0,
                    ice_air=1.0)),
    (5, False, dict(plume_root=1.0, root_q=1.0, sfc=1.0, shallow=0.03, ent_cloud=1e-4, plume_ent=1e-5)),
    (4, True, dict(plume_root=1.0, root_q=1.0, sfc=1.0, shallow=0.03, ent_cloud=1e-4, ent_ifs=1.0))])
def test_deep_plume(L, mixed, opts):
    """The deep plume (its full ascent with mixing and the cloud layer's entrainment, detrainment as cloud, frozen
    share, lightning, downdrafts; the IFS's entrainment and detrainment through the cloud layer or every tropospheric
    layer, the entrained air's vapor level by level, virtual buoyancy; the shallow plume) to the moist column's
    bars."""
    from oracle import column as ocol
    opts = dict(opts)
    N, dt, t = 40962, 900.0, 0.45 * 86400.0 + 30 * 86400.0
    print(f'case: air_column deep plume L={L} mixed={mixed} {opts} N={N}', flush=True)
    cloudy = opts.get('cloud', 0) > 0
    c = cloudy_columns(N, L, seed=160 + L) if cloudy else moist_columns(N, L, seed=160 + L, polar=30.0)
    kw = dict(ocean=True, q=c['q'], ice=c['ice'])
    if cloudy:
        kw.update(cw=c['cw'], cloudo=True)
    if opts.get('sfc'):
        kw['ch'] = np.random.default_rng(165 + L).uniform(1e-3, 3e-2, N)
    p = par(L, moist=1.0, plume_deep=1.0, **opts)
    g = gpu_air_column(p, 0, t, c, dt, mixed=mixed, **kw)
    o = ocol.air_column(p, 0, np.array([t]), c['pos'], c['uabs'], c['mu'], c['theta'], c['Ts'], dt, mix=mixed, **kw)
    assert o['conv'].max() > 0
    if opts.get('graupel_v') and L == 4:                  # (seven thin layers: the plumes stop below the freezing band)
        assert o['flash'].max() > 0
    # lightning: the storm's ice is the condensate between the first sub-levels the plume crosses 0 and -40 C at, a
    # threshold the mixed form's float32 parcel can cross one sub-level apart (measured: 1 of 40962 columns, 0.8%)
    flash_g, flash_o = g.pop('flash'), o.pop('flash')
    np.testing.assert_allclose(flash_g, flash_o, rtol=2e-2 if mixed else 1e-10, atol=1e-12 * max(flash_o.max(), 1e-30))
    g['flash'] = o['flash'] = np.zeros_like(flash_o)
    compare(g, o, 1e-5 if mixed else 1e-12, dt, f'deep plume L={L} mixed={mixed}',
            bars=dict(wx=3e-4, q=3e-5, theta=3e-5) if mixed and cloudy else None)


@pytest.mark.parametrize('L,mixed,opts', [
    (4, False, dict(ml_clear=1.0, cu_cloud=1.0)),
    (4, False, dict(plume_deep=1.0, downdraft=0.3, ent_cloud=1e-4)),
    (5, False, dict(ml_clear=1.0, cu_cloud=2.0, plume_deep=1.0, edge_vap=1.0, q_shape=1.0, sat_prof=1.0, edge_k=2.0)),
    (4, True, dict(ml_clear=1.0, cu_cloud=1.0, plume_deep=1.0, plume_ent=1e-5))])
def test_bl_q(L, mixed, opts):
    """The mixed layer's even vapor and its cloud statistics (bl_q with bl_mix and clouds; ml_clear, cu_cloud), alone
    and with the deep plume's mixed-layer limit, to the moist column's bars (the mixed form with clouds as the
    cloudy column's)."""
    from oracle import column as ocol
    N, dt, t = 40962, 900.0, 0.45 * 86400.0 + 30 * 86400.0
    print(f'case: air_column bl_q L={L} mixed={mixed} {opts} N={N}', flush=True)
    c = cloudy_columns(N, L, seed=170 + L)
    kw = dict(ocean=True, q=c['q'], ice=c['ice'], cw=c['cw'], cloudo=True)
    p = par(L, moist=1.0, cloud=1.0, bl_mix=1.0, bl_q=1.0, **opts)
    g = gpu_air_column(p, 0, t, c, dt, mixed=mixed, **kw)
    o = ocol.air_column(p, 0, np.array([t]), c['pos'], c['uabs'], c['mu'], c['theta'], c['Ts'], dt, mix=mixed, **kw)
    flash_g, flash_o = g.pop('flash'), o.pop('flash')
    np.testing.assert_allclose(flash_g, flash_o, rtol=2e-2 if mixed else 1e-10, atol=1e-12 * max(flash_o.max(), 1e-30))
    g['flash'] = o['flash'] = np.zeros_like(flash_o)
    compare(g, o, 1e-5 if mixed else 1e-12, dt, f'bl_q L={L} mixed={mixed}',
            bars=dict(wx=3e-4, q=3e-5, theta=3e-5) if mixed else None)


def test_aqua_dry_world():
    """The 'aqua_dry' world (three dry air layers over the slab, the sun, gray radiation, the column and the exchange
    every tick) against the oracle's tick, 12 ticks at level 4 in the FP64 form (<= 1e-10, as the core's own
    multi-step bar); graph replay bit-identical to direct launches; the column moved air across an interface. The
    annual-mean sun (mean_sun), the moist worlds and the moist worlds with clouds likewise."""
    from insilica.world import World
    from oracle.world import OracleWorld
    for preset, mean_sun, phys in (('aqua_dry', False, {}), ('aqua_dry', True, {}), ('aqua', False, {}),
                                   ('aqua5', False, {}), ('aqua7', False, {}),
                                   ('aqua', False, dict(cloud=1.0, ice_air=1.0, rh_crit=0.6)),
                                   ('aqua5', False, dict(cloud=1.0, rh_crit=0.6))):
        print(f'case: {preset} world level 4, 12 ticks, mean_sun={mean_sun} {phys}', flush=True)
        w = World(4, preset, precision='double', params=dict(mean_sun=mean_sun, physics=phys))
        o = OracleWorld(w)
        w.step(12, graph=False)
        o.step(12)
        got = {k: _np(v) for k, v in w.state_arrays().items()}
        ref = o.state()
        errs = {k: relmax(got[k], ref[k]) for k in ref}
        print(f'{preset} world errors', {k: float(f'{v:.3g}') for k, v in errs.items()}, flush=True)
        assert max(errs.values()) <= 1e-10
        assert np.count_nonzero(_np(w.physics.wx)) > 0
        if phys:
            assert _np(w.cw).max() > 0                    # clouds formed (the start's vapor at 70% of saturation below)
        w2 = World(4, preset, precision='double', params=dict(mean_sun=mean_sun, physics=phys))
        w2.step(12, graph=True)
        assert w2.state_hash() == w.state_hash()


def test_aqua_dry_world_ktables_and_mixed():
    """The aqua_dry world with the k-tables' radiation every hour (radiation ticks captured as their own graph kind):
    graph replay bit-identical to direct launches, the column's energy closes per cell on its last tick. The mixed
    form of the gray world stays within 1e-5 (relative L2 in h) of the FP64 form over a model day at level 5."""
    from insilica.ktables import synthetic_lw, synthetic_sw
    from insilica.world import World
    print('case: aqua_dry world level 4 k-tables, 10 ticks', flush=True)
    P = dict(longwave='ktables', lw_tables=synthetic_lw(), sw_tables=synthetic_sw())
    w = World(4, 'aqua_dry', precision='double', params=P)
    w.step(10, graph=False)
    w2 = World(4, 'aqua_dry', precision='double', params=P)
    w2.step(10, graph=True)
    assert w2.state_hash() == w.state_hash()
    ph = w.physics
    np.testing.assert_allclose(_np(ph.stored), (_np(ph.sw_abs) - _np(ph.olr)) * w.dt, rtol=0, atol=2e-3)
    assert float(ph.olr.min()) > 50.0
    print('case: aqua_dry world level 5 mixed vs FP64, one model day', flush=True)
    a, b = World(5, 'aqua_dry', precision='double'), World(5, 'aqua_dry', precision='mixed')
    k = int(round(86400.0 / a.dt))
    a.step(k)
    b.step(k)
    ha, hb = _np(a.h), _np(b.h)
    l2 = float(np.linalg.norm(hb - ha) / np.linalg.norm(ha))
    print('aqua_dry mixed vs FP64 after a day: rel L2 h', l2, 'theta', relmax(_np(b.theta), _np(a.theta)), flush=True)
    assert l2 <= 1e-5


@pytest.mark.parametrize('L,mixed,case', [(5, False, 0), (5, False, 1), (5, False, 2), (7, False, 3), (7, False, 4),
                                          (5, True, 1), (7, True, 3)])
def test_cumulus_layer(L, mixed, case):
    """The cumulus layer (cu_dth's floating edges with cu_dp and cl_float, the plumes' air taken in through it moved
    with them, the shallow plume stopping in it or the cloud layer above, the mixed layer refilled from it with
    sub_mix; the earth7 column's options together) with the gas tracers carried by the column's moves.

    FP64 bars from the rounding alone: the oracle against itself built with exp, expm1 and pow one ulp off for about
    half their arguments (a C library that rounds differently, as the device's does) differs on these columns by up
    to 3.3e-11 in vapor, 4.4e-11 in the gas tracers, 1.4e-11 in rain, 1.6e-11 in the plume mass, 1.3e-10 in cloud
    water and 3.6e-12 in the masses crossing: the condensate is a difference of the plume's vapor terms, q0 + sum Gl q
    - G qc - sum qx, and each term carries the ulps of a dozen sub-levels' exponentials. Hence q and xfer 1e-10, rain
    3e-11, cloud water and its rows 1e-9, wx 1e-11 (the oracle against its perturbed self passes these in every
    column but 2 of 40962 at the mixed layer's condensation level). The mixed form as the cloudy column's; with
    bl_lcl the condensation level's knife edge as in the mixed-layer test."""
    from oracle import column as ocol
    from test_oracle_air_column import CU_OPTS, cu_par, cumulus_columns
    L_, opts = CU_OPTS[case]
    assert L_ == L
    N, dt, t = 40962, 900.0, 0.45 * 86400.0 + 30 * 86400.0
    print(f'case: air_column cumulus L={L} mixed={mixed} case {case} N={N}', flush=True)
    c = cumulus_columns(N, L, seed=230 + L)
    cloudy = opts.get('cloud', 0) > 0
    kw = dict(ocean=True, q=c['q'], ice=c['ice'])
    if cloudy:
        kw.update(cw=c['cw'], cloudo=True)
    if opts.get('sfc'):
        kw['ch'] = np.random.default_rng(235 + L).uniform(1e-3, 3e-2, N)
    rng = np.random.default_rng(236 + L)
    kw['xfer'] = rng.uniform(0.5, 1.5, (3 * L, N)) * np.repeat([400e-6, 0.32e-6, 1.8e-6], L)[:, None]
    p = cu_par(L, **opts)
    g = gpu_air_column(p, 0, t, c, dt, mixed=mixed, **kw)
    o = ocol.air_column(p, 0, np.array([t]), c['pos'], c['uabs'], c['mu'], c['theta'], c['Ts'], dt, mix=mixed, **kw)
    assert o['conv'].max() > 0 and np.any(o['xfer'] != kw['xfer'])
    flash_g, flash_o = g.pop('flash'), o.pop('flash')
    np.testing.assert_allclose(flash_g, flash_o, rtol=2e-2 if mixed else 1e-10, atol=1e-12 * max(flash_o.max(), 1e-30))
    g['flash'] = o['flash'] = np.zeros_like(flash_o)
    lcl = opts.get('bl_lcl', 0) > 0
    fp64 = dict(q=1e-10, xfer=1e-10, rain=3e-11, rainp=3e-11, cw=1e-9, cloudo=1e-9, wx=1e-11, conv=1e-10)
    # (the mixed form lifts cu_top's parcel in float32, so the floating edges sit at float32 precision and a few
    # crossings at them switch: measured 2 of 40962 columns with cu_dp; allowed 0.1%)
    compare(g, o, 1e-5 if mixed else 1e-12, dt, f'cumulus L={L} mixed={mixed} case {case}',
            bars=(dict(wx=3e-4, q=3e-5, theta=3e-5, xfer=3e-5) if cloudy else None) if mixed else fp64,
            outliers=(2e-2 if mixed else 2e-3) if lcl else (1e-3 if mixed else 0.0))


@pytest.mark.parametrize('L,mixed,case', [(5, False, 2), (7, False, 3), (7, True, 3)])
def test_diagnostics_rows(L, mixed, case):
    """The diagnostics rows (dgo: heating and latent heating per layer, the crossing's gross up and down, the convective
    mass per layer) on the cumulus columns, to the same bars as the fields they come from (latent heating as rain and
    cloud water, the crossing and convective masses as wx and conv)."""
    from oracle import column as ocol
    from test_oracle_air_column import CU_OPTS, cu_par, cumulus_columns
    _, opts = CU_OPTS[case]
    N, dt, t = 40962, 900.0, 0.45 * 86400.0 + 30 * 86400.0
    print(f'case: air_column diagnostics L={L} mixed={mixed} case {case} N={N}', flush=True)
    c = cumulus_columns(N, L, seed=240 + L)
    kw = dict(ocean=True, q=c['q'], ice=c['ice'], cw=c['cw'], cloudo=True)
    if opts.get('sfc'):
        kw['ch'] = np.random.default_rng(245 + L).uniform(1e-3, 3e-2, N)
    p = cu_par(L, **opts)
    g = gpu_air_column(p, 0, t, c, dt, mixed=mixed, dgo=True, **kw)
    o = ocol.air_column(p, 0, np.array([t]), c['pos'], c['uabs'], c['mu'], c['theta'], c['Ts'], dt, mix=mixed,
                        dgo=True, **kw)
    assert np.all(np.isfinite(g['dgo'])) and np.any(o['dgo'][4 * L - 2:] != 0)
    for name, rows, fp64, mix in (('Q', slice(0, L), 1e-12, 1e-5), ('latent', slice(L, 2 * L), 1e-9, 3e-4),
                                  ('gross', slice(2 * L, 4 * L - 2), 1e-11, 3e-4),
                                  ('convective', slice(4 * L - 2, 5 * L - 2), 1e-10, 1e-3)):
        ce = cell_errors(g['dgo'][rows], o['dgo'][rows])
        bar = mix if mixed else fp64
        past = ce > bar
        print(f'  {name}: worst {ce.max():.3g}, past {int(past.sum())}', flush=True)
        assert past.sum() <= (2e-2 if mixed else 2e-3) * N, name
