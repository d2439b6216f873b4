"""The C ABI (CPU): libinsilica_hip.so loads, exports every entry point include/insilica_hip.h declares, and rejects
bad arguments on the host with ISW_EBADARG before any HIP call (no GPU needed for that path)."""
import ctypes
import re
from pathlib import Path

from insilica import _hip

HEADER = Path(__file__).resolve().parent.parent / 'include' / 'insilica_hip.h'


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r'^\S.*?\b(isw_[a-z0-9_]+)\(', text, flags=re.M)))


def test_exports_every_declared_symbol():
    names = declared()
    assert 'isw_layers_step' in names and 'isw_transport_step' in names
    lib = _hip.lib()
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_hip.EXPORTS)
    assert lib.isw_version().decode().endswith('gfx950')


def test_bad_arguments_rejected_on_host():
    lib = _hip.lib()
    E = _hip.ISW_EBADARG
    assert lib.isw_layers_step(None, None, None, None, None, None, 1.0, 1, None) == E
    assert lib.isw_transport_step(None, None, None, None, None, None, 3, 3, None) == E
    assert lib.isw_swe_step(None, None, None, None, None, 1.0, None) == E
    assert lib.isw_dsum(None, None, 0, None, None, None) == E
    assert lib.isw_csum(None, None, 10, None, None, None) == E
    # the static-workspace sum refuses sizes beyond its array (and empty or missing operands) before any HIP call
    one = ctypes.c_void_p(1)
    assert lib.isw_dsum_noalloc(one, None, _hip.ISW_DSUM_NOALLOC_MAX_N + 1, one, None) == E
    assert lib.isw_dsum_noalloc(one, None, 0, one, None) == E
    assert lib.isw_dsum_noalloc(None, None, 10, one, None) == E
    assert lib.isw_dsum_noalloc(one, None, 10, None, None) == E
    # the device Lloyd refuses too few cells, a chunk beyond its limit and a tolerance that never stops, before any HIP call
    it, mv = ctypes.c_int32(), ctypes.c_double()
    ok = dict(N=642, max_iters=10, tol=1e-11, chunk=64)
    for bad in (dict(N=11), dict(chunk=0), dict(chunk=_hip.ISW_LLOYD_MAX_CHUNK + 1), dict(tol=0.0), dict(max_iters=0)):
        a = dict(ok, **bad)
        assert lib.isw_mesh_lloyd_device(a['N'], one, one, one, a['max_iters'], a['tol'], one, a['chunk'],
                                         ctypes.byref(it), ctypes.byref(mv), None) == E, bad
    assert lib.isw_mesh_lloyd_device(642, one, one, one, 10, 1e-11, one, 64, None, ctypes.byref(mv), None) == E
    assert lib.isw_fill(None, 5, 0.0, None) == E
    assert lib.isw_land_column(None, 3, None, None, None, None, None, 1.0, 1.0, 1.0, None) == E
    # a grid struct with sizes but a layer count beyond the maximum, and a transport with rows not a multiple of L
    g = _hip.IswGrid()
    g.N, g.E, g.NV = 12, 30, 20
    lay = _hip.IswLayers()
    lay.L = 9
    assert lib.isw_layers_step(ctypes.byref(g), ctypes.byref(lay), None, None, None, None, 1.0, 1, None) == E
    tr = _hip.IswTransport()
    tr.L, tr.rows_max = 2, 4
    assert lib.isw_transport_step(ctypes.byref(g), ctypes.byref(tr), None, None, None, None, 3, 3, None) == E
    # radiation: missing tables, a layer count beyond the maximum, more g-points than one block holds
    air = _hip.IswAir()
    lwt, swt = _hip.IswLwTables(), _hip.IswSwTables()
    assert lib.isw_lw_setup(10, 3, None, None, None, 0, 0, 0, None, None, None, None, None, 1500.0, 2, None, 0,
                            None) == E
    assert lib.isw_lw_column(10, 3, None, ctypes.byref(lwt), None, None, None, 0, None) == E
    par = (ctypes.c_double * 140)()
    lwt.NC, lwt.NP, lwt.NT, lwt.NG, lwt.NTP, lwt.NO3 = 6, 31, 20, 300, 211, 31
    lwt.dlp = lwt.dtt = lwt.o3dlp = 1.0
    junk = (ctypes.c_float * 8)()
    lwt.logk = lwt.planck = lwt.o3 = ctypes.cast(junk, _hip._f32p)
    buf = ctypes.c_void_p(1)
    assert lib.isw_lw_column(10, 3, par, ctypes.byref(lwt), buf, buf, buf, 0, None) == E          # NG 300 > 256
    lwt.NG = 136
    assert lib.isw_lw_column(10, 9, par, ctypes.byref(lwt), buf, buf, buf, 0, None) == E          # L 9 > 8
    assert lib.isw_lw_column(10, 3, par, ctypes.byref(lwt), buf, buf, buf, 7, None) == E          # unknown form
    assert lib.isw_sw_column(10, 3, par, ctypes.byref(swt), ctypes.byref(air), buf, 0, None, 3600.0, None, buf,
                             None, buf, 0, None) == E
    # the column: one layer has no interface; the mixed bottom layer needs three; no clock pointer
    hs = (ctypes.c_double * 2)(1e-6, 1e-5)
    assert lib.isw_hs_column(10, 1, par, hs, buf, buf, buf, 600.0, buf, buf, buf, None) == E
    par[70] = 1.0
    assert lib.isw_hs_column(10, 2, par, hs, buf, buf, buf, 600.0, buf, buf, buf, None) == E
    assert lib.isw_surface_layer(10, 2, par, None, buf, buf, buf, None, buf, None, None, buf, buf, None, None) == E
    assert lib.isw_cell_speed(None, buf, buf, None) == E
    assert lib.isw_gas_src(10, 3, -1, None, None, 1.0, None, None) == E
    assert lib.isw_gas_src(10, 3, 0, None, None, 1.0, None, None) == 0          # no gases: nothing to do
    assert lib.isw_advance_time(None, 1.0, None) == E
    # the air column: every option outside the dry column, a missing sun, an unknown form, bad sizes
    par[70] = 0.0

    def air_column(L=3, mode=0, t=buf, insol=None, dgo=None, xfer=None, prec=0, dt=600.0, a=ctypes.byref(air)):
        # N, L, par, air, mode, t, insol, pos, uabs, mu, theta, qv, Ts, ice, dt; wx ... tsurf; rainp, dgo; cap ...
        # cloudo; elev, flash, radf, swf, swalb, ch, lc, firn; xfer, prec, stream
        return lib.isw_air_column(10, L, par, a, mode, t, insol, buf, buf, buf, buf, None, buf, None, dt,
                                  *([buf] * 8), None, dgo, *([None] * 10), *([None] * 8), xfer, prec, None)
    par[22] = 1.0                                            # moist air: no vapor or ice given
    assert air_column() == E
    # land without its bucket's capacity, soil, ponds or snow
    assert lib.isw_air_column(10, 3, par, ctypes.byref(air), 0, buf, None, buf, buf, buf, buf, buf, buf, buf, 600.0,
                              *([buf] * 8), None, None, None, None, None, buf, buf, buf, None, buf, None, None,
                              *([None] * 8), None, 0, None) == E
    par[22] = 0.0
    assert air_column(a=None) == E
    assert air_column(t=None) == E and air_column(mode=1) == E and air_column(mode=2, insol=buf) == E
    assert air_column(prec=5) == E and air_column(L=1) == E and air_column(L=9) == E and air_column(dt=0.0) == E


def test_request_draws_are_seeded():
    from insilica.serve import draw_requests
    a, b = draw_requests(20, seed=3), draw_requests(20, seed=3)
    assert [(r.arrival, r.level, r.days) for r in a] == [(r.arrival, r.level, r.days) for r in b]
    assert all(5 <= r.level <= 6 and 1 <= r.days <= 30 for r in a)
    assert all(x.arrival < y.arrival for x, y in zip(a, a[1:]))


def test_column_options_predicate():
    """The air column carries every option of the reference's column (the cumulus layer's cu_dth, cu_dp, cl_float,
    cu_sub and sub_mix included): the predicate accepts them all in two to eight layers, dry or moist, and refuses only
    a layer count out of range or a missing parameter array (host only, no launch)."""
    lib = _hip.lib()
    E = _hip.ISW_EBADARG
    par = (ctypes.c_double * 140)()
    for moist in (0.0, 1.0):
        par[22] = moist
        for idx in (111, 112, 114, 117, 131, 47):                   # cu_dth, cu_dp, cu_sub, cl_float, sub_mix, plume_deep
            par[idx] = 1.0
        par[100] = par[70] = 1.0                                    # with the mixed layer and its condensation level
        for L in range(2, 9):
            assert lib.isw_air_column_options(L, par) == 0, (moist, L)
    assert lib.isw_air_column_options(1, par) == E and lib.isw_air_column_options(9, par) == E
    assert lib.isw_air_column_options(3, None) == E


def test_grid_check_pair_tables():
    """The two-sub-step patch tables are refused when a patch owns more edges than the pair kernel has threads, and
    accepted at the limit (host check only, no launch)."""
    lib = _hip.lib()
    E = _hip.ISW_EBADARG
    buf = (ctypes.c_int32 * 8)()
    p32 = ctypes.cast(buf, _hip._i32p)
    f64 = ctypes.cast((ctypes.c_double * 8)(), _hip._f64p)
    g = _hip.IswGrid()
    g.N, g.E, g.NV = 642, 1920, 1280
    for k in ('eI', 'eJ', 'cellEdge', 'cellEdgeAsc'):
        setattr(g, k, p32)
    g.cellSign = ctypes.cast((ctypes.c_int8 * 8)(), _hip._i8p)
    g.le = g.de = g.areaCell = f64
    assert lib.isw_grid_check(ctypes.byref(g)) == 0                    # no patches
    g.npatch, g.pcells, g.pslots, g.pown = 11, 64, 120, 200
    for k in ('ph_off', 'ph_cell', 'pe_off', 'pe_edge', 'pe_slot'):
        setattr(g, k, p32)
    assert lib.isw_grid_check(ctypes.byref(g)) == E                    # each edge's slots in its cells' rings
    g.eslot = g.cellSign
    assert lib.isw_grid_check(ctypes.byref(g)) == 0
    g.pown = 0
    assert lib.isw_grid_check(ctypes.byref(g)) == E                    # an owned-edge count is required
    g.pown = 300
    assert lib.isw_grid_check(ctypes.byref(g)) == 0                    # one sub-step per launch loops over edges
    g.pslots2, g.pedges2 = 300, 400
    for k in ('ph2_off', 'ph2_cell', 'px_off', 'px_edge', 'px_slot', 'pc_off', 'pc_eslot'):
        setattr(g, k, p32)
    assert lib.isw_grid_check(ctypes.byref(g)) == E                    # 300 owned edges > 256 threads
    g.pown = _hip.ISW_PAIR_THREADS
    assert lib.isw_grid_check(ctypes.byref(g)) == 0
    g.pown = _hip.ISW_PAIR_THREADS + 1
    assert lib.isw_grid_check(ctypes.byref(g)) == E
    g.pown = _hip.ISW_PAIR_THREADS
    g.pslots, g.pslots2 = _hip.ISW_PAIR_THREADS + 1, 400               # patch and halo: one thread per cell too
    assert lib.isw_grid_check(ctypes.byref(g)) == E
    g.pslots = _hip.ISW_PAIR_THREADS
    assert lib.isw_grid_check(ctypes.byref(g)) == 0
    assert lib.isw_grid_check(None) == E


def test_ocean_column_refusals():
    """The ocean's column entry points refuse bad sizes and settings on the host (no launch)."""
    lib = _hip.lib()
    E = _hip.ISW_EBADARG
    from insilica.ocean import OCEAN, ocean_struct
    o = ocean_struct(OCEAN, 9.80616)
    buf = ctypes.c_void_p(1)
    assert lib.isw_ocn_overturn(10, 1, ctypes.byref(o), buf, buf, buf, None, None) == E          # one layer
    assert lib.isw_ocn_overturn(10, 9, ctypes.byref(o), buf, buf, buf, None, None) == E          # beyond LMAX
    assert lib.isw_ocn_overturn(10, 3, None, buf, buf, buf, None, None) == E
    bad = ocean_struct(dict(OCEAN, hmin=0.0), 9.80616)
    assert lib.isw_ocn_overturn(10, 3, ctypes.byref(bad), buf, buf, buf, None, None) == E         # no floor
    col = lambda oo, dt=86400.0, Tc=None: lib.isw_ocn_column(10, 3, ctypes.byref(oo), buf, buf, buf, buf, buf, buf,
                                                             buf, dt, buf, None, None, Tc, None, None)
    assert col(o, dt=0.0) == E
    dc = ocean_struct(dict(OCEAN, det_class=1.0), 9.80616)
    assert col(dc) == E                                                 # detrainment by class needs the classes
    fl = ocean_struct(dict(OCEAN, fw_local=1.0), 9.80616)
    assert lib.isw_ocn_fresh(10, ctypes.byref(fl), buf, buf, buf, buf, buf, buf, 450.0, buf, None, None, None, None,
                             None, None) == E                           # fresh water with its volume needs h, M, area
    assert lib.isw_ocn_xch_salt(10, 1, buf, buf, buf, 0, None) == E
    assert lib.isw_ocn_forcing(None, 86400.0, buf, buf, buf, buf, buf, None) == E
    assert lib.isw_ocn_diffuse(None, 3, buf, buf, buf, buf, buf, None, 1.0, 1.0, None) == E


def test_drag_refusals():
    """The drags' entry points refuse, on the host, a single layer, a step that is not the settings' dt, a spectrum
    beyond ISW_GW_MAX_SPEEDS, non-positive widths and missing arrays (no launch is made for a refused call)."""
    lib = _hip.lib()
    E = _hip.ISW_EBADARG
    buf = ctypes.c_void_p(1)
    p32 = ctypes.cast(buf, _hip._i32p)
    f64 = ctypes.cast((ctypes.c_double * 8)(), _hip._f64p)
    g = _hip.IswGrid()
    g.N, g.E, g.NV = 642, 1920, 1280
    g.eI = g.eJ = g.cellEdge = p32
    g.areaCell = f64
    oro = _hip.IswOro(grav=9.8, p0=1e5, kappa=0.2857, rd=287.0, G=0.3, cd=1.0, hnc=0.5, ric=0.25, dt=900.0, mu_min=1.0)
    o = lambda L=3, dt=900.0, par=oro, sso=buf: lib.isw_oro_drag(ctypes.byref(g), L, ctypes.byref(par) if par else None,
                                                                buf, buf, buf, buf, buf, sso, buf, buf, dt, buf, None)
    assert o(L=1) == E and o(L=_hip.LMAX + 1) == E
    assert o(dt=600.0) == E and o(par=None) == E and o(sso=None) == E
    bad = _hip.IswOro(**{k: getattr(oro, k) for k, _ in oro._fields_})
    bad.ric = 0.0
    assert o(par=bad) == E
    gw = _hip.IswGw(grav=9.8, p0=1e5, kappa=0.2857, rd=287.0, bt=0.004, cw=40.0, kw=2e-5, psrc=31500.0, cmax=99.6,
                    dc=1.2, dt=900.0, fc=0.0)
    w = lambda L=3, dt=900.0, par=gw, conv=None, cmean=None: lib.isw_wave_drag(
        ctypes.byref(g), L, ctypes.byref(par), buf, buf, buf, buf, buf, conv, cmean, buf, buf, dt, buf, None)
    assert w(L=1) == E and w(dt=450.0) == E
    assert w(conv=buf, cmean=None) == E                                # the mean's output is required with conv
    wide = _hip.IswGw(**{k: getattr(gw, k) for k, _ in gw._fields_})
    wide.dc = 2 * wide.cmax / _hip.ISW_GW_MAX_SPEEDS                  # one speed more than the limit
    assert w(par=wide) == E
    wide.dc = 0.0
    assert w(par=wide) == E
