"""ctypes bindings of libinsilica_hip.so (include/insilica_hip.h).

The library is loaded from insilica/_lib (built in-tree by insilica/build.py). There is no fallback: if the library
or a GPU is missing, every operation raises.
"""
import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / '_lib' / 'libinsilica_hip.so'
ISW_FP64, ISW_MIXED, ISW_EBADARG, LMAX, KSOIL = 0, 1, 9001, 8, 16
ISW_PAIR_THREADS = 256            # the most edges a patch may own for the two-sub-step kernel (one thread each)
ISW_DSUM_NOALLOC_MAX_N = 1 << 25  # the most values isw_dsum_noalloc sums (its static workspace's limit)
ISW_LLOYD_MAX_CHUNK = 4096        # the most Lloyd iterations isw_mesh_lloyd queues between reads of their moves

_i32p, _i8p = ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int8)
_f64p, _f32p = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float)


class IswGrid(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_int32) for k in ('N', 'E', 'NV', 'n_eoe')]
                + [(k, _i32p) for k in ('eI', 'eJ', 'nbr', 'cellEdge', 'cellEdgeAsc', 'ev1', 'ev2', 'vEdge', 'vCell',
                                        'eoeStart', 'eoe', 'd2idx')]
                + [(k, _i8p) for k in ('cellSign', 'vEdgeSign', 'vSlot')]
                + [(k, _f64p) for k in ('le', 'de', 'areaCell', 'areaV', 'fV', 'kite', 'w', 'd2wt', 'rA', 'rde', 'rAV')]
                + [(k, _f32p) for k in ('wf', 'd2wtf')]
                + [(k, ctypes.c_int32) for k in ('npatch', 'pcells', 'pslots', 'pown')]
                + [(k, _i32p) for k in ('ph_off', 'ph_cell', 'pe_off', 'pe_edge', 'pe_slot')]
                + [(k, ctypes.c_int32) for k in ('pslots2', 'pedges2')]
                + [(k, _i32p) for k in ('ph2_off', 'ph2_cell', 'px_off', 'px_edge', 'px_slot', 'pc_off', 'pc_eslot')]
                + [(k, ctypes.c_int32) for k in ('ptedges', 'pad3_')]
                + [(k, _i32p) for k in ('pi_off', 'pi_edge', 'tc_slot')]
                + [('eslot', _i8p)])


LAYER_SCRATCH = ('B', 'q', 'Tp', 'F', 'scale', 'G', 'He', 'ubs', 'ub0', 'ub', 'acc', 'eta0', 'eta', 'eta_prev', 'h1',
                 'h2', 'u1', 'u2', 'D1', 'Z1', 'L1', 'L2', 'hold', 'hedge', 'diss', 'pibar', 'S', 'muf', 'mup', 'phi',
                 'uf', 'accL', 'pk', 'ufb', 'ubb', 'ubc')


class IswOcean(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ('grav', 'alpha', 'alpha2', 'beta', 'S_ref', 'alpha_ref', 'rho_w', 'cw',
                                               'm_kt', 'n_kt', 'hmin', 'hmax', 'dTmin', 'sw_R', 'sw_z1', 'sw_z2',
                                               'kappa_d', 'dwf', 'ri_b', 'reach', 'h_ent', 'det_class', 'fw_local')]


class IswBatch(ctypes.Structure):
    _fields_ = ([('nworlds', ctypes.c_int32), ('nmax', ctypes.c_int32 * 3)]
                + [(k, _i32p) for k in ('wc', 'we', 'nsub')]
                + [(k, _f64p) for k in ('grav', 'nu4', 'tau', 'inv_n')])


class IswLayers(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_int32) for k in ('L', 'air', 'prec', 'pad_')]
                + [(k, ctypes.c_double) for k in ('grav', 'cp', 'kappa', 'p0', 'rd', 'nu4', 'floor_', 'div_damp', 'thin',
                                                  'shelf', 'vmix', 'vmix_cd', 'thin_wind')]
                + [(k, ctypes.c_double * LMAX) for k in ('gp', 'drag', 'stress')]
                + [(k, _f64p) for k in ('fV', 'b', 'wall', 'Gc')]
                + [(k, _f64p) for k in LAYER_SCRATCH]
                + [('pkf', _f32p), ('batch', ctypes.POINTER(IswBatch))]
                + [(k, _f64p) for k in ('tide', 'cbot')] + [('cd_bot', ctypes.c_double)]
                + [('cfc', _f64p), ('cfcf', _f32p)])


class IswTransport(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_int32) for k in ('L', 'rows_max', 'prec', 'pad_')]
                + [('beta', ctypes.c_double)]
                + [(k, _f64p) for k in ('td', 'p1', 'p2', 'Rp', 'Rm', 'Hf', 'A')]
                + [('pf', _f32p), ('pf2', _f32p)])


class IswAir(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_double) for k in ('m_air', 'm_air_kg', 'o2', 'ray', 'broad', 'o3', 'eps', 'rd',
                                                    'vf')]
                + [('cia', ctypes.c_int32), ('pad_', ctypes.c_int32)])


class IswLwTables(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_int32) for k in ('NC', 'NP', 'NT', 'NG', 'NTP', 'NO3')]
                + [(k, ctypes.c_double) for k in ('lp0', 'dlp', 't0', 'dtt', 'tp0', 'o3lp0', 'o3dlp')]
                + [(k, _f32p) for k in ('logk', 'planck', 'o3', 'cia')])


class IswSwTables(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_int32) for k in ('NC', 'NP', 'NT', 'NG')]
                + [(k, ctypes.c_double) for k in ('yscale', 'tau550', 'hssa', 'hg', 'inhom')]
                + [(k, _f32p) for k in ('logk', 'sun', 'ray', 'hzf', 'cssa', 'cssa_i', 'g_i', 'cia')])


class IswOro(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ('grav', 'p0', 'kappa', 'rd', 'G', 'cd', 'hnc', 'ric', 'dt', 'mu_min')]


class IswGw(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ('grav', 'p0', 'kappa', 'rd', 'bt', 'cw', 'kw', 'psrc', 'cmax', 'dc',
                                               'dt', 'fc')]


class IswSky(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ('tilt', 'year_s', 'day_s', 'ecc', 'lam_peri', 'sun_gm', 'sun_a',
                                               'moon_gm', 'moon_a', 'moon_e', 'moon_i', 'moon_T', 'node_T', 'aps_T',
                                               'moon_L0', 'node0', 'aps0', 'love', 'radius')]


ISW_GW_MAX_SPEEDS = 4096          # the most phase speeds isw_wave_drag's spectrum may hold


class HipError(RuntimeError):
    pass


_lib = None


def lib():
    """The loaded library (raises if it is missing: build it with `python __graft_entry__.py` or insilica.build)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise HipError(f'{LIB_PATH} is missing: build the HIP library first (python -c "import __graft_entry__ as '
                           f'g; g.build()")')
        L = ctypes.CDLL(str(LIB_PATH))
        vp, i32, i64, d = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        sig = {
            'isw_layers_step': [vp, vp, vp, vp, vp, vp, d, i32, vp],
            'isw_layers_lap': [vp, vp, vp, vp, vp],
            'isw_layers_gfast': [vp, vp, vp, vp, vp],
            'isw_layers_exchange': [vp, vp, vp, vp, vp, vp, vp, vp, i32, d, vp],
            'isw_vmix_heat': [vp, vp, vp, vp, vp, vp],
            'isw_mass_update': [vp, i32, vp, vp, vp, vp],
            'isw_transport_step': [vp, vp, vp, vp, vp, vp, i32, i32, vp],
            'isw_swe_step': [vp, vp, vp, vp, vp, d, vp],
            'isw_couple_stress': [vp, i32, vp, vp, vp, vp, d, d, d, vp, vp, vp, d, vp, vp],
            'isw_land_column': [vp, i32, vp, vp, vp, vp, vp, d, d, d, vp],
            'isw_ocn_cap': [vp, vp, d, vp, vp, d, vp],
            'isw_dsum': [vp, vp, i64, vp, vp, vp],
            'isw_csum': [vp, vp, i64, vp, vp, vp],
            'isw_dsum_noalloc': [vp, vp, i64, vp, vp],
            'isw_mesh_lloyd_device': [i64, vp, vp, vp, i32, d, vp, i32, vp, vp, vp],
            'isw_land_route': [vp, vp, vp, vp, vp, vp, vp, d, d, d, d, vp, vp, vp, vp],
            'isw_ice_flow': [vp, vp, vp, vp, d, d, vp, vp, vp, vp, vp],
            'isw_fill': [vp, i64, d, vp],
            'isw_lw_setup': [i32, i32, vp, vp, vp, d, d, d, vp, vp, vp, vp, vp, d, i32, vp, i32, vp],
            'isw_lw_column': [i32, i32, vp, vp, vp, vp, vp, i32, vp],
            'isw_sw_column': [i32, i32, vp, vp, vp, vp, i32, vp, d, vp, vp, vp, vp, i32, vp],
            'isw_hs_column': [i32, i32, vp, vp, vp, vp, vp, d, vp, vp, vp, vp],
            'isw_surface_layer': [i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp],
            'isw_cell_speed': [vp, vp, vp, vp],
            'isw_cell_speed_perot': [vp, vp, vp, vp, vp],
            'isw_gas_src': [i32, i32, i32, vp, vp, d, vp, vp],
            'isw_advance_time': [vp, d, vp],
            # N, L, par, air, mode, t, insol, pos, uabs, mu, theta, qv, Ts, ice, dt, then 29 arrays, prec, stream
            'isw_air_column': [i32, i32, vp, vp, i32] + [vp] * 9 + [d] + [vp] * 29 + [i32, vp],
            'isw_ocn_overturn': [i32, i32, vp, vp, vp, vp, vp, vp],
            'isw_ocn_forcing': [vp, d, vp, vp, vp, vp, vp, vp],
            'isw_ocn_column': [i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, d, vp, vp, vp, vp, vp, vp],
            'isw_ocn_xch_salt': [i32, i32, vp, vp, vp, i32, vp],
            'isw_ocn_pen': [i32, i32, vp, vp, vp, vp, vp, d, vp, vp, vp, vp, vp],
            'isw_ocn_fresh': [i32, vp, vp, vp, vp, vp, vp, vp, d, vp, vp, vp, vp, vp, vp, vp],
            'isw_ocn_diffuse': [vp, i32, vp, vp, vp, vp, vp, vp, d, d, vp],
            'isw_grid_check': [vp],
            'isw_air_column_options': [i32, vp],
            'isw_oro_drag': [vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, d, vp, vp],
            'isw_wave_drag': [vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, d, vp, vp],
            'isw_tide_pot': [vp, vp, vp, d, vp, vp, vp],
            'isw_tide_itd': [vp, i32, vp, vp, vp, vp, vp, d, vp, vp, vp],
            'isw_tide_mix': [vp, i32, vp, vp, vp, vp, vp, vp, d, d, d, d, vp, d, vp, d, vp],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = ctypes.c_int32, args
        L.isw_version.restype, L.isw_version.argtypes = ctypes.c_char_p, []
        L.isw_timing_start.restype, L.isw_timing_start.argtypes = ctypes.c_int32, []
        L.isw_timing_stop.restype, L.isw_timing_stop.argtypes = ctypes.c_int32, [ctypes.c_char_p, ctypes.c_int32]
        _lib = L
    return _lib


EXPORTS = ('isw_layers_step', 'isw_layers_lap', 'isw_layers_gfast', 'isw_layers_exchange', 'isw_vmix_heat',
           'isw_mass_update', 'isw_transport_step', 'isw_swe_step', 'isw_couple_stress', 'isw_land_column',
           'isw_ocn_cap', 'isw_dsum', 'isw_csum', 'isw_dsum_noalloc', 'isw_fill', 'isw_lw_setup', 'isw_lw_column', 'isw_sw_column',
           'isw_hs_column', 'isw_surface_layer', 'isw_cell_speed', 'isw_cell_speed_perot', 'isw_gas_src',
           'isw_advance_time', 'isw_air_column', 'isw_timing_start', 'isw_timing_stop', 'isw_version',
           'isw_grid_check', 'isw_air_column_options', 'isw_ocn_overturn', 'isw_ocn_forcing', 'isw_ocn_column',
           'isw_ocn_xch_salt', 'isw_ocn_pen', 'isw_ocn_fresh', 'isw_ocn_diffuse', 'isw_mesh_lloyd_device',
           'isw_land_route', 'isw_ice_flow', 'isw_oro_drag', 'isw_wave_drag', 'isw_tide_pot',
           'isw_tide_itd', 'isw_tide_mix')


def check(rc, what):
    if rc != 0:
        if rc == ISW_EBADARG:
            raise HipError(f'{what}: bad arguments (ISW_EBADARG)')
        raise HipError(f'{what}: HIP error {rc}')


def call(name, *args):
    check(getattr(lib(), name)(*args), name)


def ptr(t):
    """Device pointer of a tensor (None passes NULL)."""
    return None if t is None else t.data_ptr()


def stream():
    """The current torch stream as a hipStream_t."""
    import torch
    return torch.cuda.current_stream().cuda_stream


def timing_start():
    """Switch the library's per-kernel event timing on (direct launches only)."""
    lib().isw_timing_start()


def timing_stop():
    """Switch it off; per kernel name: dict(launches, ms, bytes) summed over the launches since timing_start()."""
    cap = 1 << 20
    buf = ctypes.create_string_buffer(cap)
    n = lib().isw_timing_stop(buf, cap)
    if n == -1:
        raise HipError('isw_timing_stop: event query failed')
    if n < 0:
        raise HipError(f'isw_timing_stop: the report needs {-n} bytes')
    out = {}
    for line in buf.value.decode().splitlines():
        name, cnt, ms, by = line.split()
        out[name] = dict(launches=int(cnt), ms=float(ms), bytes=float(by))
    return out
