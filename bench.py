"""Benchmark: model-hours/s of the coupled air-ocean world at level 6 on one GPU, mixed and plain FP64 forms.

Workload: the level-6 mesh (40,962 cells, 122,880 edges), three air layers over two ocean layers with a seeded 30%
land mask and random-phase harmonic perturbations (degree <= 20, +-2 K, +-5 m/s); split-explicit RK3 at dt 450 s with
15 fast sub-steps; FCT transport of theta, vapor and three gases (15 rows) in the air and of T and S in the ocean; the
exact air-sea stress with its heat, the land column, and one ocean step (dt 3600 s) per model hour. A step here is one
model hour: 8 ticks, replayed from captured HIP graphs, with the state resident in HBM. Synthetic data.

Printed: one JSON line with the mixed form's model-hours/s as `value`, the plain FP64 form's beside it, the dominant
kernel's roofline (algorithmic bytes per launch over its event-timed average duration, against 8 TB/s), and the CPU
oracle's rate on a bounded sample of the same workload on this host's cores.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (str(ROOT), str(ROOT / 'insilica-world-v2_amd')):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=12, help='model hours timed')
    ap.add_argument('--warmup', type=int, default=2, help='model hours before timing')
    ap.add_argument('--level', type=int, default=6)
    ap.add_argument('--no-fp64', action='store_true', help='skip the plain FP64 form')
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU oracle baseline')
    ap.add_argument('--cpu-hours', type=float, default=0.0,
                    help='model hours of the CPU sample (0: 48 for coupled_stack, 150 for the aquaplanets)')
    ap.add_argument('--ensemble', type=int, default=0, help='run the batched-ensemble workload with this many worlds')
    ap.add_argument('--no-ensemble', action='store_true',
                    help='skip the default line\'s ensemble leg (64 batched worlds at levels 4-6, one model day)')
    ap.add_argument('--ens-days', type=int, default=1, help='model days the ensemble advances')
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--batched', action='store_true', help='ensemble: batch the worlds of each level on one mesh')
    ap.add_argument('--serve', type=int, default=0, help='run the serving workload with this many requests')
    ap.add_argument('--serve-levels', default='5,6,7', help='serving: the request levels')
    ap.add_argument('--calibrate-copy', type=int, default=0,
                    help='MiB: first run one plain copy of this size (the reference dispatch for HBM counters)')
    ap.add_argument('--radiation', action='store_true', help='run the column-radiation workload (passes/s)')
    ap.add_argument('--preset', default='coupled_stack', choices=sorted(PRESET_LINES),
                    help='the world the default workload times (aqua: the moist aquaplanet with its column physics)')
    ap.add_argument('--cpu-cols', type=int, default=0, help='radiation: columns in the CPU oracle sample (0: all)')
    return ap.parse_args()


TRAFFIC_PROFILE = ROOT / 'profiles' / 'coupled_stack_level6_mixed_hbm_traffic.md'


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo) and its logical CPU count."""
    name = 'unknown'
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                name = line.split(':', 1)[1].strip()
                break
    except OSError:
        pass
    return f'{name}, {os.cpu_count()} logical CPUs'


def pmc_traffic(name):
    """HBM bytes per launch of kernel `name` at level 6, mixed form, from the committed FETCH_SIZE / WRITE_SIZE
    profile (copy-calibrated counters; the table's last column, MB per launch), or None if it has no row for it."""
    import re
    if not TRAFFIC_PROFILE.exists():
        return None
    for line in TRAFFIC_PROFILE.read_text().splitlines():
        m = re.match(r'\|\s*`k_(\w+)(?:<\d+>)?`\s*\|.*\|\s*([0-9.]+)\s*\|\s*$', line)
        if m and m.group(1) == name:
            return float(m.group(2)) * 1e6
    return None


# The worlds the default workload can time: the description in the line's config, and the reference's own published
# number for that world at level 6 in the mixed form where it has one (model-hours/s; its RTX 4070, SPEC.md:560).
PRESET_LINES = {
    'coupled_stack': ('3 air + 2 ocean layers, dt 450 s, nfast 15, 15 air + 4 ocean tracer rows, 30% land', None),
    'aqua': ('3 moist air layers over a slab sea with sea ice, dt 450 s, the sun, gray radiation, the air column '
             '(evaporation, condensation, plumes, the crossing) and the exchange every tick', 41.8),
    'aqua5': ('the moist aquaplanet in 5 air layers, dt 450 s, the column every tick', None),
    'aqua7': ('the moist aquaplanet in 7 air layers, dt 450 s, the column every tick', None),
}


def make_world(level, precision, preset='coupled_stack'):
    from insilica.world import World
    return World(level, preset, precision=precision, params=dict(nfast=15) if preset == 'coupled_stack' else None)


def time_world(w, steps, warmup):
    import torch
    ticks = int(round(w.ticks_per_hour))
    w.step(warmup * ticks)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    w.step(steps * ticks)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def kernel_profile(w, hours=1):
    """Per-kernel event times and algorithmic bytes over `hours` model hours of directly launched ticks."""
    import torch
    from insilica import _hip
    torch.cuda.synchronize()
    _hip.timing_start()
    w.step(hours * int(round(w.ticks_per_hour)), graph=False)
    torch.cuda.synchronize()
    return _hip.timing_stop()


def ensemble_run(n, days, seed, batched, roofline_leg=True):
    """The batched-ensemble workload: n coupled worlds with seeded levels (4/5/6), planet radius, rotation and gravity,
    mixed form, all advanced together by `days` model days. Returns (worlds, levels, hours, seconds, the per-day
    latencies, the largest batch's dominant-kernel roofline or None)."""
    import numpy as np
    import torch
    from insilica.ensemble import Ensemble, draw_planets
    from insilica.world import World
    torch.cuda.set_device(0)
    specs = draw_planets(n, seed=seed)
    t0 = time.perf_counter()
    worlds = [World(lev, 'coupled_stack', precision='mixed', params=p) for lev, p in specs]
    print(f'built {len(worlds)} worlds in {time.perf_counter() - t0:.1f} s', file=sys.stderr, flush=True)
    ens = Ensemble(worlds, batch=batched)
    ens.advance(1)                                   # warm-up: the first tick of each kind runs directly, then capture
    torch.cuda.synchronize()
    hours = 24 * days
    t0 = time.perf_counter()
    ens.advance(hours, day_marks=True)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    lat = np.concatenate([np.asarray(x) for x in ens.day_latencies()])
    levels = [lev for lev, _ in specs]
    roofline = None
    if batched and roofline_leg:
        # the largest batch's dominant kernel over one model hour of direct launches (the library's event timing)
        from insilica import _hip
        unit = max(ens.units, key=lambda u: u.grid.N if hasattr(u, 'grid') else 0)
        per_hour = int(round(unit.ticks_per_hour))
        torch.cuda.synchronize()
        _hip.timing_start()
        unit.step(per_hour, graph=False)
        torch.cuda.synchronize()
        prof = _hip.timing_stop()
        name, k = max(prof.items(), key=lambda kv: kv[1]['ms'])
        avg_s = k['ms'] / k['launches'] / 1e3
        ach = k['bytes'] / k['launches'] / avg_s / 1e9
        roofline = dict(bound='hbm', kernel=name, achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit='GB/s',
                        frac=round(ach / HBM_PEAK_GBS, 4), traffic=None,
                        bytes_per_launch=k['bytes'] / k['launches'], avg_us=round(avg_s * 1e6, 2),
                        batch=f'{len(unit)} worlds, {unit.grid.N} cells')
    del ens
    return worlds, levels, hours, el, lat, roofline


def ensemble_workload(n, levels, days, batched):
    return (f'{n} coupled_stack worlds, levels {sorted(set(levels))} ({levels.count(4)}/{levels.count(5)}/'
            f'{levels.count(6)} at 4/5/6), seeded planets (radius 0.5-1.5 Earth, rotation 0.5-2x, gravity 3.7-11), '
            f'{days} model days, ' + ('one batch (union mesh) per level, one launch per phase' if batched
                                       else 'one stream and graph per world'))


def ensemble_main(a):
    """The batched-ensemble workload on its own: one JSON line with the aggregate model-hours/s and the p50/p99 wall
    time per model day over the members' days."""
    import numpy as np
    worlds, levels, hours, el, lat, roofline = ensemble_run(a.ensemble, a.ens_days, a.seed, a.batched)
    line = dict(metric='ensemble model-hours/s and wall seconds per model day (p50/p99), 1 GPU',
                value=round(len(worlds) * hours / el, 2), unit='model-hours/s', n_gpus=1, steps=hours, warmup=1,
                ms_per_step=round(el / hours * 1e3, 3), higher_is_better=True, scaling='weak', vs_baseline=None,
                dtype='f64 state, f32 stencils (mixed)', data='synthetic',
                config=dict(workload=ensemble_workload(len(worlds), levels, a.ens_days, a.batched),
                            parallelism='1 GPU'),
                day_p50_s=round(float(np.percentile(lat, 50)), 4), day_p99_s=round(float(np.percentile(lat, 99)), 4),
                days_measured=int(lat.size), roofline=roofline)
    print(json.dumps(line), flush=True)


def serve_main(a):
    """The serving workload: `a.serve` requests with seeded Poisson arrivals (mean 2/s) at levels `a.serve_levels`
    (default 5-7), 1-30 model days each, seeded planets; one JSON line with the aggregate model-hours/s and p50/p99
    latency per model day."""
    import torch
    from insilica.serve import Server, draw_requests
    torch.cuda.set_device(0)
    from insilica import grid as gm
    levels = tuple(int(x) for x in a.serve_levels.split(','))
    for n in levels:
        t0 = time.perf_counter()
        gm.get(n)                                    # the meshes are built (or loaded) before the clock starts
        print(f'mesh level {n} ready in {time.perf_counter() - t0:.1f} s', file=sys.stderr, flush=True)
    reqs = draw_requests(a.serve, seed=a.seed, rate=2.0, levels=levels)
    done, st = Server(precision='mixed').run(reqs, log=lambda m: print(m, file=sys.stderr, flush=True))
    line = dict(metric='serving: model-hours/s and latency per model day (p50/p99), 1 GPU',
                value=round(st['model_hours_per_s'], 2), unit='model-hours/s', n_gpus=1, steps=len(done), warmup=0,
                ms_per_step=round(st['seconds'] / len(done) * 1e3, 3), higher_is_better=True, scaling='weak',
                vs_baseline=None, dtype='f64 state, f32 stencils (mixed)', data='synthetic',
                config=dict(workload=f'{len(done)} coupled_stack requests, Poisson arrivals 2/s (seed {a.seed}), levels '
                                     f'{a.serve_levels}, 1-30 model days, seeded planets; world construction inside the '
                                     f'timed loop (host part on worker threads)',
                            parallelism='1 GPU, continuous admission, one stream and graph per world'),
                day_p50_s=round(st['day_p50_s'], 4), day_p99_s=round(st['day_p99_s'], 4),
                model_hours=st['model_hours'], seconds=round(st['seconds'], 3),
                **{k: {q: round(v, 4) for q, v in st[k].items()} for k in ('queue_s', 'prep_s', 'build_s', 'compute_s')})
    print(json.dumps(line), flush=True)


# Algorithmic operations per (column, g-point, sub-level), clear sky, counted from the kernels' formulas: adds and
# multiplies 1 each, and exp, division and square root at their instruction cost (FP64: 20, 10, 10; FP32: 4, 4, 4 --
# the transcendental unit's quarter rate). Longwave: 6 bilinear ln k lookups (12 each) and exps, the optical depth (12),
# two Planck lookups per pass (6 each), the two-stream step down (11, exp, div) and up (clear and all-sky: 22, 3 exps,
# 2 divs, 2 products). Shortwave: 7 lookups and exps, the optical depth (14), scattering terms (10, 2 divs),
# Delta-Eddington (60, 2 exps, sqrt, 3 divs), the clear and all-sky albedos (20, 2 divs), the downward pass (15, div).
def rad_ops(form):
    ex, dv = (20, 10) if form == 'double' else (4, 4)
    lw = 6 * 12 + 12 + 2 * 12 + 11 + 22 + 2 + ex * (6 + 1 + 3) + dv * (1 + 2)
    sw = 7 * 12 + 14 + 10 + 60 + 20 + 15 + ex * (7 + 2) + dv * (2 + 3 + 1 + 2 + 1)
    return lw, sw


def col_line(pm, pd, em, ed, steps):
    """The moist air column's pass beside the radiation: steps/s in both forms and its algorithmic bytes over its
    measured duration (an N-sized pass bound by its per-column arithmetic, reported against HBM for scale)."""
    out = dict(what='air_column, moist with clouds over the sea, k-table fluxes, 40962 x 7',
               steps_per_s=round(steps / em, 1),
               fp64_steps_per_s=round(steps / ed, 1) if ed else None)
    for form, pr in (('mixed', pm), ('fp64', pd)):
        if pr and 'air_column' in pr:
            v = pr['air_column']
            us = v['ms'] / v['launches'] * 1e3
            out[form + '_us'] = round(us, 2)
            out[form + '_GBs'] = round(v['bytes'] / v['launches'] / (us * 1e-6) / 1e9, 1)
    return out


def radiation_main(a):
    """The column-radiation workload: 40,962 columns x 7 layers (the level-6 mesh's cells), synthetic k-tables with
    16 x 16 longwave and 14 x 16 shortwave g-points; one step = one radiation pass (lw_setup, lw_column, sw_column)
    over every column, the state resident in HBM. Mixed form as `value`, the FP64 form beside it. The moist air
    column's step with clouds (column physics with the pass's k-table fluxes) is timed beside it (`column`)."""
    import ctypes
    import numpy as np
    import torch
    from insilica import _hip
    from insilica import grid as gm
    from insilica.atmosphere import par_array
    from insilica.ktables import column_state, synthetic_lw, synthetic_sw
    from insilica.radiation import Longwave, Shortwave, _air
    torch.cuda.set_device(0)
    g = gm.get(6)
    L = 7
    lwk, swk = synthetic_lw(), synthetic_sw()
    pos = np.ascontiguousarray(g.pos, dtype=np.float64)
    c = column_state(pos, layers=L, seed=a.seed)
    par = par_array(L, dict(edge_k=2.0, theta_stars=[500.0, 380.0, 330.0, 305.0, 204.0, 200.0]))
    dev = lambda x: torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64), device='cuda')
    mu, th, q, Ts, alb, cl = (dev(c[k]) for k in ('mu', 'theta', 'q', 'Ts', 'alb', 'cloudo'))
    p = dev(pos)
    t = dev([2.0e6])
    N = g.N

    # the column physics beside it: the moist air column over the sea with clouds (the k-tables' fluxes of the pass,
    # the sun, evaporation, the cloud scheme with ice fall and snow, plumes, the crossing of interfaces between the
    # layers' median theta); cloud water 0.3 g/kg in the synthetic clouds' cover
    col_par = par_array(L, dict(moist=1.0, cloud=1.0, ice_air=1.0, edge_k=2.0,
                                theta_stars=[620.0, 485.0, 410.0, 360.0, 334.0, 312.0]))
    rng = np.random.default_rng(a.seed + 7)
    uabs = dev(rng.uniform(0.0, 15.0, N))
    cw0 = 3e-4 * np.asarray(c['cloudo'][:L])
    air = _air(None)

    def build(form):
        lw = Longwave(N, L, par, tables=lwk, precision=form)
        sw = Shortwave(N, L, par, lw, tables=swk, ice_re=30e-6, precision=form, cloud_ssa=swk['cloud_ssa'],
                       band_top=swk['band_top'])
        z = lambda *sh: torch.zeros(*sh, dtype=torch.float64, device='cuda')
        cs = dict(theta=th.clone(), q=q.clone(), Ts=Ts.clone(), ice=z(N), wx=z(L - 1, N), rainp=z(4, N),
                  cw=dev(cw0), co=z(4 * L + 3, N),
                  **{k: z(N) for k in ('olr', 'sw_abs', 'stored', 'rain', 'evap', 'conv', 'tsurf', 'flash')})
        return lw, sw, cs

    def one(lw, sw, cs=None):
        lw(mu, th, q, cl, Ts)
        sw(0, t, 3600.0, p, alb, cl)

    def column(lw, sw, cs, form):
        P = _hip.ptr
        _hip.call('isw_air_column', N, L, col_par.ctypes.data_as(ctypes.c_void_p), ctypes.byref(air), 0, P(t), None,
                  P(p), P(uabs), P(mu), P(cs['theta']), P(cs['q']), P(cs['Ts']), P(cs['ice']), 450.0, P(cs['wx']),
                  P(cs['olr']), P(cs['sw_abs']), P(cs['stored']), P(cs['rain']), P(cs['evap']), P(cs['conv']),
                  P(cs['tsurf']), P(cs['rainp']), None, None, None, None, None, None, None, None, None, P(cs['cw']),
                  P(cs['co']), None, P(cs['flash']), P(lw.radf), P(sw.swf), None, None, None, None, None,
                  1 if form == 'mixed' else 0, _hip.stream())

    def timed(form):
        lw, sw, cs = build(form)
        out = []
        for f in (one, lambda *x: column(*x, form)):
            for _ in range(a.warmup):
                f(lw, sw, cs)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(a.steps):
                f(lw, sw, cs)
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            _hip.timing_start()
            f(lw, sw, cs)
            torch.cuda.synchronize()
            out += [el, _hip.timing_stop()]
        return out

    el_m, prof_m, elc_m, profc_m = timed('mixed')
    print(f'mixed: {a.steps} passes in {el_m:.3f} s, {a.steps} column steps in {elc_m:.3f} s', file=sys.stderr,
          flush=True)
    el_d, prof_d, elc_d, profc_d = (None,) * 4 if a.no_fp64 else timed('double')
    if el_d:
        print(f'fp64: {a.steps} passes in {el_d:.3f} s, {a.steps} column steps in {elc_d:.3f} s', file=sys.stderr,
              flush=True)
    NGl, NGs, NS = lwk['logk'].shape[3], swk['logk'].shape[3], 4 * L

    def roof(prof, form):
        lw_ops, sw_ops = rad_ops(form)
        ops = dict(lw_column=lw_ops * NGl * NS * N, sw_column=sw_ops * NGs * NS * N)
        name = max(ops, key=lambda k: prof[k]['ms'])
        avg = prof[name]['ms'] / prof[name]['launches'] / 1e3
        ach = ops[name] / avg / 1e12
        peak = 78.6 if form == 'double' else 157.3
        return dict(bound='valu', kernel=name, achieved=round(ach, 2), peak=peak, unit='TFLOP/s',
                    frac=round(ach / peak, 4), traffic=None, ops_per_launch=ops[name], avg_us=round(avg * 1e6, 2),
                    kernels_us={k: round(v['ms'] / v['launches'] * 1e3, 2) for k, v in prof.items()})
    cpu = None
    if not a.no_cpu:
        from oracle import radiation as orad
        cores = int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))
        cols = np.arange(N) if a.cpu_cols <= 0 else np.arange(0, N, max(1, N // a.cpu_cols))
        n = len(cols)
        lwt = orad.LwTables(lwk)
        from insilica.radiation import ice_optics
        ng = np.asarray(swk['ng'], int)
        io = ice_optics(30.0, swk['band_top'])
        swt = orad.SwTables(swk, cssa_i=np.repeat([x[0] for x in io], ng), g_i=np.repeat([x[1] for x in io], ng))
        sub_c = {k: np.ascontiguousarray(v[..., cols]) for k, v in c.items()}
        t0 = time.perf_counter()
        sub = orad.lw_setup(n, L, par, lwt, sub_c['mu'], sub_c['theta'], sub_c['q'], sub_c['cloudo'])
        orad.lw_column(n, L, par, lwt, sub, sub_c['Ts'])
        orad.sw_column(n, L, par, swt, sub, sub_c['alb'], mode=0, tnow=2.0e6, interval=3600.0,
                       pos=np.ascontiguousarray(pos[cols]), cloudo=sub_c['cloudo'])
        el = time.perf_counter() - t0
        cpu = dict(value=round(n / N / el, 5), unit='passes/s', cores=cores, kind='port', model=cpu_model(),
                   sample=(f'{n} of the {N} columns (every {N // n}th), ' if n < N else f'all {N} columns, ') +
                          f'one FP64 oracle pass (C, OpenMP), {el:.1f} s' + (', scaled to the whole pass' if n < N else ''))
        if n == N:
            from oracle import column as ocol
            radf = orad.lw_column(n, L, par, lwt, sub, sub_c['Ts'])
            swf = orad.sw_column(n, L, par, swt, sub, sub_c['alb'], mode=0, tnow=2.0e6, interval=3600.0,
                                 pos=np.ascontiguousarray(pos), cloudo=sub_c['cloudo'])
            u_h = uabs.cpu().numpy()
            t0 = time.perf_counter()
            reps = 0
            while reps < 3 or time.perf_counter() - t0 < 2.0:
                ocol.air_column(col_par, 0, np.array([2.0e6]), pos, u_h, c['mu'], c['theta'], c['Ts'], 450.0,
                                radf=radf, swf=swf, q=c['q'], ice=np.zeros(N), cw=cw0, cloudo=True)
                reps += 1
            cpu['column_steps_per_s'] = round(reps / (time.perf_counter() - t0), 3)
    line = dict(metric='radiation passes/s: 40962 columns x 7 layers, LW 256 + SW 224 g-points (FP64 vs mixed, 1 GPU)',
                value=round(a.steps / el_m, 3), unit='passes/s', n_gpus=1, steps=a.steps, warmup=a.warmup,
                ms_per_step=round(el_m / a.steps * 1e3, 3), higher_is_better=True, scaling='weak', vs_baseline=None,
                dtype='f32 arithmetic, f64 g-sums (mixed)', data='synthetic',
                config=dict(workload='column radiation: level-6 cells (40962) x 7 layers, 4 sub-levels each; synthetic '
                                     'k-tables LW 16x16 / SW 14x16 g-points, k log-uniform 1e-6..1e2 m^2/kg; lapse-rate '
                                     'T 180-305 K, vapor exp(-z/2 km), clouds in ~1/3 of layers; step = lw_setup + '
                                     'lw_column + sw_column', parallelism='1 GPU'),
                fp64_value=round(a.steps / el_d, 3) if el_d else None,
                mixed_over_fp64=round(el_d / el_m, 3) if el_d else None,
                roofline=roof(prof_m, 'mixed'), roofline_fp64=roof(prof_d, 'double') if prof_d else None,
                column=col_line(profc_m, profc_d, elc_m, elc_d, a.steps), cpu_baseline=cpu)
    print(json.dumps(line), flush=True)


def main():
    a = parse()
    if a.radiation:
        if a.gpus != 1:
            print('bench.py runs on one GPU only', file=sys.stderr, flush=True)
            sys.exit(2)
        return radiation_main(a)
    if a.serve:
        if a.gpus != 1:
            print('bench.py runs on one GPU only', file=sys.stderr, flush=True)
            sys.exit(2)
        return serve_main(a)
    if a.ensemble:
        if a.gpus != 1:
            print('bench.py runs on one GPU only', file=sys.stderr, flush=True)
            sys.exit(2)
        return ensemble_main(a)
    if a.gpus != 1:
        print(f'bench.py runs on one GPU only (asked for --gpus {a.gpus}): this build has no multi-GPU path; '
              f'run with --gpus 1', file=sys.stderr, flush=True)
        sys.exit(2)
    import numpy as np
    import torch
    torch.cuda.set_device(0)
    if a.calibrate_copy:
        # a plain copy of known size (reads and writes calibrate_copy MiB once each), for the HBM counters' scale
        src = torch.ones(a.calibrate_copy * (1 << 20) // 8, dtype=torch.float64, device='cuda')
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    print(f'building level-{a.level} worlds', file=sys.stderr, flush=True)
    wm = make_world(a.level, 'mixed', a.preset)
    el_m = time_world(wm, a.steps, a.warmup)
    print(f'mixed: {a.steps} model hours in {el_m:.3f} s', file=sys.stderr, flush=True)
    prof = kernel_profile(wm)
    fp64 = None
    if not a.no_fp64:
        wd = make_world(a.level, 'double', a.preset)
        el_d = time_world(wd, a.steps, a.warmup)
        fp64 = a.steps / el_d
        print(f'fp64: {a.steps} model hours in {el_d:.3f} s', file=sys.stderr, flush=True)
        del wd
    # dominant kernel by summed time
    name, k = max(prof.items(), key=lambda kv: kv[1]['ms'])
    avg_s = k['ms'] / k['launches'] / 1e3
    bytes_per = k['bytes'] / k['launches']
    achieved = bytes_per / avg_s / 1e9
    total_ms = sum(v['ms'] for v in prof.values())
    total_bytes = sum(v['bytes'] for v in prof.values())
    traffic = pmc_traffic(name) if a.level == 6 and a.preset == 'coupled_stack' else None
    roofline = dict(bound='hbm', kernel=name, achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit='GB/s',
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, bytes_per_launch=bytes_per,
                    traffic_source=TRAFFIC_PROFILE.name if traffic else None,
                    avg_us=round(avg_s * 1e6, 3), launches_per_hour=k['launches'],
                    all_kernels_GBs=round(total_bytes / (total_ms / 1e3) / 1e9, 1),
                    kernel_ms_per_hour=round(total_ms, 3))
    cpu = None
    if not a.no_cpu:
        from oracle.world import OracleWorld
        cores = int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))
        o = OracleWorld(wm)
        tph = int(round(wm.ticks_per_hour))
        hours = a.cpu_hours or (48.0 if a.preset == 'coupled_stack' else 150.0)     # (about 10-20 s of CPU work)
        ticks = max(5, int(round(hours * tph)) // 5 * 5)
        o.step(tph)                                                              # (one model hour of warm-up)
        rates, el = [], 0.0
        for _ in range(5):                                                       # the median of five runs
            t0 = time.perf_counter()
            o.step(ticks // 5)
            dt_ = time.perf_counter() - t0
            el += dt_
            rates.append(ticks // 5 / tph / dt_)
        cpu = dict(value=round(float(np.median(rates)), 4), unit='model-hours/s', cores=cores, kind='port',
                   model=cpu_model(), runs=[round(r, 4) for r in rates],
                   sample=f'5 runs of {ticks // 5} ticks ({ticks // 5 / tph:g} model hours) each of the same level-'
                          f'{a.level} {a.preset} world after a model hour of warm-up, median; FP64 oracle (C, '
                          f'OpenMP), {el:.1f} s in all')
    # the metric's second half: a 64-world ensemble at levels 4/5/6, batched, one model day, p50/p99 per model day
    ens = None
    tph_w, cells_w, edges_w = int(round(wm.ticks_per_hour)), wm.grid.N, wm.grid.E
    if not a.no_ensemble and a.level == 6 and a.preset == 'coupled_stack':
        del wm
        import gc
        gc.collect()
        torch.cuda.empty_cache()
        worlds, levels, hours, el_e, lat, _ = ensemble_run(64, 1, a.seed, True, roofline_leg=False)
        ens = dict(worlds=len(worlds), model_hours_per_s=round(len(worlds) * hours / el_e, 2),
                   day_p50_s=round(float(np.percentile(lat, 50)), 4), day_p99_s=round(float(np.percentile(lat, 99)), 4),
                   days_measured=int(lat.size), seconds=round(el_e, 3),
                   workload=ensemble_workload(len(worlds), levels, 1, True))
        print(f'ensemble: {len(worlds)} worlds, a model day in {el_e:.2f} s', file=sys.stderr, flush=True)
        del worlds
        wm = None
    ms = el_m / a.steps * 1e3
    what, published = PRESET_LINES[a.preset]
    value = a.steps / el_m
    line = dict(metric=f'model-hours/s at level {a.level} (FP64 vs mixed, 1 GPU)' +
                ('' if a.preset == 'coupled_stack' else f', {a.preset}'), value=round(value, 3),
                unit='model-hours/s', n_gpus=1, steps=a.steps, warmup=a.warmup, ms_per_step=round(ms, 3),
                higher_is_better=True, scaling='weak',
                vs_baseline=round(value / published, 3) if published and a.level == 6 else None,
                dtype='f64 state, f32 stencils (mixed)', data='synthetic',
                config=dict(workload=f'{a.preset} level {a.level}: {what}; step = 1 model hour '
                                     f'({tph_w} ticks)',
                            cells=cells_w, edges=edges_w, parallelism='1 GPU'),
                baseline=(dict(value=published, unit='model-hours/s', hardware='RTX 4070 (the reference, mixed)',
                               source='SPEC.md:560') if published and a.level == 6 else None),
                fp64_value=round(fp64, 3) if fp64 else None,
                mixed_over_fp64=round(a.steps / el_m / fp64, 3) if fp64 else None,
                roofline=roofline, cpu_baseline=cpu, ensemble=ens,
                kernels={k2: dict(launches=v['launches'], ms=round(v['ms'], 4),
                                  GBs=round(v['bytes'] / (v['ms'] / 1e3) / 1e9, 1) if v['ms'] > 0 else None)
                         for k2, v in sorted(prof.items(), key=lambda kv: -kv[1]['ms'])})
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
