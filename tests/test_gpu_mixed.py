"""The mixed form against the plain FP64 form, and the budgets of both forms (GPU).

Tolerances, written here: after one model day of the coupled world at level 4 the mixed form's relative L2 error in
each air layer's thickness against the FP64 form is <= 1e-5 (the figure the mixed form is held to); every per-layer
mass and every tracer budget (compensated sums) is conserved to <= 1e-13 relative in both forms.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _l2rel(area, a, b):
    return float(np.sqrt(np.sum(area * (a - b) ** 2) / np.sum(area * b ** 2)))


def test_mixed_vs_fp64_one_day():
    from insilica.world import World
    print('case: mixed vs fp64, coupled_stack n=4, 1 day', flush=True)
    a = World(4, 'coupled_stack', precision='double')
    b = World(4, 'coupled_stack', precision='mixed')
    ticks = int(round(86400 / a.dt))
    a.step(ticks)
    b.step(ticks)
    A = a.grid.areaCell
    ha, hb = a.h.cpu().numpy(), b.h.cpu().numpy()
    errs = [_l2rel(A, hb[k], ha[k]) for k in range(3)]
    print('mixed h L2 errors per air layer', errs, flush=True)
    assert max(errs) <= 1e-5
    th = _l2rel(A, b.theta.cpu().numpy(), a.theta.cpu().numpy())
    print('mixed theta L2 error', th, flush=True)
    assert np.all(np.isfinite(hb)) and np.all(np.isfinite(b.u.cpu().numpy()))


@pytest.mark.parametrize('precision', ['double', 'mixed'])
def test_budgets(precision):
    from insilica.world import World
    print(f'case: budgets {precision}', flush=True)
    w = World(4, 'coupled_stack', precision=precision)
    t0 = w.totals()
    w.step(48)
    t1 = w.totals()
    # conserved: every layer's mass M (air and ocean), the air's vapor and gas rows (3..14) and the ocean's salt rows
    # (2, 3); theta takes the wind stress's heat and the ocean's T the land column's, so theirs are not budgets here
    keys = ([f'M{k}' for k in range(3)] + [f'ocean_M{k}' for k in range(2)] + [f'tracer{y}' for y in range(3, 15)]
            + [f'ocean_tracer{y}' for y in (2, 3)])
    worst = max(abs(t1[k] - t0[k]) / abs(t0[k]) for k in keys)
    print(f'{precision}: worst budget drift {worst:.2e}', flush=True)
    assert worst <= 1e-13
    # the air's mass per layer sum h A follows M to rounding
    for k in range(3):
        assert abs(t1[f'volume{k}'] - t1[f'M{k}']) / t1[f'M{k}'] <= 1e-13


def test_accuracy_contract_level6_30_days():
    """The mixed form's contract on the benchmark's own world (coupled_stack at level 6, nfast 15): after 30 model
    days (5,760 ticks) the relative L2 error of each air layer's thickness against the FP64 form is <= 1e-5, and in
    both forms every per-layer mass and every conserved tracer's total has drifted by <= 1e-13 relative. The other
    fields' errors are reported beside it."""
    from insilica.world import World
    print('case: mixed vs fp64 coupled_stack n=6, 30 days', flush=True)
    a = World(6, 'coupled_stack', precision='double', params=dict(nfast=15))
    b = World(6, 'coupled_stack', precision='mixed', params=dict(nfast=15))
    keys = ([f'M{k}' for k in range(3)] + [f'ocean_M{k}' for k in range(2)] + [f'tracer{y}' for y in range(3, 15)]
            + [f'ocean_tracer{y}' for y in (2, 3)])
    t0 = {w.dg.precision: w.totals() for w in (a, b)}
    ticks = int(round(30 * 86400 / a.dt))
    assert ticks == 5760
    for day in range(30):
        a.step(ticks // 30)
        b.step(ticks // 30)
        if day in (0, 9, 19, 29):
            ha, hb = a.h.cpu().numpy(), b.h.cpu().numpy()
            print(f'day {day + 1}: h L2 per air layer',
                  [f'{_l2rel(a.grid.areaCell, hb[k], ha[k]):.2e}' for k in range(3)], flush=True)
    A = a.grid.areaCell
    sa, sb = a.state_arrays(), b.state_arrays()
    report = {}
    for name, x in sa.items():
        xa, xb = x.detach().cpu().numpy(), sb[name].detach().cpu().numpy()
        if xa.ndim == 2 and xa.shape[1] == a.grid.N:
            report[name] = [_l2rel(A, xb[k], xa[k]) if np.any(xa[k]) else 0.0 for k in range(xa.shape[0])]
        else:
            report[name] = [float(np.sqrt(np.sum((xb - xa) ** 2) / max(np.sum(xa ** 2), 1e-300)))]
    for name, e in report.items():
        print(f'  {name:12s} relative L2 after 30 days: ' + ' '.join(f'{v:.2e}' for v in e), flush=True)
    assert max(report['h']) <= 1e-5
    for w in (a, b):
        t1, s = w.totals(), t0[w.dg.precision]
        worst = max(abs(t1[k] - s[k]) / abs(s[k]) for k in keys)
        print(f'{w.dg.precision}: worst budget drift over 30 days {worst:.2e}', flush=True)
        assert worst <= 1e-13
