"""Ensembles (GPU): every member of an ensemble ends bit-identical to the same world run alone (the reference's
ENSEMBLE check, tests/test_world.py:355-359), with members of different levels and planets advanced together."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_members_equal_solo():
    from insilica.ensemble import Ensemble
    from insilica.world import World
    print('case: ensemble of 3 coupled worlds, n=4 and n=3', flush=True)
    specs = [(4, dict(seed=1)), (4, dict(seed=2, planet=dict(radius=3.3895e6, omega=7.0882e-5, grav=3.721))),
             (3, dict(seed=3, planet=dict(radius=9.0e6, omega=1.2e-4, grav=10.5)))]
    ens = Ensemble(World(n, 'coupled_stack', precision='mixed', params=dict(p)) for n, p in specs)
    ens.advance(3, day_marks=False)
    ens.advance(2)
    for (n, p), w in zip(specs, ens):
        solo = World(n, 'coupled_stack', precision='mixed', params=dict(p))
        solo.step(int(round(5 * solo.ticks_per_hour)))
        assert solo.state_hash() == w.state_hash(), (n, p)


def test_day_latencies():
    from insilica.ensemble import Ensemble
    from insilica.world import World
    print('case: ensemble day marks', flush=True)
    ens = Ensemble([World(3, 'coupled_stack', precision='mixed'), World(4, 'coupled_stack', precision='mixed')])
    ens.advance(1)
    ens.advance(48, day_marks=True)
    lat = ens.day_latencies()
    assert [len(x) for x in lat] == [2, 2]
    assert all(v > 0 for x in lat for v in x)
    torch.cuda.synchronize()


@pytest.mark.parametrize('batch', [False, True])
def test_64_worlds_levels_4_5_6(batch):
    """The ensemble workload at its full size: 64 coupled worlds, levels 4, 5 and 6 in turn, seeded planets (radius
    0.5-1.5 Earth, rotation 0.5-2x, gravity 3.7-11), mixed form, one model day with per-day marks; the first member
    of each level ends bit-identical to the same world run alone, per-world graphs or batched per level."""
    import time
    from insilica.ensemble import Ensemble, draw_planets
    from insilica.world import World
    print(f'case: ensemble of 64 worlds, levels 4-6, batch={batch}', flush=True)
    specs = draw_planets(64, seed=0, levels=(4, 5, 6))
    t0 = time.perf_counter()
    ens = Ensemble([World(n, 'coupled_stack', precision='mixed', params=dict(p)) for n, p in specs], batch=batch)
    t1 = time.perf_counter()
    ens.advance(24, day_marks=True)
    t2 = time.perf_counter()
    lat = ens.day_latencies()
    print(f'built 64 worlds in {t1 - t0:.1f} s; one model day in {t2 - t1:.2f} s; slowest member '
          f'{max(x[0] for x in lat):.2f} s per day', flush=True)
    for lev in (4, 5, 6):
        i = next(k for k, (n, _) in enumerate(specs) if n == lev)
        solo = World(lev, 'coupled_stack', precision='mixed', params=dict(specs[i][1]))
        solo.step(int(round(24 * solo.ticks_per_hour)))
        assert ens.member_state(i)[1] == solo.state_hash(), (lev, i)
