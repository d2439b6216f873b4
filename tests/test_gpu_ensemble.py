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
