"""Batched worlds (GPU): worlds of one level with different planets stepped together on their union mesh end
bit-identical to the same worlds stepped alone, in both precision forms."""
import pytest

pytestmark = pytest.mark.gpu

SPECS = [dict(seed=1), dict(seed=2, planet=dict(radius=3.3895e6, omega=7.0882e-5, grav=3.721), visc4_U=8.0),
         dict(seed=3, planet=dict(radius=9.0e6, omega=1.2e-4, grav=10.5), visc4_U=21.0)]


@pytest.mark.parametrize('precision', ['double', 'mixed'])
def test_batch_equals_solo(precision):
    from insilica.batch import BatchWorld
    from insilica.world import World
    print(f'case: batch of 3 coupled worlds n=4, {precision}', flush=True)
    worlds = [World(4, 'coupled_stack', precision=precision, params=dict(p)) for p in SPECS]
    nf, nfo = [w.nfast for w in worlds], [w.sea.nfast for w in worlds]
    print(f'air nfast {nf}, ocean nfast {nfo}', flush=True)
    assert len(set(nf)) > 1 and len(set(nfo)) > 1, (nf, nfo)       # the batch's per-world sub-step counts differ
    bw = BatchWorld(worlds)
    solo = [World(4, 'coupled_stack', precision=precision, params=dict(p)) for p in SPECS]
    ticks = 2 * bw.ocean_every + 1
    bw.step(ticks, graph=False)
    bw.step(ticks)
    for s in solo:
        s.step(2 * ticks)
    for k, s in enumerate(solo):
        assert bw.state_hash(k) == s.state_hash(), k
