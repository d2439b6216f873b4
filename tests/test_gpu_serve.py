"""The serving loop (GPU): requests that arrive while others run are admitted and overlap on the device; every result
is bit-identical to the same world run alone; a budget that holds one world at a time makes requests queue, not fail."""
import pytest

pytestmark = pytest.mark.gpu


def _alone(r):
    from insilica.world import World
    w = World(r.level, 'coupled_stack', precision='mixed', params=dict(r.params))
    w.step(int(round(r.days * 24 * w.ticks_per_hour)))
    return w.state_hash()


def test_results_same_alone_and_in_the_mix():
    from insilica.serve import Request, Server
    print('case: serve 4 requests, levels 3-4', flush=True)
    reqs = [Request(0, 0.0, 4, 2, dict(seed=1)),
            Request(1, 0.02, 3, 3, dict(seed=2, planet=dict(radius=3.3895e6, omega=7.0882e-5, grav=3.721),
                                        visc4_U=8.0)),
            Request(2, 0.05, 4, 1, dict(seed=3, planet=dict(radius=9.0e6, omega=1.2e-4, grav=10.5), visc4_U=21.0)),
            Request(3, 0.3, 3, 2, dict(seed=4))]
    done, stats = Server(keep_hashes=True, lead_hours=1).run(reqs)
    assert all(r.finished is not None and r.finished >= r.admitted >= r.arrival for r in done)
    overlap = any(a.admitted < b.finished and b.admitted < a.finished for a in done for b in done if a is not b)
    assert overlap, 'no two requests ran at the same time'
    for r in done:
        assert r.hash == _alone(r), r.rid
    assert stats['model_hours'] == 24 * sum(r.days for r in reqs)


def test_budget_queues_requests():
    from insilica.serve import Request, Server
    print('case: serve under a one-world budget', flush=True)
    reqs = [Request(k, 0.0, 3, 1, dict(seed=k)) for k in range(3)]
    srv = Server(hbm_budget=1, lead_hours=1)          # below any world's need: one world at a time
    done, _ = srv.run(reqs)
    done = sorted(done, key=lambda r: r.admitted)
    for a, b in zip(done, done[1:]):
        assert b.admitted >= a.finished


def test_levels_5_to_7_same_alone():
    """Requests at levels 5, 6 and 7 served together give the same states as each run alone."""
    from insilica.serve import Request, Server, draw_requests
    print('case: serve 4 requests, levels 5-7', flush=True)
    planets = draw_requests(4, seed=11, levels=(5, 6, 7))
    reqs = [Request(k, 0.05 * k, lev, 1, dict(planets[k].params)) for k, lev in enumerate((5, 7, 6, 5))]
    done, stats = Server(keep_hashes=True, lead_hours=2).run(reqs)
    print(f'served {stats["model_hours"]} model hours at {stats["model_hours_per_s"]:.1f} model-hours/s', flush=True)
    for r in done:
        assert r.hash == _alone(r), (r.rid, r.level)
