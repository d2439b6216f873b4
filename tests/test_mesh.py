"""The mesh (CPU): bit-identity with the reference generator's arrays at level 3 (a fixture written by
tests/golden/make_mesh_golden.py from the reference's NumPy-only mesh modules), the published Lloyd iteration
counts, and the GRID invariants (reference tests/test_grid.py:14-67)."""
from pathlib import Path

import numpy as np
import pytest

from insilica import grid as gm

GOLDEN = Path(__file__).resolve().parent / 'golden' / 'mesh_n3.npz'


@pytest.mark.parametrize('order,pre', [('v1', 'v1_'), ('hilbert', 'hb_')])
def test_mesh_matches_reference_n3(order, pre):
    z = np.load(GOLDEN, allow_pickle=False)
    g, _ = gm.make(3, order=order)
    assert g.lloydIters == int(z['lloydIters'])
    for k in z.files:
        if not k.startswith(pre):
            continue
        a, b = np.asarray(getattr(g, k[len(pre):])), z[k]
        assert a.shape == b.shape, k
        assert np.array_equal(a, b), f'{order} {k}'


def test_published_lloyd_counts(grids):
    # SPEC.md:54-60 / #30: n = 6 converges in 5,412 iterations
    assert grids(4).lloydIters == 404
    assert grids(6).lloydIters == 5412


@pytest.mark.parametrize('n', [3, 4, 5])
def test_grid_invariants(grids, n):
    g = grids(n)
    assert (g.N, g.E, g.NV) == (10 * 4 ** n + 2, 30 * 4 ** n, 20 * 4 ** n)
    R = 6.371e6
    assert abs(np.sum(g.areaCell) / (4 * np.pi * R * R) - 1) < 1e-12
    assert abs(np.sum(g.areaV) / (4 * np.pi * R * R) - 1) < 1e-12
    assert np.max(np.abs(g.kite.sum(1) - g.areaCell) / g.areaCell) < 1e-12
    # curl(grad) = 0: the circulation of a gradient around every vertex vanishes
    phi = np.random.default_rng(n).standard_normal(g.N)
    grad = (phi[g.eJ] - phi[g.eI]) / g.de
    curl = np.zeros(g.NV)
    for k in range(3):
        curl += g.vEdgeSign[:, k] * g.de[g.vEdge[:, k]] * grad[g.vEdge[:, k]]
    assert np.max(np.abs(curl)) < 1e-12 * np.max(np.abs(phi))
    # W antisymmetry: w_ee' l_e d_e = -w_e'e l_e' d_e'
    start, eoe, w = g.eoeStart, g.eoe, g.w
    row = np.repeat(np.arange(g.E), np.diff(start))
    lhs = {(int(a), int(b)): v * g.le[a] * g.de[a] for a, b, v in zip(row, eoe, w)}
    assert all((b, a) in lhs for (a, b) in lhs)
    scale = max(abs(v) for v in lhs.values())
    worst = max(abs(v + lhs[(b, a)]) for (a, b), v in lhs.items()) / scale
    assert worst < 1e-12
