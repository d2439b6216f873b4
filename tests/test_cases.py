"""Initial-state helpers (CPU): the orthonormal associated Legendre functions behind the seeded harmonic
perturbations equal SciPy's spherical harmonics (real part at longitude 0) and are orthonormal on the sphere."""
import math

import numpy as np
from scipy.special import sph_harm_y

from insilica import cases


def test_legendre_matches_scipy_and_is_orthonormal(grids):
    rng = np.random.default_rng(0)
    pos = rng.normal(size=(4000, 3))
    pos /= np.linalg.norm(pos, axis=1)[:, None]
    pos[0], pos[1] = (0.0, 0.0, 1.0), (0.0, 0.0, -1.0)              # the poles
    P, lon = cases._legendre(pos, 20)
    assert len(P) == 230 and np.array_equal(lon, np.arctan2(pos[:, 1], pos[:, 0]))
    theta = math.pi / 2 - np.arcsin(np.clip(pos[:, 2], -1, 1))
    for (l, m), v in P.items():
        np.testing.assert_allclose(v, np.real(sph_harm_y(l, m, theta, 0.0)), rtol=0, atol=1e-13)
    # orthonormality of Y_lm = P_lm cos(m lon) (times sqrt 2 for m > 0) by the mesh's quadrature at level 5
    g = grids(5)
    Pg, lg = cases._legendre(np.asarray(g.pos), 6)
    A = g.areaCell / (6.371e6 ** 2)
    Y = {k: v * (np.cos(k[1] * lg) * math.sqrt(2.0) if k[1] else 1.0) for k, v in Pg.items()}
    for a in Y:
        for b in Y:
            ip = float(np.sum(A * Y[a] * Y[b]))
            assert abs(ip - (1.0 if a == b else 0.0)) < 2e-3, (a, b, ip)
