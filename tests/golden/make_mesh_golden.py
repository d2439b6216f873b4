"""Writes tests/golden/mesh_n3.npz: the reference mesh generator's arrays at level 3 (642 cells), in the
reference's default Hilbert cell order and in its v1 order, plus the published Lloyd iteration counts.

Only the mesh's numbering and bit-level geometry are pinned this way: the GRID invariants fix the mesh's
geometry to round-off but not which cell, edge or vertex gets which number, and a drop-in mesh must number them
as the reference does (reference: src/insilica/grid/icosa.py:101-353, grid/trisk.py:141-269).

Run in the build container only (it reads the reference tree under /root/reference/src, NumPy only):
    python tests/golden/make_mesh_golden.py
"""
import sys
from pathlib import Path

import numpy as np

REF = Path('/root/reference/src')
OUT = Path(__file__).resolve().parent / 'mesh_n3.npz'
KEYS = ('pos', 'nbr', 'deg', 'dist', 'edgeLen', 'area', 'eI', 'eJ', 'eKi', 'eKj', 'cornerV', 'cellEdge', 'cellSign',
        'ev1', 'ev2', 'de', 'le', 'kite', 'areaCell', 'areaV', 'circ1', 'circ2', 'fV', 'eoeStart', 'eoe', 'w', 'xm', 'ne',
        'cellEdgeAsc', 'vEdge', 'vEdgeSign', 'vCell', 'vSlot', 'vPos')


def main(n=3):
    sys.path.insert(0, str(REF))
    from insilica.grid import icosa, trisk          # NumPy-only modules of the reference
    base = icosa.build_grid(n, xp=np, max_iters=200000)
    out = {'n': n, 'lloydIters': base['lloydIters'], 'lloydMoved': base['lloydMoved']}
    v1 = dict(base)
    v1.update(trisk.build_trisk(v1))
    for k in KEYS:
        out['v1_' + k] = np.asarray(v1[k])
    keys = ('n', 'N', 'pos', 'lat', 'lon', 'area', 'east', 'north', 'nbr', 'deg', 'dist', 'edgeLen', 'edgeDir',
            'corners', 'totalArea', 'lloydIters', 'lloydMoved')
    hb = icosa.reorder({k: base[k] for k in keys}, icosa.hilbert_order(base['pos']))
    hb.update(trisk.build_trisk(hb))
    for k in KEYS + ('perm',):
        out['hb_' + k] = np.asarray(hb[k])
    np.savez_compressed(OUT, **out)
    print('wrote', OUT, 'lloyd iterations', base['lloydIters'])


if __name__ == '__main__':
    main()
