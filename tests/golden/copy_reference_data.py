"""Copy the reference's radiation data files in as test fixtures (run once, in a container holding the reference).

Data, not source: the line-by-line fluxes the reference's longwave test compares against (tests/data/lbl/*.npz: pressure,
temperature, gas mixing ratios and up/down fluxes per band on AFGL columns) and the k-tables that test evaluates them
with (src/insilica/data/{lw,sw}_ktables.npz: ln k per component, pressure, temperature and g-point, the Planck table per
g-point, the solar weights). Loaded with numpy's default (allow_pickle=False).
"""
import shutil
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
HERE = Path(__file__).resolve().parent

if __name__ == '__main__':
    for f in sorted((REF / 'tests' / 'data' / 'lbl').glob('lbl_*.npz')):
        shutil.copyfile(f, HERE / 'lbl' / f.name)
    for name in ('lw_ktables.npz', 'sw_ktables.npz'):
        shutil.copyfile(REF / 'src' / 'insilica' / 'data' / name, HERE / 'ktables' / name)
    print('copied')
