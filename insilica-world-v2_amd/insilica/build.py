"""Build the native libraries in-tree: libinsilica_hip.so (hipcc, gfx950) and libinsilica_mesh.so (g++).

core.hip and radiation.hip are compiled twice each: the plain FP64 form with -ffp-contract=off (every product rounds on its own) and the
mixed form with -ffp-contract=fast (radiation.hip: fast-honor-pragmas, so that its set-up kernel's own
`fp contract(off)` holds). column.hip is compiled twice too, its cloudy air-column kernels (ISW_COLUMN_CLOUDS=1) apart
from the rest, so that the two build in parallel; ocean.hip (the ocean's column processes) mesh.hip (Lloyd's iteration) and land.hip (the land's water and ice) drag.hip (the gravity waves' drag) and tides.hip (the tides) once, FP64. All HIP code is compiled with packed-FP32 instructions disabled.
"""
import os
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = ROOT / 'csrc'
INCLUDE = ROOT.parent / 'include'
LIBDIR = PKG / '_lib'
OBJDIR = ROOT / '_build'
ARCH = os.environ.get('ISW_OFFLOAD_ARCH', 'gfx950')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

HIP_FLAGS = ['-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}', '-I', str(INCLUDE),
             '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops', '-Wno-unused-command-line-argument']


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('build failed: ' + ' '.join(cmd) + '\n' + r.stdout + r.stderr)
    return r


def _stale(src, out):
    """Whether out is older than its source or any header it may include (csrc/*.hpp, include/*.h)."""
    if not out.exists():
        return True
    deps = [CSRC / src, *CSRC.glob('*.hpp'), *INCLUDE.glob('*.h')]
    return max(d.stat().st_mtime for d in deps) > out.stat().st_mtime


def build_hip(verbose=False, force=True):
    """Compile csrc/*.hip into insilica/_lib/libinsilica_hip.so (force=False recompiles only stale units)."""
    OBJDIR.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    objs = []
    units = [('core.hip', 'core_f64.o', ['-DISW_FORM_MIXED=0', '-ffp-contract=off']),
             ('core.hip', 'core_mix.o', ['-DISW_FORM_MIXED=1', '-ffp-contract=fast']),
             ('radiation.hip', 'radiation_f64.o', ['-DISW_FORM_MIXED=0', '-ffp-contract=off']),
             ('radiation.hip', 'radiation_mix.o', ['-DISW_FORM_MIXED=1', '-ffp-contract=fast-honor-pragmas']),
             ('column.hip', 'column.o', ['-ffp-contract=off']),
             ('column.hip', 'column_clouds.o', ['-DISW_COLUMN_CLOUDS=1', '-ffp-contract=off']),
             ('ocean.hip', 'ocean.o', ['-ffp-contract=off']),
             ('mesh.hip', 'mesh.o', ['-ffp-contract=off']),
             ('land.hip', 'land.o', ['-ffp-contract=off']),
             ('drag.hip', 'drag.o', ['-ffp-contract=off']),
             ('tides.hip', 'tides.o', ['-ffp-contract=off']),
             ('capi.hip', 'capi.o', ['-ffp-contract=off'])]
    # the units compile in parallel (each hipcc is one process; at most ISW_BUILD_JOBS at a time, default one per unit
    # up to the CPU count)
    from concurrent.futures import ThreadPoolExecutor
    jobs = max(1, int(os.environ.get('ISW_BUILD_JOBS', str(min(len(units), os.cpu_count() or 4)))))

    def one(unit):
        src, obj, extra = unit
        out = OBJDIR / obj
        if not force and not _stale(src, out):
            return str(out)
        _run([HIPCC, '-c', *HIP_FLAGS, *extra, str(CSRC / src), '-o', str(out)])
        if verbose:
            print('compiled', src, '->', obj, flush=True)
        return str(out)
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(one, units))
    tmp = LIBDIR / 'libinsilica_hip.so.tmp'
    _run([HIPCC, '-shared', f'--offload-arch={ARCH}', '-o', str(tmp), *objs])
    os.replace(tmp, LIBDIR / 'libinsilica_hip.so')
    return LIBDIR / 'libinsilica_hip.so'


def build_mesh():
    """Compile the mesh builder (host C++, OpenMP) into insilica/_lib/libinsilica_mesh.so."""
    LIBDIR.mkdir(exist_ok=True)
    tmp = LIBDIR / 'libinsilica_mesh.so.tmp'
    _run(['g++', '-O2', '-ffp-contract=off', '-fopenmp', '-fPIC', '-shared', '-o', str(tmp),
          str(CSRC / 'mesh.cpp')])
    os.replace(tmp, LIBDIR / 'libinsilica_mesh.so')
    return LIBDIR / 'libinsilica_mesh.so'


def build_all(verbose=False, force=True):
    build_mesh()
    return build_hip(verbose, force)


if __name__ == '__main__':
    import sys
    print(build_all(verbose=True, force='--changed' not in sys.argv))
