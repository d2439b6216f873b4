"""bench.py's host logic on the CPU: the dominant kernel's HBM traffic read from the committed counter profile, and
the refusal of more than one GPU (no GPU needed)."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def test_traffic_rows_from_the_profile():
    """The traffic table's last column (MB per launch) for a kernel by name, templated names included (the first
    instantiation listed: the air's); None for a kernel it does not list."""
    import bench
    rows = {}
    for line in bench.TRAFFIC_PROFILE.read_text().splitlines():
        if line.startswith('| `k_'):
            cells = [c.strip() for c in line.strip('|').split('|')]
            name = cells[0].strip('`')[2:].split('<')[0]
            rows.setdefault(name, float(cells[-1]) * 1e6)
    assert 'af_patch2' in rows and 'slow_e' in rows
    for name, want in rows.items():
        assert bench.pmc_traffic(name) == want, name
    assert bench.pmc_traffic('no_such_kernel') is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / 'bench.py'), '--gpus', '2'], capture_output=True, text=True,
                       timeout=120, cwd=str(ROOT))
    assert r.returncode == 2 and 'one GPU' in r.stderr and not r.stdout.strip()
