"""isw_dsum_noalloc through the C ABI: the tree sum on the library's static device workspace equals isw_dsum's (its
caller-given workspace) bit for bit, on two calls in a row, the second on a smaller input (the array keeps the first
call's partials, which the second must not read)."""
import ctypes

import numpy as np
import pytest

from insilica import _hip

pytestmark = pytest.mark.gpu


def _dsum(lib, x, y, n, st):
    import torch
    per = 2048
    ws = torch.zeros((n + per - 1) // per + 1 + (n + per * per - 1) // (per * per) + 1, dtype=torch.float64,
                     device='cuda')
    out = torch.full((1,), np.nan, dtype=torch.float64, device='cuda')
    _hip.check(lib.isw_dsum(x.data_ptr(), None if y is None else y.data_ptr(), n, ws.data_ptr(), out.data_ptr(), st),
               'isw_dsum')
    return out


def _noalloc(lib, x, y, n, st):
    import torch
    out = torch.full((1,), np.nan, dtype=torch.float64, device='cuda')
    _hip.check(lib.isw_dsum_noalloc(x.data_ptr(), None if y is None else y.data_ptr(), n, out.data_ptr(), st),
               'isw_dsum_noalloc')
    return out


def test_noalloc_equals_dsum_twice_second_smaller():
    import torch
    lib = _hip.lib()
    st = _hip.stream()
    rng = np.random.default_rng(7)
    # first call: three passes of the tree (5,000,001 > 2048^2), products x*y; second: one pass, x alone
    cases = [(5_000_001, True), (3_001, False)]
    for n, with_y in cases:
        print(f'isw_dsum_noalloc n={n} y={with_y}', flush=True)
        xh = rng.standard_normal(n) * 1e3 + 1.0
        yh = rng.random(n) if with_y else None
        x = torch.from_numpy(xh).cuda()
        y = torch.from_numpy(yh).cuda() if with_y else None
        got = _noalloc(lib, x, y, n, st)
        want = _dsum(lib, x, y, n, st)
        torch.cuda.synchronize()
        g, w = float(got.item()), float(want.item())
        assert g == w, (n, g, w)
        ref = float(np.sum(xh * yh)) if with_y else float(np.sum(xh))
        assert abs(g - ref) <= 1e-9 * np.sum(np.abs(xh * yh if with_y else xh)), (n, g, ref)


def test_noalloc_at_the_largest_size():
    import torch
    lib = _hip.lib()
    st = _hip.stream()
    n = _hip.ISW_DSUM_NOALLOC_MAX_N
    print(f'isw_dsum_noalloc n={n}', flush=True)
    x = torch.linspace(-1.0, 2.0, n, dtype=torch.float64, device='cuda')
    got = _noalloc(lib, x, None, n, st)
    want = _dsum(lib, x, None, n, st)
    torch.cuda.synchronize()
    assert float(got.item()) == float(want.item())
    assert lib.isw_dsum_noalloc(x.data_ptr(), None, n + 1, got.data_ptr(), st) == _hip.ISW_EBADARG
