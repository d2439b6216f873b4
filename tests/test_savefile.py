"""The save-file container (CPU): the reference's .iw2 layout (src/insilica/io/save.py:1-34; REPRO's byte identity,
tests/test_world.py:361-371): a round trip keeps every array bit for bit, the same state saved twice gives the same
bytes, and a file of another format version or holding a pickle is refused."""
import io
import json
import zipfile

import numpy as np
import pytest

from insilica import savefile


def test_round_trip_and_identical_bytes(tmp_path):
    rng = np.random.default_rng(0)
    arrays = {'h': rng.standard_normal((3, 101)), 'u': rng.standard_normal((3, 300)), 'time': np.array([1.5e6])}
    meta = dict(kind='insilica-world-v2', n=3, preset='jet', tick=17, params={'nfast': 6, 'planet': {'grav': 3.7}})
    a, b = tmp_path / 'a.iw2', tmp_path / 'b.iw2'
    savefile.write(a, meta, arrays)
    savefile.write(b, dict(reversed(list(meta.items()))), dict(reversed(list(arrays.items()))))
    assert a.read_bytes() == b.read_bytes()
    m, got = savefile.read(a)
    assert m['tick'] == 17 and m['params'] == meta['params'] and m['arrays'] == sorted(arrays)
    for k, v in arrays.items():
        assert got[k].dtype == v.dtype and got[k].shape == v.shape and np.array_equal(got[k], v)
    with zipfile.ZipFile(a) as z:
        assert [i.filename for i in z.infolist()] == ['meta.json', 'h.npy', 'time.npy', 'u.npy']
        assert all(i.compress_type == zipfile.ZIP_STORED and i.date_time == (1980, 1, 1, 0, 0, 0)
                   for i in z.infolist())


def test_refuses_other_versions_and_pickles(tmp_path):
    p = tmp_path / 'old.iw2'
    with zipfile.ZipFile(p, 'w') as z:
        z.writestr('meta.json', json.dumps(dict(formatVersion=0, arrays=[])))
    with pytest.raises(ValueError):
        savefile.read(p)
    q = tmp_path / 'pickled.iw2'
    buf = io.BytesIO()
    np.save(buf, np.array([{'x': 1}], dtype=object), allow_pickle=True)
    with zipfile.ZipFile(q, 'w') as z:
        z.writestr('meta.json', json.dumps(dict(formatVersion=1, arrays=['x'])))
        z.writestr('x.npy', buf.getvalue())
    with pytest.raises(ValueError):
        savefile.read(q)
