"""World save files (the reference's .iw2 container, src/insilica/io/save.py:1-34): an uncompressed zip holding
`meta.json` and one `.npy` file per state array (raw float64 bytes, no pickles).

Deterministic: entries carry one fixed timestamp and go in sorted order, the JSON is written with sorted keys and a
fixed indent, and arrays are written C-contiguous, so saving the same state twice gives the same bytes. Reading
checks the format version and loads arrays with allow_pickle=False.
"""
import io
import json
import zipfile

import numpy as np

FORMAT_VERSION = 1
_STAMP = (1980, 1, 1, 0, 0, 0)            # the earliest date a zip entry can carry


def _entry(name):
    info = zipfile.ZipInfo(name, date_time=_STAMP)
    info.compress_type = zipfile.ZIP_STORED
    info.external_attr = 0o644 << 16
    return info


def write(path, meta, arrays):
    """Write meta (a JSON-able dict) and arrays (name -> ndarray) to path."""
    names = sorted(arrays)
    head = dict(meta, formatVersion=FORMAT_VERSION, arrays=names)
    with zipfile.ZipFile(path, 'w', compression=zipfile.ZIP_STORED) as z:
        z.writestr(_entry('meta.json'), json.dumps(head, sort_keys=True, indent=1))
        for name in names:
            buf = io.BytesIO()
            np.lib.format.write_array(buf, np.ascontiguousarray(arrays[name]), allow_pickle=False)
            z.writestr(_entry(f'{name}.npy'), buf.getvalue())


def read(path):
    """(meta, arrays) from a save file; raises ValueError for another format version."""
    with zipfile.ZipFile(path) as z:
        meta = json.loads(z.read('meta.json'))
        if meta.get('formatVersion') != FORMAT_VERSION:
            raise ValueError(f'save file format {meta.get("formatVersion")}, expected {FORMAT_VERSION}')
        arrays = {name: np.lib.format.read_array(io.BytesIO(z.read(f'{name}.npy')), allow_pickle=False)
                  for name in meta['arrays']}
    return meta, arrays
