"""Reader for the "MOCDUMP1" record files written by oracle/ref_driver.cpp.

Test infrastructure: only tests, the smoke check and the bench CPU-baseline
leg read these files.
"""
import struct

import numpy as np

_DT = {b"d": np.float64, b"f": np.float32, b"i": np.int32, b"l": np.int64}


def read(path):
    out = {}
    with open(path, "rb") as f:
        if f.read(8) != b"MOCDUMP1":
            raise ValueError(f"{path}: not a MOCDUMP1 file")
        while True:
            hdr = f.read(4)
            if not hdr:
                break
            (n,) = struct.unpack("<I", hdr)
            name = f.read(n).decode()
            dt = _DT[f.read(1)]
            (cnt,) = struct.unpack("<Q", f.read(8))
            arr = np.frombuffer(f.read(cnt * np.dtype(dt).itemsize), dtype=dt)
            out[name] = arr[0] if cnt == 1 and name not in _ARRAYS else arr
    return out


# names that stay arrays even when they hold a single element
_ARRAYS = {"flux", "fsr_volume", "fsr_material", "seg_length", "seg_fsr",
           "seg_material", "trk_azim", "trk_num_segments", "quad_phi"}
