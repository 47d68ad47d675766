"""Reader for the flat binary dumps written by oracle/ref_harness.cpp (test infrastructure)."""
import struct

import numpy as np

_TYPES = {b"d": np.float64, b"f": np.float32, b"i": np.int64}


def read_dump(path):
    """Return {name: numpy array or python scalar} for one dump file."""
    out = {}
    with open(path, "rb") as f:
        if f.read(12) != b"MITHRADUMP1\n":
            raise ValueError(f"{path}: not a MITHRA dump")
        while True:
            head = f.read(4)
            if not head:
                break
            (n,) = struct.unpack("<I", head)
            name = f.read(n).decode()
            t = f.read(1)
            (cnt,) = struct.unpack("<Q", f.read(8))
            dt = np.dtype(_TYPES[t])
            arr = np.frombuffer(f.read(cnt * dt.itemsize), dtype=dt)
            out[name] = arr[0].item() if cnt == 1 and name not in ("power", "undulators") else arr
    return out
