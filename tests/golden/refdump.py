"""Reader for the tagged binary records written by oracle/ref_dump.c.

Each record is: int32 name length, name bytes, one type byte ('d' float64 or
'i' int32), int64 element count, then the elements (little endian).
"""
import numpy as np


def read_dump(path):
    out = {}
    with open(path, "rb") as f:
        buf = f.read()
    off = 0
    while off < len(buf):
        (ln,) = np.frombuffer(buf, np.int32, 1, off)
        off += 4
        name = buf[off:off + ln].decode()
        off += ln
        t = chr(buf[off])
        off += 1
        (cnt,) = np.frombuffer(buf, np.int64, 1, off)
        off += 8
        dt = np.float64 if t == "d" else np.int32
        arr = np.frombuffer(buf, dt, int(cnt), off).copy()
        off += arr.nbytes
        out[name] = arr
    return out
