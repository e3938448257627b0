"""Reader and writer of .conan files.

Layout (src/base/misc.hh:27-78, src/base/array.hh:75-105,
src/base/header.cc:30-74, src/base/history.cc:25-69; also read by the
reference's scripts/thulsa.py:26-82): blocks framed by a uint64 byte count
before and after; block 1 is the parameter header ("key=value\\n", keys sorted),
block 2 the history ("time:command\\n"), then (record header, data) pairs where
the record header carries name= and dtype= (files written by thulsa carry
type= instead; only name= is needed to find a record).
"""
from __future__ import annotations

import struct
import time

import numpy as np

DTYPES = {
    "int32": (np.int32, 1), "uint32": (np.uint32, 1), "float32": (np.float32, 1),
    "float64": (np.float64, 1), "complex64": (np.complex64, 1), "complex128": (np.complex128, 1),
    "3i4": (np.int32, 3), "2i4": (np.int32, 2), "3f4": (np.float32, 3), "2f4": (np.float32, 2),
    "3f8": (np.float64, 3), "2f8": (np.float64, 2),
}


class ConanFormatError(RuntimeError):
    pass


def read_block(f) -> bytes:
    s = f.read(8)
    if len(s) < 8:
        raise ConanFormatError("Could not read from file")
    n = struct.unpack("<Q", s)[0]
    data = f.read(n)
    t = f.read(8)
    if len(data) != n or len(t) != 8 or struct.unpack("<Q", t)[0] != n:
        raise ConanFormatError("The block seems to be corrupted")
    return data


def write_block(f, data: bytes) -> None:
    f.write(struct.pack("<Q", len(data)))
    f.write(data)
    f.write(struct.pack("<Q", len(data)))


def parse_header(raw: bytes) -> dict:
    out = {}
    for line in raw.decode().split("\n"):
        if not line:
            continue
        k, _, v = line.partition("=")
        out.setdefault(k, v)
    return out


def format_header(h: dict) -> bytes:
    return "".join(f"{k}={str(h[k]).split(chr(10))[0]}\n" for k in sorted(h)).encode()


def parse_history(raw: bytes) -> list:
    out = []
    for line in raw.decode().split("\n"):
        if not line:
            continue
        t, _, c = line.partition(":")
        out.append((int(t), c))
    return sorted(out, key=lambda x: x[0])


def format_history(hist) -> bytes:
    return "".join(f"{t}:{c.split(chr(10))[0]}\n" for t, c in sorted(hist, key=lambda x: x[0])).encode()


def read(path: str):
    """Returns (header, history, records) with records = {name: (subheader, array)}."""
    with open(path, "rb") as f:
        header = parse_header(read_block(f))
        history = parse_history(read_block(f))
        records = {}
        while True:
            pos = f.tell()
            if not f.read(1):
                break
            f.seek(pos)
            sub = parse_header(read_block(f))
            raw = read_block(f)
            dt = sub.get("dtype", sub.get("type", "float64"))
            base, width = DTYPES.get(dt, (np.float64, 1))
            arr = np.frombuffer(raw, dtype=base)
            if width > 1:
                arr = arr.reshape(-1, width)
            records[sub.get("name", "")] = (sub, arr)
    return header, history, records


def write(path: str, header: dict, history, records: dict) -> None:
    """records: {name: (dtype_name, array)} written in insertion order."""
    with open(path, "wb") as f:
        write_block(f, format_header(header))
        write_block(f, format_history(history))
        for name, (dt, arr) in records.items():
            base, width = DTYPES[dt]
            write_block(f, format_header({"name": name, "dtype": dt}))
            write_block(f, np.ascontiguousarray(arr, dtype=base).tobytes())


def write_noise(path: str, noise, dim: int, mbits: int, seed_note: str = "stored white noise") -> None:
    """A stored white-noise field for ``conan ic --noise path``."""
    N = 1 << mbits
    write(path, {"dim": dim, "mbits": mbits, "N": N}, [(int(time.time()), seed_note)],
          {"noise": ("float64", np.asarray(noise, dtype=np.float64).ravel())})
