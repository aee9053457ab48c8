"""Helpers shared by the tests: packing tuples like the reference's Schema
(no alignment) and moving them to and from the GPU through torch."""
import numpy as np

FMT = {"int": "<i4", "long": "<i8", "dec": "<f8", "date": "<i8", "pointer": "<u8"}


def fmt(c):
    """numpy field format of a column: a type name, or ("char", width)."""
    return "S%d" % c[1] if isinstance(c, tuple) else FMT[c]


def dtype_of(cols):
    """The packed (unaligned) numpy record type of a schema."""
    return np.dtype([("f%d" % i, fmt(c)) for i, c in enumerate(cols)])


def pack(cols, rows):
    """cols: type names; rows: list of tuples. Returns a uint8 array."""
    dt = dtype_of(cols)
    a = np.zeros(len(rows), dtype=dt)
    for i, r in enumerate(rows):
        a[i] = tuple(r)
    return a.view(np.uint8).reshape(-1).copy()


def unpack(cols, buf):
    dt = dtype_of(cols)
    return np.frombuffer(np.ascontiguousarray(buf).reshape(-1).tobytes(), dtype=dt)


def to_gpu(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).reshape(-1).view(np.uint8)).to("cuda")


def from_gpu(t, nbytes):
    return t[:nbytes].cpu().numpy()
