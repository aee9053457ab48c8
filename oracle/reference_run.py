"""Binding of oracle/_ref/libpythiaref.so: the reference HashJoinOp and
TupleHasher compiled from the reference's own sources (oracle/ref/Makefile).

Checker and CPU baseline only. The library is built where the reference tree
is present and travels with the repository as a built file; when it is absent
every caller falls back to the oracle restatement.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "_ref", "libpythiaref.so")
CT = {"int": 0, "long": 1, "dec": 2, "char": 3, "date": 4, "pointer": 5}
WIDTH = {"int": 4, "long": 8, "dec": 8, "date": 8, "pointer": 8}

_lib = None


def available():
    return os.path.exists(PATH)


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(PATH)
        L.ref_hashjoin_run.restype = C.c_int
        L.ref_hashjoin_run.argtypes = [
            C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int),
            C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int,
            C.POINTER(C.c_char_p), C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
            C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
            C.POINTER(C.c_uint64), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.ref_tuple_hash.restype = C.c_uint
        L.ref_tuple_hash.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                     C.c_int]
        L.ref_aggregate_run.restype = C.c_int
        L.ref_aggregate_run.argtypes = [
            C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int),
            C.c_int, C.c_char_p, C.c_int, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
            C.POINTER(C.c_uint64), C.POINTER(C.c_uint), C.POINTER(C.c_double)]
        L.ref_hashjoin_consume.restype = C.c_int
        L.ref_hashjoin_consume.argtypes = [
            C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int),
            C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int,
            C.POINTER(C.c_char_p), C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int,
            C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
            C.POINTER(C.c_uint32), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.ref_tuple_hash_bytes.restype = C.c_uint
        L.ref_tuple_hash_bytes.argtypes = [C.c_int, C.c_void_p, C.c_int]
        _lib = L
    return _lib


def default_threads():
    """One thread per core the process may use, at most 16 (the host
    cores that come with one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def _ints(v):
    return (C.c_int * max(len(v), 1))(*v)


def _schema(cols):
    types = [CT[c[0]] if isinstance(c, tuple) else CT[c] for c in cols]
    widths = [c[1] if isinstance(c, tuple) else WIDTH.get(c, 0) for c in cols]
    return types, widths, sum(widths)


def run_join(build_cols, probe_cols, buildjattr, probejattr, projection, build, probe, threads=4,
             fn="modulo", buckets=1024, tuplesperbucket=2, buffsize=1 << 20, range_min=0, range_max=0,
             out_cap=None, prefetch=-1):
    """Runs the reference HashJoinOp under MergeOp(threads). build/probe are
    packed uint8 arrays; prefetch >= 0 sets its probeprefetchstep and
    buildprefetchstep (-1: none, the reference's default). Returns (rows
    uint8 [n, w], build seconds, total seconds)."""
    bt, bw, bsz = _schema(build_cols)
    pt, pw, psz = _schema(probe_cols)
    ow = 0
    for p in projection:
        side, a = p.split("$")
        cols = build_cols if side[-1] == "B" else probe_cols
        c = cols[int(a)]
        ow += c[1] if isinstance(c, tuple) else WIDTH[c]
    nb = build.size // bsz
    npr = probe.size // psz
    cap = out_cap if out_cap is not None else max(npr, 1)
    while True:
        out = np.zeros(cap * ow, np.uint8)
        n = C.c_uint64()
        tb = C.c_double()
        tt = C.c_double()
        proj = (C.c_char_p * len(projection))(*[p.encode() for p in projection])
        rc = lib().ref_hashjoin_run(threads, len(bt), _ints(bt), _ints(bw), len(pt), _ints(pt), _ints(pw),
                                    len(buildjattr), _ints(buildjattr), _ints(probejattr), len(projection), proj,
                                    fn.encode(), buckets, tuplesperbucket, buffsize, range_min, range_max, prefetch,
                                    build.ctypes.data_as(C.c_void_p) if build.size else None, nb,
                                    probe.ctypes.data_as(C.c_void_p) if probe.size else None, npr,
                                    out.ctypes.data_as(C.c_void_p), cap, C.byref(n), C.byref(tb), C.byref(tt))
        if rc == -2 and out_cap is None:
            cap = n.value
            continue
        if rc != 0:
            raise RuntimeError("reference HashJoinOp failed: %d" % rc)
        return out[: n.value * ow].reshape(-1, ow), tb.value, tt.value


def run_pkfk(build, probe, threads, prefetch=-1):
    """The headline join (int64 key, int64 payload) x (int64 fk, int64
    payload), projection [B$0, B$1, P$1], modulo hash over next_pow2(|R|)
    buckets of 2 tuples. Returns seconds and the output count."""
    nb = build.size // 2
    buckets = 1 << max(1, (nb - 1).bit_length())
    rows, tb, tt = run_join(["long", "long"], ["long", "long"], [0], [0], ["B$0", "B$1", "P$1"],
                            build.view(np.uint8), probe.view(np.uint8), threads=threads, fn="modulo",
                            buckets=buckets, tuplesperbucket=2, buffsize=1 << 20, prefetch=prefetch)
    return {"seconds": tt, "build_seconds": tb, "out_tuples": rows.shape[0], "rows": rows}


def run_consume(build_cols, probe_cols, buildjattr, probejattr, projection, build, probe, threads,
                fn="modulo", buckets=1024, tuplesperbucket=2, buffsize=1 << 20, prefetch=-1):
    """The reference plan MergeOp(threads) -> ConsumeOp -> HashJoinOp
    (operators/consume.cpp): every thread drains its own output and XORs its
    32-bit words. Returns (xor32 of all output words, build seconds, total
    seconds)."""
    bt, bw, bsz = _schema(build_cols)
    pt, pw, psz = _schema(probe_cols)
    nb, npr = build.size // bsz, probe.size // psz
    x, tb, tt = C.c_uint32(), C.c_double(), C.c_double()
    proj = (C.c_char_p * len(projection))(*[p.encode() for p in projection])
    rc = lib().ref_hashjoin_consume(threads, len(bt), _ints(bt), _ints(bw), len(pt), _ints(pt), _ints(pw),
                                    len(buildjattr), _ints(buildjattr), _ints(probejattr), len(projection), proj,
                                    fn.encode(), buckets, tuplesperbucket, buffsize, prefetch,
                                    build.ctypes.data_as(C.c_void_p) if build.size else None, nb,
                                    probe.ctypes.data_as(C.c_void_p) if probe.size else None, npr,
                                    C.byref(x), C.byref(tb), C.byref(tt))
    if rc != 0:
        raise RuntimeError("reference HashJoinOp under ConsumeOp failed: %d" % rc)
    return x.value, tb.value, tt.value


def consume_pkfk(build, probe, threads, prefetch=-1):
    """run_consume on the headline join (projection [B$0, B$1, P$1], modulo
    hash over next_pow2(|R|) buckets of 2 tuples, 1 MiB pages)."""
    nb = build.size // 2
    buckets = 1 << max(1, (nb - 1).bit_length())
    x, tb, tt = run_consume(["long", "long"], ["long", "long"], [0], [0], ["B$0", "B$1", "P$1"],
                            build.view(np.uint8), probe.view(np.uint8), threads, fn="modulo", buckets=buckets,
                            tuplesperbucket=2, prefetch=prefetch)
    return {"xor32": x, "build_seconds": tb, "seconds": tt, "probe_seconds": tt - tb}


def tuple_hash(fn, buckets, coltype, value, range_min=0, range_max=0, offset=0, skipbits=0):
    a = np.array([value], dtype=np.int32 if coltype == "int" else np.int64)
    return lib().ref_tuple_hash(fn.encode(), buckets, CT[coltype], a.ctypes.data_as(C.c_void_p), range_min,
                                range_max, offset, skipbits)


def bytes_hash(buckets, data):
    data = np.ascontiguousarray(data, dtype=np.uint8)
    return lib().ref_tuple_hash_bytes(buckets, data.ctypes.data_as(C.c_void_p), data.size)


def run_aggregate(cols, fields, op, sumfield, tuples, threads=4, fn="modulo", buckets=1024):
    """The reference AggregateSum (op "sum") or AggregateCount (op "count")
    in global mode under MergeOp(threads), grouped by `fields` (hash fn over
    the first grouping column; "bytes" over the field range for two, or
    when fn is "bytes").
    Returns (rows uint8 [n, w], seconds)."""
    types, widths, size = _schema(cols)
    data = np.ascontiguousarray(tuples, dtype=np.uint8)
    n = data.size // size if size else 0
    cap = max(n, 1)
    while True:
        out = np.zeros(cap * 64, np.uint8)
        n_out, w, secs = C.c_uint64(), C.c_uint(), C.c_double()
        rc = lib().ref_aggregate_run(threads, 0 if op == "sum" else 1, len(cols), _ints(types), _ints(widths),
                                     len(fields), _ints(list(fields)), sumfield if sumfield is not None else 0,
                                     ("bytes" if len(fields) > 1 else fn).encode(), buckets,
                                     data.ctypes.data_as(C.c_void_p), n, out.ctypes.data_as(C.c_void_p), cap,
                                     C.byref(n_out), C.byref(w), C.byref(secs))
        if rc == -2:
            cap = n_out.value
            continue
        if rc != 0:
            raise RuntimeError("reference aggregate failed")
        ow = w.value
        return out[: n_out.value * ow].reshape(n_out.value, ow), secs.value


def main(argv):
    """python reference_run.py consume BUILD.npy PROBE.npy THREADS PREFETCH
    One run of the headline plan (consume_pkfk) in a process of its own, so
    that each timed run starts with the reference allocator's fresh state
    (its lookaside arenas are used up by the first runs of a process).
    Prints the result as JSON.

    python reference_run.py rows BUILD.npy PROBE.npy THREADS OUT.npy
    The headline plan's output rows (run_pkfk) into OUT.npy, in a process of
    its own for the same reason; prints the seconds as JSON."""
    import json
    usage = ("usage: reference_run.py consume BUILD.npy PROBE.npy THREADS PREFETCH | "
             "rows BUILD.npy PROBE.npy THREADS OUT.npy")
    if len(argv) != 6 or argv[1] not in ("consume", "rows"):
        raise SystemExit(usage)
    b = np.load(argv[2], mmap_mode="r")
    p = np.load(argv[3], mmap_mode="r")
    if argv[1] == "consume":
        print(json.dumps(consume_pkfk(b, p, int(argv[4]), prefetch=int(argv[5]))))
        return
    r = run_pkfk(np.ascontiguousarray(b), np.ascontiguousarray(p), int(argv[4]))
    np.save(argv[5], r.pop("rows"))
    print(json.dumps(r))


if __name__ == "__main__":
    import sys
    main(sys.argv)
