"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of the
reference build in oracle/_ref when present.

Checker only: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
use it; the product never does.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libpythiaref.so")

CT = {"int": 0, "long": 1, "dec": 2, "char": 3, "date": 4, "pointer": 5}
FN = {"modulo": 1, "parammodulo": 2, "knuth": 3, "range": 4, "exactrange": 5, "bytes": 6,
      "tpchorderkey": 7, "willis": 8, "alwayszero": 9, "tpchq1magic": 10}
MAXC = 32


class Schema(C.Structure):
    _fields_ = [("ncols", C.c_int), ("type", C.c_int * MAXC), ("width", C.c_int * MAXC),
                ("offset", C.c_int * MAXC), ("size", C.c_int)]


class Hasher(C.Structure):
    _fields_ = [("fn", C.c_int), ("k", C.c_int64), ("min", C.c_int64), ("max", C.c_int64),
                ("bucketrange", C.c_int64), ("skipbits", C.c_uint), ("offset", C.c_uint),
                ("size", C.c_uint)]


class JoinDesc(C.Structure):
    _fields_ = [("build", Schema), ("probe", Schema), ("nkeys", C.c_int),
                ("buildjattr", C.c_uint16 * 8), ("probejattr", C.c_uint16 * 8),
                ("nproj", C.c_int), ("proj_side", C.c_int * MAXC), ("proj_attr", C.c_int * MAXC),
                ("hash_fn", C.c_int), ("buckets", C.c_uint), ("range_min", C.c_int64),
                ("range_max", C.c_int64), ("hash_offset", C.c_uint), ("skipbits", C.c_uint),
                ("tuplesperbucket", C.c_uint), ("buffsize", C.c_uint)]


class JoinStats(C.Structure):
    _fields_ = [("out_tuples", C.c_uint64), ("pages", C.c_uint64), ("spills", C.c_uint64),
                ("out_tuple_size", C.c_uint)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: make -f oracle/Makefile")
        L = C.CDLL(LIB_PATH)
        sig = {
            "or_schema_init": (C.c_int, [C.POINTER(Schema), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "or_getlogarithm": (C.c_uint, [C.c_uint]),
            "or_hasher_init": (C.c_int, [C.POINTER(Hasher), C.c_int, C.c_uint, C.c_int64, C.c_int64,
                                         C.c_uint, C.c_uint]),
            "or_hasher_buckets": (C.c_uint, [C.POINTER(Hasher)]),
            "or_hasher_hash": (C.c_uint, [C.POINTER(Hasher), C.c_void_p, C.c_size_t]),
            "or_hasher_hash_value": (C.c_uint, [C.POINTER(Hasher), C.c_int64]),
            "or_exactrange_minimumforbucket": (C.c_int64, [C.POINTER(Hasher), C.c_uint]),
            "or_parammodulo_generate": (C.c_uint, [C.c_int64, C.c_uint, C.c_uint, C.c_uint,
                                                   C.POINTER(C.c_uint), C.POINTER(C.c_uint)]),
            "or_find_max_contiguous": (C.c_uint, [C.POINTER(C.c_uint16), C.POINTER(C.c_uint16), C.c_uint]),
            "or_join_out_width": (C.c_int, [C.POINTER(JoinDesc)]),
            "or_hashjoin": (C.c_int, [C.POINTER(JoinDesc), C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
                                      C.c_void_p, C.c_uint64, C.POINTER(JoinStats)]),
            "or_nested_loop_join": (C.c_int, [C.POINTER(JoinDesc), C.c_void_p, C.c_uint64, C.c_void_p,
                                              C.c_uint64, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
            "or_partition": (C.c_int, [C.POINTER(Schema), C.c_int, C.c_int64, C.c_int64, C.c_uint,
                                       C.c_void_p, C.POINTER(C.c_uint64), C.c_void_p, C.POINTER(C.c_uint64)]),
            "or_tuple_mix": (C.c_uint64, [C.c_void_p, C.c_uint]),
            "or_checksum": (C.c_uint64, [C.c_void_p, C.c_uint64, C.c_uint]),
            "or_xor32": (C.c_uint32, [C.c_void_p, C.c_uint64, C.c_uint]),
            "or_expected_pkfk": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                           C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
            "or_gen_pkfk": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                      C.c_void_p, C.c_void_p]),
            "or_expected_pairs_zipf": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p, C.c_int,
                                                 C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
            "or_gen_pairs_zipf": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p,
                                            C.c_void_p]),
            "or_expected_clustered": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_int,
                                                C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
            "or_zipf_ranks_guided": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p]),
            "or_gen_pkfk_probe_range": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                                  C.c_uint64, C.c_uint64, C.c_void_p]),
            "or_expected_batch_join": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64, C.c_int,
                                                 C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def schema(cols):
    s = Schema()
    n = len(cols)
    types = (C.c_int * MAXC)()
    widths = (C.c_int * MAXC)()
    for i, c in enumerate(cols):
        if isinstance(c, tuple):
            types[i], widths[i] = CT[c[0]], c[1]
        else:
            types[i] = CT[c]
    if lib().or_schema_init(C.byref(s), n, types, widths) != 0:
        raise ValueError("bad schema")
    return s


def hasher(fn, buckets, range_min=0, range_max=0, offset=0, skipbits=0):
    h = Hasher()
    if lib().or_hasher_init(C.byref(h), FN[fn], buckets, range_min, range_max, offset, skipbits) != 0:
        raise ValueError("bad hasher")
    return h


def join_desc(build_cols, probe_cols, buildjattr, probejattr, projection, fn="modulo", buckets=16,
              tuplesperbucket=2, buffsize=1 << 20, range_min=0, range_max=0, offset=0, skipbits=0):
    d = JoinDesc()
    d.build = schema(build_cols)
    d.probe = schema(probe_cols)
    d.nkeys = len(buildjattr)
    for i, (a, b) in enumerate(zip(buildjattr, probejattr)):
        d.buildjattr[i] = a
        d.probejattr[i] = b
    d.nproj = len(projection)
    for i, p in enumerate(projection):
        side, attr = p.split("$")
        d.proj_side[i] = 0 if side[-1] == "B" else 1
        d.proj_attr[i] = int(attr)
    d.hash_fn = FN[fn]
    d.buckets = buckets
    d.range_min, d.range_max = range_min, range_max
    d.hash_offset, d.skipbits = offset, skipbits
    d.tuplesperbucket = tuplesperbucket
    d.buffsize = buffsize
    return d


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None and a.size else None


def hashjoin(desc, build, probe, out_cap=None):
    """build/probe: contiguous uint8 arrays of packed tuples. Returns
    (out uint8 array [n, width], stats)."""
    w = lib().or_join_out_width(C.byref(desc))
    nb = build.size // desc.build.size if desc.build.size else 0
    npr = probe.size // desc.probe.size if desc.probe.size else 0
    cap = out_cap if out_cap is not None else max(npr, 1)
    while True:
        out = np.zeros(cap * w, dtype=np.uint8)
        st = JoinStats()
        rc = lib().or_hashjoin(C.byref(desc), _ptr(build), nb, _ptr(probe), npr, _ptr(out), cap, C.byref(st))
        if rc == -2 and out_cap is None:
            cap *= 4
            continue
        if rc != 0:
            raise RuntimeError("or_hashjoin failed: %d" % rc)
        return out[: st.out_tuples * w].reshape(-1, w), st


def nested_loop_join(desc, build, probe):
    w = lib().or_join_out_width(C.byref(desc))
    nb = build.size // desc.build.size
    npr = probe.size // desc.probe.size
    cap = max(nb * npr, 1)
    out = np.zeros(cap * w, dtype=np.uint8)
    n = C.c_uint64()
    rc = lib().or_nested_loop_join(C.byref(desc), _ptr(build), nb, _ptr(probe), npr, _ptr(out), cap, C.byref(n))
    if rc != 0:
        raise RuntimeError("or_nested_loop_join failed: %d" % rc)
    return out[: n.value * w].reshape(-1, w)


def partition(sch, attr, rmin, rmax, thread_inputs):
    """thread_inputs: list of uint8 arrays (one per thread). Returns list of
    per-partition uint8 arrays."""
    T = len(thread_inputs)
    counts = (C.c_uint64 * T)(*[a.size // sch.size for a in thread_inputs])
    allin = np.concatenate(thread_inputs) if T else np.zeros(0, np.uint8)
    out = np.zeros_like(allin)
    sizes = (C.c_uint64 * T)()
    rc = lib().or_partition(C.byref(sch), attr, rmin, rmax, T, _ptr(allin), counts, _ptr(out), sizes)
    if rc != 0:
        raise RuntimeError("or_partition failed")
    parts, o = [], 0
    for j in range(T):
        parts.append(out[o * sch.size:(o + sizes[j]) * sch.size])
        o += sizes[j]
    return parts


def checksum(rows, width):
    rows = np.ascontiguousarray(rows, dtype=np.uint8).reshape(-1)
    n = rows.size // width
    return lib().or_checksum(_ptr(rows), n, width), lib().or_xor32(_ptr(rows), n, width)


def expected_pkfk(seed, domain, n_build, n_probe, n_miss, threads=8):
    s, c = C.c_uint64(), C.c_uint64()
    lib().or_expected_pkfk(seed, domain, n_build, n_probe, n_miss, threads, C.byref(s), C.byref(c))
    return s.value, c.value


def expected_batch_join(seed, j, n_build, n_probe, threads=8):
    """(checksum, count) of join j of the batch workload."""
    s, c = C.c_uint64(), C.c_uint64()
    lib().or_expected_batch_join(seed, j, n_build, n_probe, threads, C.byref(s), C.byref(c))
    return s.value, c.value


def gen_pkfk(seed, domain, n_build, n_probe, n_miss, probe_prefix=None):
    """Build relation and probe relation of the seeded PK-FK workload, as
    (key, payload) uint64 pairs. probe_prefix: return only the first that
    many probe tuples of the full probe relation."""
    b = np.zeros(n_build * 2, dtype=np.uint64)
    if probe_prefix is None:
        p = np.zeros(n_probe * 2, dtype=np.uint64)
        lib().or_gen_pkfk(seed, domain, n_build, n_probe, n_miss, _ptr(b), _ptr(p))
        return b, p
    lib().or_gen_pkfk(seed, domain, n_build, 0, 0, _ptr(b), None)
    p = np.zeros(probe_prefix * 2, dtype=np.uint64)
    lib().or_gen_pkfk_probe_range(seed, domain, n_build, n_probe, n_miss, 0, probe_prefix, _ptr(p))
    return b, p


def sorted_rows(rows):
    """Rows (n x w uint8) in a canonical order, for multiset comparison."""
    rows = np.ascontiguousarray(rows)
    if rows.size == 0:
        return rows
    v = rows.view(np.dtype((np.void, rows.shape[1]))).ravel()
    return rows[np.argsort(v, kind="stable")]


def zipf_cdf(n, z):
    """The Zipf CDF as the reference's generator builds it
    (operators/generator_zipf.cpp:91-113): running sum of (i+1)^-z in double,
    then divided by the total. Computed here independently of the product."""
    i = np.arange(1, n + 1, dtype=np.float64)
    cdf = np.cumsum(i ** (-z))
    return cdf / cdf[-1]


def gen_pairs_zipf(seed, n_build, n_probe, cdf):
    b = np.zeros(n_build * 16, np.uint8)
    p = np.zeros(n_probe * 16, np.uint8)
    lib().or_gen_pairs_zipf(seed, n_build, n_probe, cdf.ctypes.data_as(C.c_void_p), _ptr(b), _ptr(p))
    return b, p


def expected_pairs_zipf(seed, n_build, n_probe, cdf, threads=8):
    s, c = C.c_uint64(), C.c_uint64()
    lib().or_expected_pairs_zipf(seed, n_build, n_probe, cdf.ctypes.data_as(C.c_void_p), threads, C.byref(s),
                                 C.byref(c))
    return s.value, c.value


def expected_clustered(table, n_probe, cdf, threads=8):
    s, c = C.c_uint64(), C.c_uint64()
    lib().or_expected_clustered(table.ctypes.data_as(C.c_void_p), n_probe, cdf.ctypes.data_as(C.c_void_p), threads,
                                C.byref(s), C.byref(c))
    return s.value, c.value


_NP = {"int": "<i4", "long": "<i8", "date": "<i8", "dec": "<f8"}


def aggregate(cols, fields, op, sumfield=None, tuples=None):
    """The reference's global-mode aggregation over packed tuples of schema
    cols: GenericAggregate::remember (operators/genericaggregate.cpp:247-289)
    keeps one output tuple per distinct value of the grouping columns
    `fields` (the whole input is one group when there are none; an empty
    input gives no tuple), its schema the grouping columns followed by the
    fold's (genericaggregate.cpp:51-95). op "sum": AggregateSum
    (operators/aggregatesum.cpp:37-107) adds column sumfield in that column's
    type, INTEGER sums wrapping at 32 bits and LONG sums at 64; op "count":
    AggregateCount (operators/aggregatecount.cpp:37-52) counts tuples as a
    LONG. Returns the output tuples (n x width uint8) in no particular order.
    DECIMAL sums depend on the order of the additions and are not restated.
    """
    dt = np.dtype([("c%d" % i, _NP[c]) for i, c in enumerate(cols)])
    rows = np.ascontiguousarray(tuples, dtype=np.uint8).reshape(-1)
    t = rows.view(dt) if rows.size else np.zeros(0, dt)
    if op == "sum":
        if cols[sumfield] not in ("int", "long"):
            raise ValueError("sums of INTEGER or LONG columns")
        agg_t = _NP[cols[sumfield]]
    elif op == "count":
        agg_t = "<i8"
    else:
        raise ValueError("op is sum or count")
    out_dt = np.dtype([("g%d" % k, _NP[cols[f]]) for k, f in enumerate(fields)] + [("a", agg_t)])
    if t.size == 0:
        return np.zeros((0, out_dt.itemsize), np.uint8)
    if fields:
        keys = np.zeros(t.size, np.dtype([("g%d" % k, _NP[cols[f]]) for k, f in enumerate(fields)]))
        for k, f in enumerate(fields):
            keys["g%d" % k] = t["c%d" % f]
        uniq, inv = np.unique(keys, return_inverse=True)
        inv = inv.reshape(-1)
        ng = uniq.size
    else:
        uniq, inv, ng = None, np.zeros(t.size, np.int64), 1
    if op == "count":
        acc = np.bincount(inv, minlength=ng).astype(np.int64)
    else:
        acc = np.zeros(ng, np.int64)
        np.add.at(acc, inv, t["c%d" % sumfield].astype(np.int64))   # wraps at 64 bits
        if agg_t == "<i4":
            acc = acc.astype(np.int32)   # the low 32 bits: the INTEGER sum's wrap
    out = np.zeros(ng, out_dt)
    for k in range(len(fields)):
        out["g%d" % k] = uniq["g%d" % k]
    out["a"] = acc
    return out.view(np.uint8).reshape(ng, out_dt.itemsize)
