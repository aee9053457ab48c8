"""ctypes binding of libpjoin.so (include/pjoin.h).

This is the binding a Python caller (tests, bench.py) uses; the C++ operators
in host/ call the same C ABI directly. Loading fails loudly when the library
has not been built: there is no fallback path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# PJOIN_LIB selects another build of the library (default: the in-tree one).
LIB_PATH = os.environ.get("PJOIN_LIB") or os.path.join(HERE, "lib", "libpjoin.so")

PJ_OK = 0
PJ_ERR_INVALID = -1
PJ_ERR_UNSUPPORTED = -2
PJ_ERR_DEVICE = -3
PJ_ERR_CAPACITY = -4
PJ_ERR_STATE = -5

CT = {"int": 0, "long": 1, "dec": 2, "char": 3, "date": 4, "pointer": 5}
WIDTH = {"int": 4, "long": 8, "dec": 8, "date": 8, "pointer": 8}
PJ_MAX_COLS = 32
PJ_MAX_KEYS = 8
PART = {"radix": 0, "learned": 1, "auto": 2}
PJ_FLAG_TIMING = 1
PJ_FLAG_NARROW_ROW_HASH = 2
PJ_FLAG_LAYOUT_EXACT = 4
PJ_FLAG_LAYOUT_APPEND = 8
PJ_FLAG_LAYOUT_TWO_PASS = 16


class Schema(C.Structure):
    _fields_ = [("ncols", C.c_uint32), ("type", C.c_uint32 * PJ_MAX_COLS),
                ("width", C.c_uint32 * PJ_MAX_COLS)]


class JoinDesc(C.Structure):
    _fields_ = [("build", Schema), ("probe", Schema), ("nkeys", C.c_uint32),
                ("buildjattr", C.c_uint16 * PJ_MAX_KEYS), ("probejattr", C.c_uint16 * PJ_MAX_KEYS),
                ("nproj", C.c_uint32), ("proj_side", C.c_uint8 * PJ_MAX_COLS),
                ("proj_attr", C.c_uint16 * PJ_MAX_COLS), ("partitioner", C.c_uint32),
                ("radix_bits", C.c_int32), ("flags", C.c_uint32)]


class Timings(C.Structure):
    _fields_ = [("build_partition_ms", C.c_float), ("probe_partition_ms", C.c_float),
                ("join_ms", C.c_float), ("model_ms", C.c_float), ("total_ms", C.c_float),
                ("partitions", C.c_uint32), ("subpartitions", C.c_uint32),
                ("work_items", C.c_uint32), ("used_learned", C.c_uint32),
                ("spilled_items", C.c_uint64), ("build_hist_ms", C.c_float), ("probe_hist_ms", C.c_float),
                ("probe_appended", C.c_uint32), ("probe_passes", C.c_uint32)]


PJ_MODEL_LEAVES = 4096
PJ_MODEL_MAX_SEGS = 1024


class Model(C.Structure):
    """pj_model: the learned partitioner's two-stage piecewise-linear CDF."""
    _fields_ = [("nleaves", C.c_uint32), ("nseg", C.c_uint32), ("n", C.c_uint64), ("n_keys", C.c_uint64),
                ("window", C.c_uint32), ("iters", C.c_uint32), ("lam", C.c_float), ("loss_lsq", C.c_float),
                ("loss", C.c_float), ("reserved", C.c_uint32),
                ("seg_lo", C.c_uint64 * PJ_MODEL_MAX_SEGS), ("seg_hi", C.c_uint64 * PJ_MODEL_MAX_SEGS),
                ("seg_lw", C.c_uint64 * PJ_MODEL_MAX_SEGS), ("seg_leaf0", C.c_uint32 * (PJ_MODEL_MAX_SEGS + 1)),
                ("leaf_seg", C.c_uint16 * PJ_MODEL_LEAVES), ("leaf_start", C.c_uint32 * (PJ_MODEL_LEAVES + 1)),
                ("a", C.c_float * PJ_MODEL_LEAVES), ("b", C.c_float * PJ_MODEL_LEAVES)]


class BatchJoin(C.Structure):
    _fields_ = [("build", C.c_void_p), ("nbuild", C.c_uint64), ("probe", C.c_void_p), ("nprobe", C.c_uint64),
                ("out", C.c_void_p), ("out_cap", C.c_uint64)]


class BatchCfg(C.Structure):
    _fields_ = [("bits", C.c_uint8), ("passes", C.c_uint8), ("piece", C.c_uint16), ("load_factor", C.c_float)]


class AggDesc(C.Structure):
    _fields_ = [("schema", Schema), ("op", C.c_uint32), ("nfields", C.c_uint32), ("fields", C.c_uint16 * 2),
                ("sumfield", C.c_uint16), ("reserved", C.c_uint16), ("max_groups", C.c_uint64)]


class Cursor(C.Structure):
    _fields_ = [("impl", C.c_void_p), ("pending", C.c_uint64)]


class PJError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("pjoin error %d: %s" % (code, msg))
        self.code = code
        self.msg = msg


_lib = None

_SIGS = {
    "pj_last_error": (C.c_char_p, []),
    "pj_version": (C.c_char_p, []),
    "pj_join_create": (C.c_int, [C.POINTER(JoinDesc), C.POINTER(C.c_void_p)]),
    "pj_join_out_width": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "pj_join_destroy": (C.c_int, [C.c_void_p]),
    "pj_join_run_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
                                     C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64), C.c_void_p]),
    "pj_join_run_device_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
                                     C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64), C.c_void_p]),
    "pj_join_build_append": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "pj_join_build_finish": (C.c_int, [C.c_void_p]),
    "pj_join_probe": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(Cursor), C.c_void_p,
                                C.c_uint64, C.POINTER(C.c_uint64)]),
    "pj_cursor_release": (C.c_int, [C.POINTER(Cursor)]),
    "pj_join_reset": (C.c_int, [C.c_void_p]),
    "pj_join_last_timings": (C.c_int, [C.c_void_p, C.POINTER(Timings)]),
    "pj_partition_exactrange_device": (C.c_int, [C.POINTER(Schema), C.c_uint32, C.c_int64, C.c_int64,
                                                 C.c_uint32, C.c_void_p, C.c_uint64, C.c_void_p,
                                                 C.POINTER(C.c_uint64), C.c_void_p]),
    "pj_checksum_device": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(C.c_uint64),
                                     C.POINTER(C.c_uint32), C.c_void_p]),
    "pj_gen_pkfk": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                              C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "pj_partition_exactrange": (C.c_int, [C.POINTER(Schema), C.c_uint32, C.c_int64, C.c_int64, C.c_uint32,
                                          C.c_void_p, C.c_uint64, C.c_void_p, C.POINTER(C.c_uint64)]),
    "pj_query_run": (C.c_int, [C.c_char_p, C.c_uint32, C.POINTER(C.c_char_p), C.POINTER(C.c_void_p),
                               C.POINTER(C.c_uint64), C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64),
                               C.POINTER(C.c_uint32)]),
    "pj_set_error": (None, [C.c_char_p]),
    "pj_zipf_cdf": (C.c_int, [C.c_uint64, C.c_double, C.c_void_p]),
    "pj_clusters_make": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32, C.c_double,
                                   C.c_void_p]),
    "pj_gen_clustered": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                   C.c_void_p]),
    "pj_cluster_table_size": (C.c_uint64, []),
    "pj_gen_pairs_zipf": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_int, C.c_void_p]),
    "pj_batch_create": (C.c_int, [C.POINTER(JoinDesc), C.POINTER(C.c_void_p)]),
    "pj_batch_destroy": (C.c_int, [C.c_void_p]),
    "pj_batch_set_config": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(BatchCfg)]),
    "pj_batch_get_config": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(BatchCfg)]),
    "pj_batch_tune": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(BatchJoin), C.c_uint32, C.c_void_p]),
    "pj_batch_prepare": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(BatchJoin), C.c_void_p]),
    "pj_batch_run": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "pj_batch_results": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pj_batch_run_times": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pj_agg_create": (C.c_int, [C.POINTER(AggDesc), C.POINTER(C.c_void_p)]),
    "pj_agg_out_width": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "pj_binary_tuples": (C.c_int, [C.c_char_p, C.c_uint32, C.POINTER(C.c_uint64)]),
    "pj_hash_buckets": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "pj_set_device_limit": (C.c_int, [C.c_uint64]),
    "pj_measure": (C.c_int, [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32,
                             C.c_void_p, C.c_void_p]),
    "pj_device_bytes": (C.c_int, [C.POINTER(C.c_uint64)]),
    "pj_hash_generate": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "pj_tuple_hash_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32,
                                       C.c_void_p, C.c_void_p]),
    "pj_join_count_by_key": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p,
                                       C.c_uint64, C.POINTER(C.c_uint64), C.c_void_p]),
    "pj_smj_create": (C.c_int, [C.POINTER(JoinDesc), C.POINTER(C.c_void_p)]),
    "pj_smj_out_width": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "pj_smj_run_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p,
                                    C.c_uint64, C.POINTER(C.c_uint64), C.c_void_p]),
    "pj_smj_last_timings": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pj_smj_destroy": (C.c_int, [C.c_void_p]),
    "pj_load_binary": (C.c_int, [C.c_char_p, C.c_uint32, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64),
                                 C.c_void_p]),
    "pj_agg_run_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
                                    C.POINTER(C.c_uint64), C.c_void_p]),
    "pj_agg_destroy": (C.c_int, [C.c_void_p]),
    "pj_batch_sizes": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64, C.c_void_p]),
    "pj_gen_batch_join": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p,
                                    C.c_int, C.c_void_p]),
    "pj_model_size": (C.c_uint64, []),
    "pj_sort_u64": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p]),
    "pj_model_fit": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint32, C.POINTER(Model), C.c_void_p]),
    "pj_model_grad": (C.c_int, [C.POINTER(Model), C.c_void_p, C.c_uint64, C.c_void_p, C.POINTER(C.c_double),
                                C.c_void_p]),
    "pj_model_bounds": (C.c_int, [C.POINTER(Model), C.c_uint32, C.c_void_p, C.c_uint32,
                                  C.POINTER(C.c_uint32)]),
}

# Entry points added by later components; bound when the library has them.
_OPTIONAL_SIGS = {}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libpjoin.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        # torch ships its own HIP runtime under another file name. Loaded
        # first, it is the runtime libpjoin binds to, so the process has one
        # and torch's device memory and streams are valid in the library;
        # loaded after libpjoin, a second runtime would come up beside it.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in list(_SIGS.items()) + list(_OPTIONAL_SIGS.items()):
            if not hasattr(L, name):
                if name in _SIGS:
                    raise RuntimeError("libpjoin.so lacks %s" % name)
                continue
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != PJ_OK:
        raise PJError(rc, lib().pj_last_error().decode())
    return rc


def make_schema(cols):
    """cols: list of type names ("long", "int", "dec", "date", "pointer") or
    ("char", width) tuples."""
    s = Schema()
    s.ncols = len(cols)
    for i, c in enumerate(cols):
        if isinstance(c, tuple):
            s.type[i] = CT[c[0]]
            s.width[i] = c[1]
        else:
            s.type[i] = CT[c]
            s.width[i] = WIDTH.get(c, 0)
    return s


def make_desc(build_cols, probe_cols, buildjattr, probejattr, projection,
              partitioner="radix", radix_bits=-1, timing=False, flags=0):
    """projection: list of "B$i" / "P$j" strings, as in a plan file; flags:
    PJ_FLAG_* bits besides PJ_FLAG_TIMING."""
    d = JoinDesc()
    d.build = make_schema(build_cols)
    d.probe = make_schema(probe_cols)
    d.nkeys = len(buildjattr)
    for i, (a, b) in enumerate(zip(buildjattr, probejattr)):
        d.buildjattr[i] = a
        d.probejattr[i] = b
    d.nproj = len(projection)
    for i, p in enumerate(projection):
        side, attr = p.split("$")
        d.proj_side[i] = 0 if side[-1] == "B" else 1
        d.proj_attr[i] = int(attr)
    d.partitioner = PART[partitioner]
    d.radix_bits = radix_bits
    d.flags = (PJ_FLAG_TIMING if timing else 0) | flags
    return d


class Join:
    """Owns a pj_join_t handle."""

    def __init__(self, desc):
        self.h = C.c_void_p()
        check(lib().pj_join_create(C.byref(desc), C.byref(self.h)))
        w = C.c_uint32()
        check(lib().pj_join_out_width(self.h, C.byref(w)))
        self.out_width = w.value

    def run_device(self, d_build, nbuild, d_probe, nprobe, d_out, out_cap, stream=None):
        n = C.c_uint64()
        rc = lib().pj_join_run_device(self.h, C.c_void_p(d_build), nbuild, C.c_void_p(d_probe), nprobe,
                                      C.c_void_p(d_out), out_cap, C.byref(n), C.c_void_p(stream))
        if rc == PJ_ERR_CAPACITY:
            raise PJError(rc, "need %d output tuples" % n.value)
        check(rc)
        return n.value

    def run_device_noalloc(self, d_build, nbuild, d_probe, nprobe, d_out, out_cap, stream=None):
        """pj_join_run_device_noalloc: the same join in the library's fixed
        device scratch (no device allocation)."""
        n = C.c_uint64()
        rc = lib().pj_join_run_device_noalloc(self.h, C.c_void_p(d_build), nbuild, C.c_void_p(d_probe), nprobe,
                                              C.c_void_p(d_out), out_cap, C.byref(n), C.c_void_p(stream))
        if rc == PJ_ERR_CAPACITY:
            raise PJError(rc, "need %d output tuples" % n.value)
        check(rc)
        return n.value

    def count_by_key(self, d_build, nbuild, d_probe, nprobe, d_groups, groups_cap, stream=None):
        """aggregate_count grouped by the build key over the join, without the
        join's rows; returns the number of groups."""
        g = C.c_uint64()
        rc = lib().pj_join_count_by_key(self.h, C.c_void_p(d_build), nbuild, C.c_void_p(d_probe), nprobe,
                                        C.c_void_p(d_groups), groups_cap, C.byref(g), C.c_void_p(stream))
        if rc == PJ_ERR_CAPACITY:
            raise PJError(rc, "need %d groups" % g.value)
        check(rc)
        return g.value

    def timings(self):
        t = Timings()
        check(lib().pj_join_last_timings(self.h, C.byref(t)))
        return t

    def close(self):
        if self.h:
            lib().pj_join_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Owns a pj_batch_t: many joins of one descriptor, run as one pipeline
    (PLAIN, or GRAPH: one graph launch) or as whole single joins
    (SEQUENTIAL, CONCURRENT, CONCURRENT_GRAPH)."""

    def __init__(self, desc):
        self.h = C.c_void_p()
        check(lib().pj_batch_create(C.byref(desc), C.byref(self.h)))
        self.n = 0

    @staticmethod
    def _joins(joins):
        arr = (BatchJoin * len(joins))()
        for i, (b, nb, p, np_, o, cap) in enumerate(joins):
            arr[i] = BatchJoin(b, nb, p, np_, o, cap)
        return arr

    def tune(self, joins, trial=4, stream=None):
        arr = self._joins(joins)
        check(lib().pj_batch_tune(self.h, len(joins), arr, trial, C.c_void_p(stream)))

    PIECE_UNIT = 4096   # probe records of a work item are set in these units

    def config(self, size_class):
        """(radix bits, passes, load factor, probe records a work item); 0
        bits: from the load factor, or the single join's choice when that is
        0; 0 probe records: the default."""
        c = BatchCfg()
        check(lib().pj_batch_get_config(self.h, size_class, C.byref(c)))
        return c.bits, c.passes, c.load_factor, c.piece * self.PIECE_UNIT

    def set_config(self, size_class, bits, passes=1, load_factor=0.0, piece=0):
        c = BatchCfg(bits, passes, piece // self.PIECE_UNIT, load_factor)
        check(lib().pj_batch_set_config(self.h, size_class, C.byref(c)))

    def prepare(self, joins, stream=None):
        """joins: list of (d_build, nbuild, d_probe, nprobe, d_out, out_cap)."""
        arr = self._joins(joins)
        check(lib().pj_batch_prepare(self.h, len(joins), arr, C.c_void_p(stream)))
        self.n = len(joins)

    PLAIN, GRAPH, SEQUENTIAL, CONCURRENT, CONCURRENT_GRAPH = 0, 1, 2, 3, 4

    def run(self, mode=GRAPH, stream=None):
        check(lib().pj_batch_run(self.h, mode, C.c_void_p(stream)))

    def results(self, stream=None):
        import numpy as np
        cnt = np.zeros(self.n, np.uint64)
        done = np.zeros(self.n, np.float64)
        check(lib().pj_batch_results(self.h, cnt.ctypes.data_as(C.c_void_p), done.ctypes.data_as(C.c_void_p),
                                     C.c_void_p(stream)))
        return cnt, done

    def run_times(self):
        """Each join's own device time (us) in the last sequential or
        concurrent run."""
        import numpy as np
        t = np.zeros(self.n, np.float64)
        check(lib().pj_batch_run_times(self.h, t.ctypes.data_as(C.c_void_p)))
        return t

    def close(self):
        if self.h:
            lib().pj_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_sizes(seed, njoins, nmin, nmax):
    import numpy as np
    n = np.zeros(njoins, np.uint64)
    check(lib().pj_batch_sizes(seed, njoins, nmin, nmax, n.ctypes.data_as(C.c_void_p)))
    return n


def run_plan(plan, sources, out_cap_bytes=1 << 20):
    """Runs a plan through the C++ operators. sources: {node name: (uint8
    array of packed tuples, tuple count)}. Returns (uint8 rows [n, w])."""
    import numpy as np
    names = list(sources)
    arrs = [np.ascontiguousarray(sources[n][0]) for n in names]
    cn = (C.c_char_p * max(len(names), 1))(*[n.encode() for n in names])
    cd = (C.c_void_p * max(len(names), 1))(*[a.ctypes.data for a in arrs])
    ct = (C.c_uint64 * max(len(names), 1))(*[sources[n][1] for n in names])
    while True:
        out = np.zeros(max(out_cap_bytes, 1), np.uint8)
        nb = C.c_uint64()
        w = C.c_uint32()
        rc = lib().pj_query_run(plan.encode(), len(names), cn, cd, ct, out.ctypes.data_as(C.c_void_p),
                                out_cap_bytes, C.byref(nb), C.byref(w))
        if rc == PJ_ERR_CAPACITY:
            out_cap_bytes = nb.value
            continue
        check(rc)
        return out[: nb.value].reshape(-1, w.value) if w.value else out[:0]


def checksum_device(d_ptr, n, width, stream=None):
    s = C.c_uint64()
    x = C.c_uint32()
    check(lib().pj_checksum_device(C.c_void_p(d_ptr), n, width, C.byref(s), C.byref(x), C.c_void_p(stream)))
    return s.value, x.value


def zipf_cdf(n, z):
    import numpy as np
    cdf = np.zeros(n, np.float64)
    check(lib().pj_zipf_cdf(n, z, cdf.ctypes.data_as(C.c_void_p)))
    return cdf


def model_fit(d_keys, n, n_keys, iters, stream=None):
    """Fits the learned partitioner's model to n ordered keys in device memory."""
    m = Model()
    check(lib().pj_model_fit(C.c_void_p(d_keys), n, n_keys, iters, C.byref(m), C.c_void_p(stream)))
    return m


def model_grad(model, d_sorted, n, stream=None):
    """(loss, gradient) of the training loss at the model's parameters; the
    gradient is [dL/da[0..L), dL/db[0..L)]."""
    import numpy as np
    g = np.zeros(2 * model.nleaves, np.float32)
    loss = C.c_double()
    check(lib().pj_model_grad(C.byref(model), C.c_void_p(d_sorted), n, g.ctypes.data_as(C.c_void_p),
                              C.byref(loss), C.c_void_p(stream)))
    return loss.value, g


def model_bounds(model, max_coarse=2048):
    import numpy as np
    cap = 2 * max_coarse * 4 + PJ_MODEL_MAX_SEGS
    b = np.zeros(cap, np.uint64)
    nb = C.c_uint32()
    check(lib().pj_model_bounds(C.byref(model), max_coarse, b.ctypes.data_as(C.c_void_p), cap, C.byref(nb)))
    return b[: nb.value].copy()


def sort_u64(d_keys, n, stream=None):
    check(lib().pj_sort_u64(C.c_void_p(d_keys), n, C.c_void_p(stream)))


def cluster_table(seed, n_keys, nclusters=256, log2_wmin=10, log2_wmax=24, weight_z=1.2):
    """The clustered-key workload's cluster table (an opaque byte buffer)."""
    import numpy as np
    t = np.zeros(lib().pj_cluster_table_size(), np.uint8)
    check(lib().pj_clusters_make(seed, nclusters, n_keys, log2_wmin, log2_wmax, weight_z,
                                 t.ctypes.data_as(C.c_void_p)))
    return t


PJ_AGG_SUM, PJ_AGG_COUNT = 0, 1


class Aggregate:
    """Owns a pj_agg_t: aggregate_sum / aggregate_count over device tuples."""

    def __init__(self, cols, fields, op, sumfield=None, max_groups=1 << 20):
        d = AggDesc()
        d.schema = make_schema(cols)
        d.op = PJ_AGG_SUM if op == "sum" else PJ_AGG_COUNT
        d.nfields = len(fields)
        for i, f in enumerate(fields):
            d.fields[i] = f
        d.sumfield = sumfield or 0
        d.max_groups = max_groups
        self.h = C.c_void_p()
        check(lib().pj_agg_create(C.byref(d), C.byref(self.h)))
        w = C.c_uint32()
        check(lib().pj_agg_out_width(self.h, C.byref(w)))
        self.out_width = w.value

    def run_device(self, d_tuples, n, d_groups, groups_cap, stream=None):
        g = C.c_uint64()
        check(lib().pj_agg_run_device(self.h, C.c_void_p(d_tuples), n, C.c_void_p(d_groups), groups_cap,
                                      C.byref(g), C.c_void_p(stream)))
        return g.value

    def close(self):
        if self.h:
            lib().pj_agg_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def binary_tuples(pattern, width):
    """Tuples of `width` bytes in the raw binary files `pattern` globs."""
    n = C.c_uint64()
    check(lib().pj_binary_tuples(pattern.encode(), width, C.byref(n)))
    return n.value


def load_binary(pattern, width, d_dst, cap, stream=None):
    """Loads the raw binary files `pattern` globs into device memory at d_dst
    (room for cap tuples); returns the number of tuples."""
    n = C.c_uint64()
    rc = lib().pj_load_binary(pattern.encode(), width, C.c_void_p(d_dst), cap, C.byref(n), C.c_void_p(stream))
    if rc == PJ_ERR_CAPACITY:
        raise PJError(rc, "need room for %d tuples" % n.value)
    check(rc)
    return n.value


class SmjTimings(C.Structure):
    _fields_ = [("extract_ms", C.c_float), ("sort_ms", C.c_float), ("merge_ms", C.c_float),
                ("total_ms", C.c_float), ("build_key_bits", C.c_uint32), ("probe_key_bits", C.c_uint32)]


class SortMergeJoin:
    """Owns a pj_smj_t: the sort-merge join over the same descriptor and
    output rows as Join."""

    def __init__(self, desc):
        self.h = C.c_void_p()
        check(lib().pj_smj_create(C.byref(desc), C.byref(self.h)))
        w = C.c_uint32()
        check(lib().pj_smj_out_width(self.h, C.byref(w)))
        self.out_width = w.value

    def run_device(self, d_build, nbuild, d_probe, nprobe, d_out, out_cap, stream=None):
        n = C.c_uint64()
        rc = lib().pj_smj_run_device(self.h, C.c_void_p(d_build), nbuild, C.c_void_p(d_probe), nprobe,
                                     C.c_void_p(d_out), out_cap, C.byref(n), C.c_void_p(stream))
        if rc == PJ_ERR_CAPACITY:
            raise PJError(rc, "need %d output tuples" % n.value)
        check(rc)
        return n.value

    def timings(self):
        t = SmjTimings()
        check(lib().pj_smj_last_timings(self.h, C.byref(t)))
        return t

    def close(self):
        if self.h:
            lib().pj_smj_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


HASH_FNS = {"modulo": 0, "parammodulo": 1, "knuth": 2, "range": 3, "exactrange": 4, "bytes": 5,
            "tpchorderkey": 6, "willis": 7, "alwayszero": 8, "tpchq1magic": 9}


class HashSpec(C.Structure):
    _fields_ = [("fn", C.c_uint32), ("buckets", C.c_uint32), ("range_min", C.c_int64), ("range_max", C.c_int64),
                ("offset", C.c_uint32), ("skipbits", C.c_uint32)]


def hash_spec(fn, buckets, range_min=0, range_max=0, offset=0, skipbits=0):
    return HashSpec(HASH_FNS[fn], buckets, range_min, range_max, offset, skipbits)


def hash_buckets(spec):
    b = C.c_uint32()
    check(lib().pj_hash_buckets(C.byref(spec), C.byref(b)))
    return b.value


MEASURES = {"read": 0, "write": 1, "copy": 2, "gather": 3, "scatter": 4, "histogram": 5, "atomic": 6, "lookup": 7, "probe": 8}


class MeasureResult(C.Structure):
    _fields_ = [("ms", C.c_float), ("bytes", C.c_uint64), ("ops", C.c_uint64), ("gbps", C.c_double),
                ("gops", C.c_double)]


def measure(kind, d_a, d_b, d_idx, n, param=0, reps=5, stream=None):
    """One measurement kernel (pj_measure): the fastest of reps launches."""
    r = MeasureResult()
    check(lib().pj_measure(MEASURES[kind], C.c_void_p(d_a), C.c_void_p(d_b), C.c_void_p(d_idx), n, param, reps,
                           C.byref(r), C.c_void_p(stream)))
    return dict(ms=r.ms, bytes=r.bytes, ops=r.ops, gbps=r.gbps, gops=r.gops)


def set_device_limit(nbytes):
    """Caps the device memory the library's buffers may hold (0: no cap)."""
    check(lib().pj_set_device_limit(nbytes))


def device_bytes():
    v = C.c_uint64()
    check(lib().pj_device_bytes(C.byref(v)))
    return v.value


def hash_generate(spec, passes):
    """The passes functions of a multi-pass radix split of a parammodulo
    spec, high bits first (ParameterizedModuloValueHasher::generate,
    hash.cpp:251-281)."""
    out = (HashSpec * max(passes, 1))()
    check(lib().pj_hash_generate(C.byref(spec), passes, out))
    return list(out)[:passes]


def tuple_hash_device(spec, d_tuples, n, width, key_offset, key_size, d_buckets, stream=None):
    """Bucket ids of n device tuples under a reference hash function."""
    check(lib().pj_tuple_hash_device(C.byref(spec), C.c_void_p(d_tuples), n, width, key_offset, key_size,
                                     C.c_void_p(d_buckets), C.c_void_p(stream)))
