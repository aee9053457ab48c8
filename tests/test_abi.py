"""CPU-side checks of the C ABI: the library loads, exports every entry point
include/pjoin.h declares, and rejects bad descriptors before touching a GPU."""
import ctypes as C
import os
import re

import pytest

import pjoin as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            names |= set(re.findall(r"^\s*(?:int|const char\*|uint64_t|void)\s+(pj_\w+)\(", txt, re.M))
    return names


def test_library_exports_every_declared_symbol():
    L = P.lib()
    names = declared_entry_points()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_version_string():
    assert b"gfx950" in P.lib().pj_version()


def _desc(**kw):
    args = dict(build_cols=["long", "long"], probe_cols=["long", "long"], buildjattr=[0], probejattr=[0],
                projection=["B$0", "B$1", "P$1"])
    args.update(kw)
    return P.make_desc(**args)


def test_out_width_of_the_headline_projection():
    j = P.Join(_desc())
    assert j.out_width == 24
    j.close()
    j = P.Join(_desc(build_cols=["int", "int", "long"], probe_cols=["int", "int", "long"],
                     buildjattr=[0, 1], probejattr=[0, 1], projection=["B$0", "B$1", "B$2", "P$2"]))
    assert j.out_width == 24
    j.close()
    j = P.Join(_desc(build_cols=["long", "int"], probe_cols=["long", "int"], projection=["B$0", "B$1", "P$1"]))
    assert j.out_width == 16
    j.close()


@pytest.mark.parametrize("kw,code", [
    (dict(projection=["B$5"]), P.PJ_ERR_INVALID),                      # attribute out of range
    (dict(buildjattr=[3]), P.PJ_ERR_INVALID),
    (dict(build_cols=["dec", "long"]), P.PJ_ERR_UNSUPPORTED),          # DECIMAL key
    (dict(build_cols=[("char", 6), "long"]), P.PJ_ERR_UNSUPPORTED),    # CHAR key, odd width
    (dict(build_cols=["date", "long"]), P.PJ_ERR_UNSUPPORTED),         # DATE against LONG
    (dict(build_cols=["long", "dec"], probe_cols=["long", "dec"], buildjattr=[0, 1], probejattr=[0, 1]),
     P.PJ_ERR_UNSUPPORTED),                                           # DECIMAL in a composite key
])
def test_bad_descriptors_are_rejected(kw, code):
    h = C.c_void_p()
    d = _desc(**kw)
    rc = P.lib().pj_join_create(C.byref(d), C.byref(h))
    assert rc == code
    assert P.lib().pj_last_error()


@pytest.mark.parametrize("kw,width", [
    # the reference's own composite test: (long, long) keys (queryhashjoincomposite.cpp:156-200)
    (dict(build_cols=["long", "long", "long"], probe_cols=["long", "long", "dec"], buildjattr=[0, 1],
          probejattr=[0, 1], projection=["B$2", "P$2"]), 16),
    # three key columns, INTEGER against LONG
    (dict(build_cols=["int", "long", "int", "long"], probe_cols=["long", "int", "long", "long"],
          buildjattr=[0, 1, 2], probejattr=[0, 1, 2], projection=["B$0", "B$3", "P$3"]), 20),
    # more than 8 payload bytes a side, a 40-byte row
    (dict(build_cols=["long", "long", "long"], projection=["B$1", "B$2", "P$1", "B$0", "P$0"]), 40),
    # CHAR columns and a tuple width that is not a multiple of 4
    (dict(build_cols=["long", ("char", 7), "int"], probe_cols=["long", ("char", 3)],
          projection=["B$1", "P$1", "B$2", "B$0"]), 22),
])
def test_row_id_descriptors_are_accepted(kw, width):
    j = P.Join(_desc(**kw))
    assert j.out_width == width
    j.close()


def test_probe_before_build_finish_is_a_state_error():
    j = P.Join(_desc())
    cur = P.Cursor()
    out = (C.c_uint8 * 24)()
    n = C.c_uint64()
    tup = (C.c_uint64 * 2)(1, 2)
    rc = P.lib().pj_join_probe(j.h, tup, 1, C.byref(cur), out, 1, C.byref(n))
    assert rc == P.PJ_ERR_STATE
    j.close()


def test_batch_modes_and_configs_before_prepare():
    """The batch handle checks its mode, state and configurations on the host
    before any device work: running or reading results before prepare is a
    state error, an unknown mode is invalid, two partition passes are
    refused, and so are radix bits above 12 and load factors outside
    0.05..1."""
    bt = P.Batch(_desc())
    assert P.lib().pj_batch_run(bt.h, P.Batch.SEQUENTIAL, None) == P.PJ_ERR_STATE
    assert P.lib().pj_batch_run(bt.h, P.Batch.GRAPH, None) == P.PJ_ERR_STATE
    assert P.lib().pj_batch_run(bt.h, 7, None) == P.PJ_ERR_INVALID
    cnt = (C.c_uint64 * 1)()
    assert P.lib().pj_batch_results(bt.h, cnt, None, None) == P.PJ_ERR_STATE
    for bits, passes, lf, code in [(8, 2, 0.0, P.PJ_ERR_UNSUPPORTED), (13, 1, 0.0, P.PJ_ERR_INVALID),
                                   (0, 1, 1.5, P.PJ_ERR_INVALID), (0, 1, 0.01, P.PJ_ERR_INVALID)]:
        with pytest.raises(P.PJError) as e:
            bt.set_config(10, bits, passes, lf)
        assert e.value.code == code
    bt.set_config(10, 0, 1, 0.7, piece=65536)
    assert bt.config(10) == (0, 1, pytest.approx(0.7), 65536)
    assert bt.config(11) == (0, 1, 0.0, 0)   # untuned: the single join's choice
    bt.close()


@pytest.mark.parametrize("kw,code", [
    (dict(op="sum", fields=[0], sumfield=5), P.PJ_ERR_INVALID),               # sumfield out of range
    (dict(op="sum", fields=[0], sumfield=2), P.PJ_ERR_UNSUPPORTED),           # DECIMAL sum
    (dict(op="count", fields=[0, 1, 2]), P.PJ_ERR_UNSUPPORTED),               # three grouping columns
    (dict(op="count", fields=[0, 1]), P.PJ_ERR_UNSUPPORTED),                  # composite of LONG columns
])
def test_bad_aggregates_are_rejected(kw, code):
    d = P.AggDesc()
    d.schema = P.make_schema(["long", "long", "dec"])
    d.op = P.PJ_AGG_SUM if kw["op"] == "sum" else P.PJ_AGG_COUNT
    d.nfields = len(kw["fields"])
    for i, f in enumerate(kw["fields"][:2]):
        d.fields[i] = f
    d.sumfield = kw.get("sumfield", 0)
    d.max_groups = 100
    h = C.c_void_p()
    assert P.lib().pj_agg_create(C.byref(d), C.byref(h)) == code
    assert P.lib().pj_last_error()


def test_aggregate_output_width():
    a = P.Aggregate(["int", "long", "int"], [0], "sum", 1)
    assert a.out_width == 12          # INTEGER group, LONG sum
    a.close()
    a = P.Aggregate(["int", "int", "long"], [0, 1], "count", None)
    assert a.out_width == 16          # two INTEGER groups, LONG count
    a.close()


def test_sortmerge_join_validates_like_the_hash_join():
    d = _desc(projection=["B$0", "B$9"])
    h = C.c_void_p()
    assert P.lib().pj_smj_create(C.byref(d), C.byref(h)) == P.PJ_ERR_INVALID
    j = P.SortMergeJoin(_desc())
    assert j.out_width == P.Join(_desc()).out_width == 24
    j.close()


def test_exactrange_partition_arguments_are_checked_before_the_device():
    """PartitionOp's exactrange bucket width (max - min + B) / B (hash.h:455-460)
    is LONG arithmetic: spans that would overflow it are refused, as are
    empty ranges and bucket counts PartitionOp::init refuses (partition.cpp:44-105)."""
    sch = P.make_schema(["long", "long"])
    sizes = (C.c_uint64 * 8)()
    lo, hi = -(1 << 63), (1 << 63) - 1
    for rmin, rmax, b in ((lo, hi, 8), (0, hi - 7, 8), (5, 4, 8), (0, 10, 0), (0, 10, 128)):
        rc = P.lib().pj_partition_exactrange_device(C.byref(sch), 0, rmin, rmax, b, None, 0, None, sizes, None)
        assert rc == P.PJ_ERR_INVALID, (rmin, rmax, b)
    # the widest span that fits, on an empty relation: no device work, all sizes zero
    sizes[0] = 99
    assert P.lib().pj_partition_exactrange_device(C.byref(sch), 0, 0, hi - 8, 8, None, 0, None, sizes, None) == P.PJ_OK
    assert list(sizes) == [0] * 8


def test_measure_refuses_bad_arguments_before_any_launch():
    # pj_measure checks its buffers and parameters on the host first
    for kind, a, b, idx, n, param in (("read", 0, 0, 0, 1 << 20, 0), ("copy", 1 << 20, 0, 0, 1 << 20, 0),
                                      ("gather", 1 << 20, 1 << 20, 0, 1 << 20, 0),
                                      ("histogram", 1 << 20, 1 << 20, 0, 1 << 20, 16),
                                      ("histogram", 1 << 20, 1 << 20, 0, 16, 11),
                                      ("atomic", 1 << 20, 1 << 20, 0, 1 << 20, 3)):
        with pytest.raises(P.PJError) as e:
            P.measure(kind, a, b, idx, n, param)
        assert e.value.code == P.PJ_ERR_INVALID
    r = P.MeasureResult()
    assert P.lib().pj_measure(99, None, None, None, 1, 0, 1, C.byref(r), None) == P.PJ_ERR_INVALID
