"""The C++ Operator layer (HashJoinOp, PartitionOp, MergeOp and the plan
builder) driven by plan text, as the reference's unit tests drive theirs
(unit_tests/queryhashjoin.cpp, querypartition.cpp). Plans that fail in init
are checked on the CPU; plans that run need the GPU."""
import os
import tempfile

import numpy as np
import pytest

import oracle as O
import pjoin as P
from _util import pack, unpack

QUERYHASHJOIN = """
buffsize = %(buffsize)d;
build: { type = "memscan"; schema = ( "long", "long" ); };
probe: { type = "memscan"; schema = ( "long", "dec" ); };
join:
{
  type = "hashjoin";
  hash: { fn = "%(fn)s"; range = [ 1, 20 ]; buckets = 16; };
  threadgroups = ( [ %(groups)s ] );
  tuplesperbucket = 2;
  allocpolicy = "local";
  buildjattr = 0;
  probejattr = 0;
  projection = ( "B$1", "P$1" );
};
merge: { type = "merge"; threads = %(threads)d; };
treeroot: { name: "merge"; input: { name: "join"; build: { name: "build"; }; probe: { name: "probe"; }; }; };
"""


def qhj_plan(fn="modulo", threads=4, buffsize=16):
    return QUERYHASHJOIN % dict(fn=fn, threads=threads, buffsize=buffsize,
                                groups=", ".join(str(i) for i in range(threads)))


def test_plan_syntax_error_is_invalid():
    with pytest.raises(P.PJError) as e:
        P.run_plan("buffsize = 16; join: { type = ", {})
    assert e.value.code == P.PJ_ERR_INVALID and "syntax" in str(e.value)


def test_unknown_hash_function_is_rejected_in_init():
    # UnknownHashException (hash.cpp:238-241)
    with pytest.raises(P.PJError) as e:
        P.run_plan(qhj_plan(fn="sha256"), {"build": (np.zeros(16, np.uint8), 1), "probe": (np.zeros(16, np.uint8), 1)})
    assert e.value.code == P.PJ_ERR_INVALID and "sha256" in str(e.value)


def test_missing_parameter_is_rejected_in_init():
    plan = qhj_plan().replace("tuplesperbucket = 2;", "")
    with pytest.raises(P.PJError) as e:
        P.run_plan(plan, {"build": (np.zeros(16, np.uint8), 1), "probe": (np.zeros(16, np.uint8), 1)})
    assert "tuplesperbucket" in str(e.value)


def test_unknown_operator_type():
    plan = qhj_plan().replace('type = "merge"', 'type = "shmwriter"')
    with pytest.raises(P.PJError) as e:
        P.run_plan(plan, {"build": (np.zeros(16, np.uint8), 1), "probe": (np.zeros(16, np.uint8), 1)})
    assert "shmwriter" in str(e.value)


@pytest.mark.gpu
@pytest.mark.parametrize("fn", ["modulo", "knuth", "bytes"])
def test_queryhashjoin_plan(fn):
    # unit_tests/queryhashjoin.cpp: 4 threads, 16-byte pages (one output
    # tuple a page), every key exactly once with d == v + 0.1
    T = 20
    build = pack(["long", "long"], [(k, k) for k in range(1, T + 1)])
    probe = pack(["long", "dec"], [(k, k + 0.1) for k in range(1, T + 1)])
    rows = P.run_plan(qhj_plan(fn=fn), {"build": (build, T), "probe": (probe, T)})
    got = sorted((int(v), float(d)) for v, d in unpack(["long", "dec"], rows))
    assert got == [(k, k + 0.1) for k in range(1, T + 1)]


@pytest.mark.gpu
@pytest.mark.parametrize("threads,buffsize", [(1, 24), (4, 24 * 7), (16, 1 << 16)])
def test_hashjoin_plan_matches_oracle(threads, buffsize):
    # resumable output pages under many threads against the restated
    # HashJoinOp on the same inputs
    b, p = O.gen_pkfk(31, 1 << 40, 50000, 400000, 40000)
    plan = """
buffsize = %d;
b: { type = "memscan"; schema = ( "long", "long" ); };
p: { type = "memscan"; schema = ( "long", "long" ); };
j: { type = "hashjoin"; hash: { fn = "modulo"; buckets = 65536; }; threadgroups = ( [ %s ] );
     tuplesperbucket = 2; allocpolicy = "local"; buildjattr = 0; probejattr = 0;
     projection = ( "B$0", "B$1", "P$1" ); probebatch = 100000; };
m: { type = "merge"; threads = %d; };
treeroot: { name = "m"; input: { name = "j"; build: { name = "b"; }; probe: { name = "p"; }; }; };
""" % (buffsize, ", ".join(str(i) for i in range(threads)), threads)
    rows = P.run_plan(plan, {"b": (b.view(np.uint8), 50000), "p": (p.view(np.uint8), 400000)}, 360000 * 24)
    d = O.join_desc(["long", "long"], ["long", "long"], [0], [0], ["B$0", "B$1", "P$1"], buckets=65536)
    exp, _ = O.hashjoin(d, b.view(np.uint8), p.view(np.uint8))
    assert rows.shape == exp.shape
    assert np.array_equal(O.sorted_rows(rows), O.sorted_rows(exp))


@pytest.mark.gpu
def test_composite_int_keys_plan():
    # queryhashjoincomposite.cpp's shape on INTEGER key columns: build
    # (i, 1..4, i), probe (i, {5, 2, 6, 7}, i + 0.1); only (i, 2) matches
    T = 50
    cols_b, cols_p = ["int", "int", "long"], ["int", "int", "dec"]
    build = pack(cols_b, [(i, j, i) for i in range(1, T + 1) for j in (1, 2, 3, 4)])
    probe = pack(cols_p, [(i, j, i + 0.1) for i in range(1, T + 1) for j in (5, 2, 6, 7)])
    for fn in ("modulo", "bytes"):
        plan = """
buffsize = 32;
b: { type = "memscan"; schema = ( "int", "int", "long" ); };
p: { type = "memscan"; schema = ( "int", "int", "dec" ); };
j: { type = "hashjoin"; hash: { fn = "%s"; range = [ 1, 50 ]; buckets = 16; }; threadgroups = ( [0, 1, 2, 3] );
     tuplesperbucket = 2; allocpolicy = "local"; buildjattr = [0, 1]; probejattr = [0, 1];
     projection = ( "B$2", "P$2" ); };
m: { type = "merge"; threads = 4; };
treeroot: { name = "m"; input: { name = "j"; build: { name = "b"; }; probe: { name = "p"; }; }; };
""" % fn
        rows = P.run_plan(plan, {"b": (build, 4 * T), "p": (probe, 4 * T)})
        got = sorted((int(v), float(d)) for v, d in unpack(["long", "dec"], rows))
        assert got == [(i, i + 0.1) for i in range(1, T + 1)]


@pytest.mark.gpu
def test_querypartition_plan():
    # unit_tests/querypartition.cpp: 2 x 128K random (v, v ^ 0xABCDEF), v in
    # 1..128K; exact per-value counts, one partition per value, partitions
    # monotone in key; thread count drawn from a fixed seed
    TUP = 128 * 1024
    rng = np.random.default_rng(7)
    for threads in (1, 5, 13, 15):
        vals = rng.integers(1, TUP + 1, size=2 * TUP)
        data = np.stack([vals, vals ^ 0xABCDEF], 1).astype(np.int64)
        plan = """
buffsize = 4096;
s: { type = "memscan"; schema = ( "long", "long" ); };
part: { type = "partition"; attr = 0; range = [ 1, %d ]; buckets = %d; maxtuples = %d; sort = "no"; };
tid: { type = "threadidprepend"; };
m: { type = "merge"; threads = %d; };
treeroot: { name = "m"; input: { name = "tid"; input: { name = "part"; input: { name = "s"; }; }; }; };
""" % (TUP, threads, 2 * TUP * threads, threads)
        rows = P.run_plan(plan, {"s": (data.view(np.uint8).reshape(-1), 2 * TUP)}, 2 * TUP * 20)
        out = unpack(["int", "long", "long"], rows)
        tid = out["f0"].astype(np.int64)
        v = out["f1"].astype(np.int64)
        assert np.all((v ^ 0xABCDEF) == out["f2"])
        assert np.array_equal(np.bincount(v, minlength=TUP + 1), np.bincount(vals, minlength=TUP + 1))
        src = np.full(TUP + 1, -1)
        for t in range(threads):
            vs = np.unique(v[tid == t])
            assert np.all(src[vs] == -1)
            src[vs] = t
        seen = src[1:][src[1:] >= 0]
        assert np.all(np.diff(seen) >= 0)


@pytest.mark.gpu
def test_binary_scan_plan(tmp_path):
    # raw binary tables, the reference's MemMappedTable layout: packed tuples,
    # no header
    b, p = O.gen_pkfk(32, 1 << 40, 20000, 100000, 0)
    (tmp_path / "r.bin").write_bytes(b.tobytes())
    (tmp_path / "s.bin").write_bytes(p.tobytes())
    plan = """
path = "%s/";
buffsize = 1048576;
b: { type = "scan"; filetype = "binary"; file = "r.bin"; schema = ( "long", "long" ); };
p: { type = "parallelscan"; filetype = "binary"; file = "s.bin"; schema = ( "long", "long" ); };
j: { type = "hashjoin"; hash: { fn = "modulo"; buckets = 32768; }; threadgroups = ( [0, 1, 2] );
     tuplesperbucket = 2; allocpolicy = "local"; buildjattr = 0; probejattr = 0;
     projection = ( "B$0", "B$1", "P$1" ); };
m: { type = "merge"; threads = 3; };
treeroot: { name = "m"; input: { name = "j"; build: { name = "b"; }; probe: { name = "p"; }; }; };
""" % tmp_path
    rows = P.run_plan(plan, {}, 100000 * 24)
    s, c = O.expected_pkfk(32, 1 << 40, 20000, 100000, 0)
    assert rows.shape[0] == c and O.checksum(rows, 24)[0] == s


SMJ_PLAN = """
buffsize = 65536;
build: { type = "memscan"; schema = ( %(bschema)s ); };
probe: { type = "memscan"; schema = ( %(pschema)s ); };
join: { type = "%(type)s"; threadgroups = ( [0, 1, 2, 3] ); maxbuildtuples = %(nb)d; maxprobetuples = %(np)d;
        buildjattr = 0; probejattr = 0; projection = ( %(proj)s ); };
merge: { type = "merge"; threads = 4; };
treeroot: { name = "merge"; input: { name = "join"; build: { name = "build"; }; probe: { name = "probe"; }; }; };
"""


def smj_plan(bcols, pcols, proj, nb, npr, kind="sortmergejoin"):
    q = lambda cols: ", ".join('"%s"' % c for c in cols)
    return SMJ_PLAN % dict(bschema=q(bcols), pschema=q(pcols), type=kind, nb=nb, np=npr,
                           proj=", ".join('"%s"' % p for p in proj))


def test_sortmerge_plan_needs_its_staging_sizes():
    # SortMergeJoinOp::init reads maxbuildtuples / maxprobetuples (join.cpp:934-956)
    p = smj_plan(["long", "long"], ["long", "long"], ["B$0", "P$1"], 10, 10).replace("maxbuildtuples = 10;", "")
    with pytest.raises(P.PJError) as e:
        P.run_plan(p, {"build": (np.zeros(16, np.uint8), 1), "probe": (np.zeros(16, np.uint8), 1)})
    assert "maxbuildtuples" in str(e.value)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["sortmergejoin", "mpsmjoin"])
def test_sortmerge_plan_equals_oracle(kind):
    b, p = O.gen_pkfk(34, 1 << 40, 30000, 120000, 12000)
    proj = ["B$0", "B$1", "P$1"]
    got = P.run_plan(smj_plan(["long", "long"], ["long", "long"], proj, 30000, 120000, kind),
                     {"build": (b.view(np.uint8), 30000), "probe": (p.view(np.uint8), 120000)})
    d = O.join_desc(["long", "long"], ["long", "long"], [0], [0], proj, buckets=1 << 15)
    exp, _ = O.hashjoin(d, b.view(np.uint8), p.view(np.uint8))
    assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))


@pytest.mark.gpu
def test_sortmerge_plan_many_to_many():
    rng = np.random.default_rng(35)
    nb, npr = 2000, 9000
    build = pack(["int", "long"], [(int(k), int(v)) for k, v in zip(rng.integers(-200, 200, nb),
                                                                    rng.integers(0, 1 << 62, nb))])
    probe = pack(["long", "int"], [(int(k), int(v)) for k, v in zip(rng.integers(-200, 200, npr),
                                                                    rng.integers(0, 1 << 31, npr))])
    proj = ["P$1", "B$1", "B$0"]
    got = P.run_plan(smj_plan(["int", "long"], ["long", "int"], proj, nb, npr),
                     {"build": (build, nb), "probe": (probe, npr)})
    d = O.join_desc(["int", "long"], ["long", "int"], [0], [0], proj, buckets=1 << 10)
    exp, _ = O.hashjoin(d, build, probe)
    assert got.shape[0] > npr
    assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))
