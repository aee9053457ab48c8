"""Pins the oracle against the reference itself: HashJoinOp and TupleHasher
compiled from the reference's sources into oracle/_ref (oracle/ref/Makefile).
Skipped where that build is absent."""
import numpy as np
import pytest

import oracle as O
import reference_run as RR
from _util import pack, unpack

pytestmark = pytest.mark.skipif(not RR.available(), reason="oracle/_ref not built (reference tree absent)")


def test_reference_queryhashjoin_each_key_once():
    # unit_tests/queryhashjoin.cpp run through the real operator: 4 threads,
    # 16-byte output pages
    T = 20
    build = pack(["long", "long"], [(k, k) for k in range(1, T + 1)])
    probe = pack(["long", "dec"], [(k, k + 0.1) for k in range(1, T + 1)])
    rows, _, _ = RR.run_join(["long", "long"], ["long", "dec"], [0], [0], ["B$1", "P$1"], build, probe,
                             threads=4, fn="modulo", buckets=16, tuplesperbucket=2, buffsize=16)
    got = sorted((int(v), float(d)) for v, d in unpack(["long", "dec"], rows))
    assert got == [(k, k + 0.1) for k in range(1, T + 1)]


@pytest.mark.parametrize("fn,threads", [("modulo", 1), ("modulo", 4), ("knuth", 3), ("bytes", 2), ("willis", 4)])
def test_oracle_hashjoin_equals_reference(fn, threads):
    rng = np.random.default_rng(41)
    cols = ["long", "long", "int"]
    nb, npr = 3000, 9000
    build = pack(cols, [(int(k), int(v), int(w)) for k, v, w in zip(
        rng.integers(-500, 500, nb), rng.integers(0, 1 << 62, nb), rng.integers(-(1 << 31), 1 << 31, nb))])
    probe = pack(cols, [(int(k), int(v), int(w)) for k, v, w in zip(
        rng.integers(-500, 500, npr), rng.integers(0, 1 << 62, npr), rng.integers(-(1 << 31), 1 << 31, npr))])
    proj = ["P$2", "B$1", "B$0", "P$1", "B$2"]
    ref, _, _ = RR.run_join(cols, cols, [0], [0], proj, build, probe, threads=threads, fn=fn, buckets=256,
                            tuplesperbucket=2, buffsize=4096)
    d = O.join_desc(cols, cols, [0], [0], proj, fn=fn, buckets=256, tuplesperbucket=2, buffsize=4096)
    ora, _ = O.hashjoin(d, build, probe)
    assert ref.shape[0] > 10000
    assert np.array_equal(O.sorted_rows(ref), O.sorted_rows(ora))


def test_oracle_composite_equals_reference():
    # queryhashjoincomposite.cpp shape with duplicates on both sides
    rng = np.random.default_rng(42)
    cols = ["int", "int", "long"]
    build = pack(cols, [(int(a), int(b), int(c)) for a, b, c in zip(
        rng.integers(0, 30, 2000), rng.integers(0, 30, 2000), rng.integers(0, 1 << 62, 2000))])
    probe = pack(cols, [(int(a), int(b), int(c)) for a, b, c in zip(
        rng.integers(0, 35, 2000), rng.integers(0, 35, 2000), rng.integers(0, 1 << 62, 2000))])
    for fn in ("modulo", "bytes"):
        ref, _, _ = RR.run_join(cols, cols, [0, 1], [0, 1], ["B$0", "B$1", "B$2", "P$2"], build, probe,
                                threads=3, fn=fn, buckets=64)
        d = O.join_desc(cols, cols, [0, 1], [0, 1], ["B$0", "B$1", "B$2", "P$2"], fn=fn, buckets=64)
        ora, _ = O.hashjoin(d, build, probe)
        assert ref.shape[0] > 1000
        assert np.array_equal(O.sorted_rows(ref), O.sorted_rows(ora))


@pytest.mark.parametrize("fn", ["modulo", "knuth", "parammodulo", "exactrange", "range", "tpchorderkey", "willis"])
def test_value_hashers_equal_reference(fn):
    rng = np.random.default_rng(43)
    kw = dict(range_min=-1000, range_max=(1 << 30)) if fn in ("exactrange", "range") else {}
    if fn in ("knuth", "parammodulo"):
        kw = dict(offset=77, skipbits=3)
    h = O.hasher(fn, 1000, **kw)
    for ct in ("long", "int"):
        if fn in ("exactrange", "range"):
            vals = rng.integers(-1000, 1 << 30, 500)
        elif ct == "int":
            vals = rng.integers(-(1 << 31), 1 << 31, 500)
        else:
            vals = rng.integers(-(1 << 62), 1 << 62, 500)
        for v in vals:
            a = np.array([v], dtype=np.int32 if ct == "int" else np.int64)
            mine = O.lib().or_hasher_hash(h, a.ctypes.data, a.itemsize)
            assert mine == RR.tuple_hash(fn, 1000, ct, int(v), **kw), (fn, ct, int(v))


def test_byte_hasher_equals_reference():
    # ByteHasher's sign-extended bytes and xor-fold: random keys, bytes >= 0x80
    # included, 8 and 16 byte keys, several bucket counts
    rng = np.random.default_rng(44)
    for buckets in (16, 1000, 1 << 20):
        h = O.hasher("bytes", buckets)
        for size in (8, 16):
            for _ in range(300):
                data = rng.integers(0, 256, size, dtype=np.uint8)
                assert O.lib().or_hasher_hash(h, data.ctypes.data, size) == RR.bytes_hash(buckets, data)


def test_reference_pkfk_matches_expected_checksum():
    b, p = O.gen_pkfk(77, 1 << 40, 50000, 200000, 20000)
    r = RR.run_pkfk(b, p, threads=4)
    s, c = O.expected_pkfk(77, 1 << 40, 50000, 200000, 20000)
    assert r["out_tuples"] == c
    assert O.checksum(r["rows"], 24)[0] == s


def test_reference_prefetch_steps_change_no_rows():
    # probeprefetchstep / buildprefetchstep (join.cpp:386-404) only prefetch:
    # the CPU baseline sweeps them, so each step must give the same rows
    b, p = O.gen_pkfk(11, 1 << 40, 1 << 12, 1 << 15, 1 << 12)
    base = RR.run_pkfk(b, p, 4)["rows"]
    for pf in (0, 8, 32):
        rows = RR.run_pkfk(b, p, 4, prefetch=pf)["rows"]
        assert np.array_equal(O.sorted_rows(rows), O.sorted_rows(base)), pf
    assert base.shape[0] == (1 << 15) - (1 << 12)


def test_oracle_row_id_shapes_equal_reference():
    # the shapes the device joins in its row-id form: (long, long) keys
    # (queryhashjoincomposite.cpp:156-200), three key columns across widths,
    # CHAR columns in the projection (writeData's strncpy) and odd widths
    rng = np.random.default_rng(45)
    cases = [
        (["long", "long", "long"], ["long", "long", "dec"], [0, 1], [0, 1], ["B$2", "P$2", "B$0", "P$1"],
         lambda r, n: [(int(a), int(b), int(c)) for a, b, c in zip(r.integers(0, 300, n), r.integers(-2, 2, n),
                                                                    r.integers(-(1 << 62), 1 << 62, n))],
         lambda r, n: [(int(a), int(b), float(c)) for a, b, c in zip(r.integers(0, 330, n), r.integers(-2, 2, n),
                                                                      r.standard_normal(n))]),
        (["int", "long", "int"], ["long", "int", "long"], [0, 1, 2], [0, 1, 2], ["B$0", "P$2", "B$2", "P$1"],
         lambda r, n: [(int(a), int(b), int(c)) for a, b, c in zip(r.integers(-9, 9, n), r.integers(-9, 9, n),
                                                                    r.integers(-3, 3, n))],
         lambda r, n: [(int(a), int(b), int(c)) for a, b, c in zip(
             np.where(r.random(n) < 0.1, r.integers(1 << 32, 1 << 40, n), r.integers(-9, 9, n)),
             r.integers(-9, 9, n), r.integers(-3, 3, n))]),
    ]
    for bc, pc, bj, pj, proj, gb, gp in cases:
        build = pack(bc, gb(rng, 2000))
        probe = pack(pc, gp(rng, 6000))
        for fn in ("modulo", "bytes") if bc == pc[:2] + [bc[2]] else ("modulo",):
            ref, _, _ = RR.run_join(bc, pc, bj, pj, proj, build, probe, threads=3, fn=fn, buckets=64)
            d = O.join_desc(bc, pc, bj, pj, proj, fn=fn, buckets=64)
            ora, _ = O.hashjoin(d, build, probe)
            assert ref.shape[0] > 1000
            assert np.array_equal(O.sorted_rows(ref), O.sorted_rows(ora))


def test_oracle_char_projection_equals_reference():
    from _util import dtype_of
    rng = np.random.default_rng(46)
    bc, pc = [("char", 5), "long", "int"], ["int", ("char", 3)]
    n = 3000
    b = np.zeros(n, dtype=dtype_of(bc))
    raw = rng.integers(0, 256, (n, 5), dtype=np.uint8)
    raw[rng.random((n, 5)) < 0.25] = 0
    b["f0"] = raw.view("S5").reshape(n)
    b["f1"] = rng.integers(-400, 400, n)
    b["f2"] = rng.integers(-(1 << 31), 1 << 31, n)
    p = np.zeros(3 * n, dtype=dtype_of(pc))
    p["f0"] = rng.integers(-450, 450, 3 * n)
    raw = rng.integers(0, 256, (3 * n, 3), dtype=np.uint8)
    raw[rng.random((3 * n, 3)) < 0.25] = 0
    p["f1"] = raw.view("S3").reshape(3 * n)
    build, probe = b.view(np.uint8).reshape(-1).copy(), p.view(np.uint8).reshape(-1).copy()
    proj = ["B$0", "P$1", "B$2", "B$1", "P$0"]
    ref, _, _ = RR.run_join(bc, pc, [1], [0], proj, build, probe, threads=2, fn="modulo", buckets=128)
    ora, _ = O.hashjoin(O.join_desc(bc, pc, [1], [0], proj, fn="modulo", buckets=128), build, probe)
    assert ref.shape[0] > 5000 and ref.shape[1] == 24
    assert np.array_equal(O.sorted_rows(ref), O.sorted_rows(ora))


def test_reference_consume_plan_xor_equals_output_xor():
    # MergeOp -> ConsumeOp -> HashJoinOp, the CPU baseline's plan: its XOR of
    # 32-bit words equals the XOR over the rows the reference writes
    b, p = O.gen_pkfk(78, 1 << 40, 30000, 120000, 9000)
    for threads in (1, 4):
        r = RR.consume_pkfk(b, p, threads)
        rows = RR.run_pkfk(b, p, threads)["rows"]
        assert r["xor32"] == O.checksum(rows, 24)[1]
        assert r["xor32"] != 0
