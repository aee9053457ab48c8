"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

Each check names what fixes the expected value: a known-answer test of the
reference's unit_tests, a closed form, a brute-force nested-loop join, or the
reference itself built into oracle/_ref (tests/test_reference.py).
"""
import math
import random

import numpy as np
import pytest

import oracle as O
from _util import pack, unpack


# ---------------------------------------------------------------- hashes

def test_getlogarithm_is_ceil_log2():
    # unit_tests/testhash.cpp:71-94 (getlogarithm == ceil(log2 k)); exhaustive
    # over the range the reference draws from.
    L = O.lib()
    for k in range(1, 1 << 12):
        assert L.or_getlogarithm(k) == math.ceil(math.log2(k)), k
    assert L.or_getlogarithm(1 << 20) == 20
    assert L.or_getlogarithm((1 << 20) + 1) == 21


def test_generate_three_passes():
    # unit_tests/testhash.cpp:53-69: parammodulo(0, 1024, skipbits 2).generate(3)
    # gives 8, 8 and 16 buckets.
    import ctypes as C
    b = (C.c_uint * 3)()
    s = (C.c_uint * 3)()
    assert O.lib().or_parammodulo_generate(0, 1024, 2, 3, b, s) == 3
    assert list(b) == [8, 8, 16]
    # the passes take disjoint bit ranges, high bits first (hash.cpp:256-268)
    assert list(s) == [2 + 10 - 3, 2 + 10 - 6, 2]
    assert O.lib().or_hasher_buckets(O.hasher("parammodulo", 1024, skipbits=2)) == 1024


def test_modulo_bounds_and_closed_form():
    # unit_tests/testhash.cpp:96-106 plus the closed form v & (2^k - 1)
    h = O.hasher("modulo", 1000)     # rounds up to 1024 (hash.cpp:61-69)
    assert O.lib().or_hasher_buckets(h) == 1024
    for v in list(range(0, 10001)) + [-1, -(1 << 40), (1 << 62) + 5]:
        assert O.lib().or_hasher_hash_value(h, v) == (v & 1023)


def test_knuth_closed_form():
    h = O.hasher("knuth", 256, offset=12345, skipbits=3)
    rng = random.Random(5)
    for _ in range(2000):
        v = rng.randrange(-(1 << 63), 1 << 63)
        exp = (((v * 2654435761) % (1 << 64)) & (255 << 3)) >> 3
        assert O.lib().or_hasher_hash_value(h, v) == exp


def test_alwayszero():
    # unit_tests/testhash.cpp:108-118
    h = O.hasher("alwayszero", 0)
    assert O.lib().or_hasher_buckets(h) == 1
    for v in range(100):
        assert O.lib().or_hasher_hash_value(h, v) == 0


def test_exactrange_boundaries():
    # unit_tests/testhash.cpp:120-144, with a fixed seed instead of time()
    rng = random.Random(1234)
    L = O.lib()
    for _ in range(2000):
        mx = rng.randrange(0, 1 << 31) & ~1
        mn = rng.randrange(0, 1 << 31)
        while mn >= mx:
            mn = rng.randrange(0, 1 << 31)
            mx = max(mx, 2)
        h = O.hasher("exactrange", 80, range_min=mn, range_max=mx)
        assert L.or_hasher_buckets(h) == 80
        assert L.or_hasher_hash_value(h, mx) <= 79
        for j in range(1, 80):
            m = L.or_exactrange_minimumforbucket(h, j)
            assert L.or_hasher_hash_value(h, m - 1) == j - 1
            assert L.or_hasher_hash_value(h, m) == j


def test_find_max_contiguous():
    # unit_tests/testmaxcontiguousregion.cpp:43-103
    import ctypes as C

    def f(a, b):
        A = (C.c_uint16 * len(a))(*a)
        B = (C.c_uint16 * len(b))(*b)
        return O.lib().or_find_max_contiguous(A, B, len(a))

    assert f([3, 4, 5], [0, 1, 5]) == 1
    assert f([3, 4, 5], [0, 1, 2]) == 2
    assert f([0, 3], [0, 1]) == 0
    assert f([0], [0]) == 0


def test_bytes_hash_known_values():
    # Over bytes < 0x80, where sign extension changes nothing, ByteHasher is
    # FNV-1a-64 (hash.h:165-191) followed by the xor-fold to k bits
    # (hash.h:195). Published FNV-1a-64 vectors: "a" -> 0xaf63dc4c8601ec8c,
    # "foobar" -> 0x85944171f73967e8.
    h = O.hasher("bytes", 1 << 16)
    data = np.frombuffer(b"a", dtype=np.uint8).copy()
    full = 0xAF63DC4C8601EC8C
    exp = ((full >> 16) ^ full) & 0xFFFF
    assert O.lib().or_hasher_hash(h, data.ctypes.data, 1) == exp
    data = np.frombuffer(b"foobar", dtype=np.uint8).copy()
    full = 0x85944171F73967E8
    assert O.lib().or_hasher_hash(h, data.ctypes.data, 6) == (((full >> 16) ^ full) & 0xFFFF)
    # a byte >= 0x80 is XORed sign-extended (hash.h:183): differs from
    # textbook FNV-1a, which XORs 0x80 into the low byte only.
    data = np.array([0x80], dtype=np.uint8)
    off, prime = 0xCBF29CE484222325, 0x100000001B3
    signext = ((off ^ 0xFFFFFFFFFFFFFF80) * prime) % (1 << 64)
    assert O.lib().or_hasher_hash(h, data.ctypes.data, 1) == (((signext >> 16) ^ signext) & 0xFFFF)


# ----------------------------------------------------------- hash join

def _queryhashjoin_desc(fn="modulo", buffsize=16):
    # unit_tests/queryhashjoin.cpp:116-182: modulo, 16 buckets, 2 tuples a
    # bucket, 16-byte output pages, projection B$1, P$1
    return O.join_desc(["long", "long"], ["long", "dec"], [0], [0], ["B$1", "P$1"], fn=fn,
                       buckets=16, tuplesperbucket=2, buffsize=buffsize, range_min=1, range_max=20)


def test_queryhashjoin_each_key_once():
    # unit_tests/queryhashjoin.cpp:85-92, 213-218
    T = 20
    build = pack(["long", "long"], [(k, k) for k in range(1, T + 1)])
    probe = pack(["long", "dec"], [(k, k + 0.1) for k in range(1, T + 1)])
    out, st = O.hashjoin(_queryhashjoin_desc(), build, probe)
    rows = unpack(["long", "dec"], out)
    seen = {}
    for v, d in rows:
        assert 1 <= v <= T
        assert d == v + 0.1
        seen[v] = seen.get(v, 0) + 1
    assert seen == {k: 1 for k in range(1, T + 1)}
    # a 16-byte page holds one tuple: every output tuple ends a page
    assert st.pages == T + 1


def test_queryhashjoincomposite_only_pair_2():
    # unit_tests/queryhashjoincomposite.cpp:98-122, 248-252: build (i, 1..4, i),
    # probe (i, {5, 2, 6, 7}, i + 0.1); only (i, 2) matches, under both
    # "modulo" (hash on a, filter on (a, b)) and "bytes" (hash on both keys).
    T = 20
    build = pack(["long"] * 3, [(i, j, i) for i in range(1, T + 1) for j in (1, 2, 3, 4)])
    probe = pack(["long", "long", "dec"], [(i, j, i + 0.1) for i in range(1, T + 1) for j in (5, 2, 6, 7)])
    for fn in ("modulo", "bytes"):
        d = O.join_desc(["long"] * 3, ["long", "long", "dec"], [0, 1], [0, 1], ["B$2", "P$2"], fn=fn,
                        buckets=16, tuplesperbucket=2, buffsize=32)
        out, _ = O.hashjoin(d, build, probe)
        rows = unpack(["long", "dec"], out)
        assert sorted(int(v) for v, _ in rows) == list(range(1, T + 1))
        assert all(d == v + 0.1 for v, d in rows)


def test_chain_overflow_each_value_once():
    # unit_tests/testhashtable.cpp:44-85: one bucket of two slots, N inserts;
    # every inserted value is found exactly once through the overflow chain.
    rng = np.random.default_rng(3)
    vals = rng.choice(1 << 40, size=500, replace=False).astype(np.int64)
    build = pack(["long", "long"], [(int(v), int(v) ^ 7) for v in vals])
    probe = pack(["long", "long"], [(int(v), 1) for v in vals])
    d = O.join_desc(["long", "long"], ["long", "long"], [0], [0], ["B$0", "B$1"], buckets=1,
                    tuplesperbucket=2)
    out, st = O.hashjoin(d, build, probe)
    assert st.spills == len(vals) // 2 - 1
    rows = unpack(["long", "long"], out)
    assert sorted(int(a) for a, _ in rows) == sorted(int(v) for v in vals)
    assert all(int(b) == int(a) ^ 7 for a, b in rows)


def _random_relation(rng, cols, n, keyrange, key_cols):
    rows = []
    for _ in range(n):
        r = []
        for i, c in enumerate(cols):
            if i in key_cols:
                r.append(int(rng.integers(-keyrange, keyrange)))
            elif c == "dec":
                r.append(float(rng.normal()))
            elif c == "int":
                r.append(int(rng.integers(-(1 << 31), 1 << 31)))
            else:
                r.append(int(rng.integers(-(1 << 62), 1 << 62)))
        rows.append(tuple(r))
    return pack(cols, rows)


@pytest.mark.parametrize("fn", ["modulo", "knuth", "bytes", "willis", "tpchorderkey", "parammodulo"])
def test_hashjoin_matches_nested_loop(fn):
    # brute force: a many-to-many join (duplicate keys both sides) gives the
    # same multiset through every hash function and bucket count
    rng = np.random.default_rng(11)
    bcols, pcols = ["int", "long", "dec"], ["long", "int", "long"]
    if fn == "bytes":
        # ByteHasher hashes raw key bytes, so an INTEGER and a LONG key of
        # the same value land in different buckets (hash.h:150-157); the
        # hash-independence below holds only for equal key widths.
        bcols = ["long", "long", "dec"]
    build = _random_relation(rng, bcols, 300, 40, {0})
    probe = _random_relation(rng, pcols, 700, 40, {0})
    for buckets, tpb in ((1, 1), (16, 2), (1024, 4)):
        d = O.join_desc(bcols, pcols, [0], [0], ["B$0", "B$2", "P$2", "P$1", "B$1"], fn=fn,
                        buckets=buckets, tuplesperbucket=tpb, buffsize=7 * 40)
        got, _ = O.hashjoin(d, build, probe)
        exp = O.nested_loop_join(d, build, probe)
        assert got.shape == exp.shape and got.shape[0] > 100
        assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))


def test_composite_int_keys_match_nested_loop():
    rng = np.random.default_rng(12)
    cols = ["int", "int", "long"]
    build = _random_relation(rng, cols, 400, 6, {0, 1})
    probe = _random_relation(rng, cols, 400, 6, {0, 1})
    for fn in ("modulo", "bytes"):
        d = O.join_desc(cols, cols, [0, 1], [0, 1], ["B$0", "B$1", "B$2", "P$2"], fn=fn, buckets=64)
        got, _ = O.hashjoin(d, build, probe)
        exp = O.nested_loop_join(d, build, probe)
        assert got.shape[0] > 100
        assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))


def test_int_long_cross_width_keys():
    # rawcompfns.cpp:78, 141: an INTEGER key equals the LONG key of the same
    # value, negative values included (sign extension)
    build = pack(["int", "long"], [(-5, 1), (7, 2), (-(1 << 31), 3)])
    probe = pack(["long", "long"], [(-5, 10), (7, 20), (-(1 << 31), 30), ((1 << 32) - 5, 40)])
    d = O.join_desc(["int", "long"], ["long", "long"], [0], [0], ["B$1", "P$1"], buckets=8)
    out, _ = O.hashjoin(d, build, probe)
    rows = sorted(tuple(int(x) for x in r) for r in unpack(["long", "long"], out))
    assert rows == [(1, 10), (2, 20), (3, 30)]


def test_bytes_hash_drops_cross_width_matches():
    # the reference's own behaviour: "bytes" over an INTEGER build key and a
    # LONG probe key hashes 4 vs 8 bytes, so equal values rarely meet
    build = pack(["int", "long"], [(k, k) for k in range(200)])
    probe = pack(["long", "long"], [(k, k) for k in range(200)])
    d = O.join_desc(["int", "long"], ["long", "long"], [0], [0], ["B$1", "P$1"], fn="bytes", buckets=256)
    out, _ = O.hashjoin(d, build, probe)
    assert out.shape[0] < 200
    d = O.join_desc(["int", "long"], ["long", "long"], [0], [0], ["B$1", "P$1"], fn="modulo", buckets=256)
    assert O.hashjoin(d, build, probe)[0].shape[0] == 200


def test_empty_inputs():
    d = _queryhashjoin_desc(buffsize=1 << 10)
    e = np.zeros(0, np.uint8)
    b = pack(["long", "long"], [(1, 1)])
    p = pack(["long", "dec"], [(1, 1.1)])
    assert O.hashjoin(d, e, p)[0].shape[0] == 0
    assert O.hashjoin(d, b, e)[0].shape[0] == 0


# ------------------------------------------------------------ partition

def test_querypartition_invariants():
    # unit_tests/querypartition.cpp:84-98, 204-231: per-value counts exact,
    # each value in one partition, partitions monotone in key.
    T = 128 * 1024
    rng = np.random.default_rng(21)
    sch = O.schema(["long", "long"])
    for threads in (1, 3, 7, 15):
        vals = rng.integers(1, T + 1, size=2 * T)
        rows = np.stack([vals, vals ^ 0xABCDEF], axis=1).astype(np.int64)
        split = np.array_split(rows, threads)
        parts = O.partition(sch, 0, 1, T, [s.view(np.uint8).reshape(-1) for s in split])
        seen = np.full(T + 1, -1)
        counts = np.zeros(T + 1, np.int64)
        prev_max = 0
        for j, p in enumerate(parts):
            r = p.view(np.int64).reshape(-1, 2)
            assert np.all((r[:, 0] ^ 0xABCDEF) == r[:, 1])
            if r.shape[0]:
                assert r[:, 0].min() >= prev_max
                prev_max = r[:, 0].max()
                assert np.all((seen[r[:, 0]] == -1) | (seen[r[:, 0]] == j))
                seen[r[:, 0]] = j
                np.add.at(counts, r[:, 0], 1)
        assert np.array_equal(counts[1:], np.bincount(vals, minlength=T + 1)[1:])


def test_partition_scatter_order_is_thread_major():
    # partition.cpp:293-306: inside a partition, thread 0's tuples come first,
    # each thread's in input order
    sch = O.schema(["long", "long"])
    t0 = np.array([[5, 0], [1, 1], [6, 2]], np.int64)
    t1 = np.array([[2, 3], [7, 4]], np.int64)
    parts = O.partition(sch, 0, 0, 9, [t0.view(np.uint8).reshape(-1), t1.view(np.uint8).reshape(-1)])
    p0 = parts[0].view(np.int64).reshape(-1, 2)
    p1 = parts[1].view(np.int64).reshape(-1, 2)
    assert p0[:, 1].tolist() == [1, 3]
    assert p1[:, 1].tolist() == [0, 2, 4]


# ----------------------------------------------- checksums and workloads

def test_checksum_is_order_independent_and_sensitive():
    rng = np.random.default_rng(4)
    rows = rng.integers(0, 1 << 63, size=(1000, 3), dtype=np.int64).view(np.uint8).reshape(1000, 24)
    s1, x1 = O.checksum(rows, 24)
    s2, x2 = O.checksum(rows[rng.permutation(1000)], 24)
    assert (s1, x1) == (s2, x2)
    flipped = rows.copy()
    flipped[17, 9] ^= 1
    assert O.checksum(flipped, 24)[0] != s1
    # a transposed column pair changes it, the XOR of words does not
    swapped = rows.copy()
    swapped[3, 0:8], swapped[3, 8:16] = rows[3, 8:16], rows[3, 0:8]
    assert O.checksum(swapped, 24)[0] != s1
    assert O.checksum(swapped, 24)[1] == x1


def test_expected_pkfk_matches_restated_hashjoin():
    # the generator-derived closed form agrees with the restated HashJoinOp
    for seed, dom, nb, npr, nmiss in ((1, 1 << 14, 1 << 14, 1 << 16, 0),
                                     (2, 1 << 40, 5000, 30000, 3000)):
        b, p = O.gen_pkfk(seed, dom, nb, npr, nmiss)
        keys = b[0::2]
        assert len(np.unique(keys)) == nb and keys.max() < dom
        if dom == nb:
            assert np.array_equal(np.sort(keys), np.arange(nb, dtype=np.uint64))
        d = O.join_desc(["long", "long"], ["long", "long"], [0], [0], ["B$0", "B$1", "P$1"],
                        buckets=1 << 14)
        out, _ = O.hashjoin(d, b.view(np.uint8), p.view(np.uint8))
        exp_sum, exp_cnt = O.expected_pkfk(seed, dom, nb, npr, nmiss)
        assert out.shape[0] == exp_cnt == npr - nmiss
        assert O.checksum(out, 24)[0] == exp_sum


# ------------------------------------------------ Zipf composite workload

def test_zipf_cdf_product_equals_reference_formula():
    # the product's CDF (pj_zipf_cdf) equals generator_zipf.cpp:91-113's
    # running sum in double, computed independently here
    import pjoin as P
    for n, z in ((1000, 0.5), (100000, 1.0), (1 << 20, 1.5)):
        assert np.array_equal(P.zipf_cdf(n, z), O.zipf_cdf(n, z))


def test_zipf_sampling_follows_the_cdf():
    # P(rank 0) = 1 / H_n(z): about 38% of probes hit the top key at z = 1.5
    n = 1 << 16
    cdf = O.zipf_cdf(n, 1.5)
    b, p = O.gen_pairs_zipf(9, n, 200000, cdf)
    bk = b.reshape(-1, 16)[:, :8].copy().view(np.uint64).ravel()
    pk = p.reshape(-1, 16)[:, :8].copy().view(np.uint64).ravel()
    assert len(np.unique(bk)) == n
    vals, counts = np.unique(pk, return_counts=True)
    top = counts.max() / pk.size
    assert abs(top - cdf[0]) < 0.005, (top, cdf[0])
    assert np.isin(pk, bk).all()


def test_expected_pairs_zipf_matches_restated_hashjoin():
    n, m = 20000, 100000
    cdf = O.zipf_cdf(n, 1.0)
    b, p = O.gen_pairs_zipf(5, n, m, cdf)
    cols = ["int", "int", "long"]
    d = O.join_desc(cols, cols, [0, 1], [0, 1], ["B$0", "B$1", "B$2", "P$2"], fn="bytes", buckets=1 << 14)
    out, _ = O.hashjoin(d, b, p)
    s, c = O.expected_pairs_zipf(5, n, m, cdf)
    assert out.shape[0] == c == m
    assert O.checksum(out, 24)[0] == s


# ------------------------------------------------- clustered-key workload

def _ranks_binary_search(cdf, seed, m):
    # pj_zipf_rank restated with numpy: same uniforms (splitmix64 streams),
    # first CDF entry > u; u >= the last entry never occurs here
    import pjoin as P  # noqa: F401
    out = []
    M = (1 << 64) - 1

    def fin(x):
        x ^= x >> 30
        x = (x * 0xBF58476D1CE4E5B9) & M
        x ^= x >> 27
        x = (x * 0x94D049BB133111EB) & M
        return x ^ (x >> 31)
    for j in range(m):
        a = 0
        while True:
            r = fin((fin(seed ^ (((5 + a) * 0xD1B54A32D192ED03) & M)) + j * 0x9E3779B97F4A7C15) & M) >> 2
            u = r / float(1 << 62)
            if u < cdf[-1]:
                break
            a += 1
        out.append(int(np.searchsorted(cdf, u, side="right")))
    return out


def test_guided_zipf_rank_equals_binary_search():
    import ctypes as C
    cdf = O.zipf_cdf(50000, 1.0)
    ranks = np.zeros(3000, np.uint64)
    O.lib().or_zipf_ranks_guided(cdf.ctypes.data_as(C.c_void_p), cdf.size, 17, 3000, ranks.ctypes.data)
    assert ranks.tolist() == _ranks_binary_search(cdf, 17, 3000)


def test_cluster_table_and_expected_clustered():
    import ctypes as C
    import pjoin as P
    n, m = 200000, 300000
    t = P.cluster_table(41, n, nclusters=64, log2_wmin=10, log2_wmax=16)
    cdf = O.zipf_cdf(n, 1.0)
    b = np.zeros(2 * n, np.uint64)
    p = np.zeros(2 * m, np.uint64)
    P.check(P.lib().pj_gen_clustered(t.ctypes.data_as(C.c_void_p), m, cdf.ctypes.data_as(C.c_void_p),
                                     b.ctypes.data_as(C.c_void_p), p.ctypes.data_as(C.c_void_p), 0, None))
    keys = b[0::2]
    assert len(np.unique(keys)) == n and keys.max() < (1 << 63)
    d = O.join_desc(["long", "long"], ["long", "long"], [0], [0], ["B$0", "B$1", "P$1"], buckets=1 << 18)
    out, _ = O.hashjoin(d, b.view(np.uint8), p.view(np.uint8))
    s, c = O.expected_clustered(t, m, cdf)
    assert out.shape[0] == c == m
    assert O.checksum(out, 24)[0] == s
