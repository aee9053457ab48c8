"""GPU parity of the device hash join against the CPU oracle.

Every case compares the HIP path, called through the C ABI, with the restated
HashJoinOp (oracle/oracle.c) or the brute-force nested-loop join on the same
seeded inputs: bit-exact output multisets (sorted row compare) at small
sizes, count + 64-bit checksum against the generator's closed form at the
full 1M x 16M and 16M x 256M shapes.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle as O
import pjoin as P
from _util import pack, to_gpu, unpack

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    P.lib()
    yield


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_join(desc, build, probe, nb, npr, cap=None):
    """Host uint8 arrays in, output rows (n x width uint8) out."""
    j = P.Join(desc)
    db = to_gpu(build) if build.size else torch.zeros(16, dtype=torch.uint8, device="cuda")
    dp = to_gpu(probe) if probe.size else torch.zeros(16, dtype=torch.uint8, device="cuda")
    cap = cap if cap is not None else max(npr, 1)
    out = torch.empty(max(cap, 1) * j.out_width, dtype=torch.uint8, device="cuda")
    n = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, out.data_ptr(), cap,
                     torch.cuda.current_stream().cuda_stream)
    rows = out[: n * j.out_width].cpu().numpy().reshape(-1, j.out_width)
    t = j.timings()
    j.close()
    return rows, t


def oracle_desc(pdesc_args):
    a = dict(pdesc_args)
    return O.join_desc(a["build_cols"], a["probe_cols"], a["buildjattr"], a["probejattr"], a["projection"],
                       buckets=1 << 12)


def compare(args, build, probe, nb, npr, **kw):
    exp, _ = O.hashjoin(oracle_desc(args), build, probe)
    got, t = run_join(P.make_desc(**args, **kw), build, probe, nb, npr, cap=max(exp.shape[0], 1))
    assert got.shape == exp.shape, (got.shape, exp.shape)
    assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))
    return got, t


HEADLINE = dict(build_cols=["long", "long"], probe_cols=["long", "long"], buildjattr=[0], probejattr=[0],
                projection=["B$0", "B$1", "P$1"])


def test_device_generator_equals_host_generator():
    _say("case: generator device == host")
    nb, npr, miss = 5000, 40000, 4000
    db = torch.empty(nb * 16, dtype=torch.uint8, device="cuda")
    dp = torch.empty(npr * 16, dtype=torch.uint8, device="cuda")
    P.check(P.lib().pj_gen_pkfk(9, 1 << 40, nb, npr, miss, C.c_void_p(db.data_ptr()), C.c_void_p(dp.data_ptr()),
                                1, None))
    hb, hp = O.gen_pkfk(9, 1 << 40, nb, npr, miss)
    assert np.array_equal(db.cpu().numpy(), hb.view(np.uint8))
    assert np.array_equal(dp.cpu().numpy(), hp.view(np.uint8))


def test_queryhashjoin_shape():
    # unit_tests/queryhashjoin.cpp: (k, k) x (k, k + 0.1), projection B$1, P$1
    _say("case: queryhashjoin 20 keys")
    T = 20
    args = dict(build_cols=["long", "long"], probe_cols=["long", "dec"], buildjattr=[0], probejattr=[0],
                projection=["B$1", "P$1"])
    build = pack(["long", "long"], [(k, k) for k in range(1, T + 1)])
    probe = pack(["long", "dec"], [(k, k + 0.1) for k in range(1, T + 1)])
    got, _ = compare(args, build, probe, T, T)
    rows = unpack(["long", "dec"], got)
    assert sorted(int(v) for v, _ in rows) == list(range(1, T + 1))
    assert all(d == v + 0.1 for v, d in rows)


@pytest.mark.parametrize("nb,npr,bits", [(1 << 16, 1 << 20, -1), (100000, 300000, 0), (100000, 300000, 14),
                                         (3, 1000, -1), (9000, 50000, 0), (3, 1000, 14),
                                         (1 << 20, 3 << 20, 11)])
def test_pkfk_vs_oracle(nb, npr, bits):
    _say("case: pkfk nb=%d np=%d bits=%d" % (nb, npr, bits))
    b, p = O.gen_pkfk(nb + npr, 1 << 40, nb, npr, npr // 10)
    got, t = compare(HEADLINE, b.view(np.uint8), p.view(np.uint8), nb, npr, radix_bits=bits, timing=True)
    assert got.shape[0] == npr - npr // 10


def test_many_to_many_cross_width_keys():
    _say("case: many-to-many int/long keys")
    rng = np.random.default_rng(31)
    nb, npr = 3000, 20000
    bk = rng.integers(-300, 300, nb)
    pk = rng.integers(-300, 300, npr)
    build = pack(["int", "long", "int"], [(int(k), int(v), int(w)) for k, v, w in
                                          zip(bk, rng.integers(0, 1 << 62, nb), rng.integers(0, 1 << 31, nb))])
    probe = pack(["long", "int"], [(int(k), int(v)) for k, v in zip(pk, rng.integers(-(1 << 31), 1 << 31, npr))])
    args = dict(build_cols=["int", "long", "int"], probe_cols=["long", "int"], buildjattr=[0], probejattr=[0],
                projection=["P$1", "B$1", "B$0", "P$0"])
    got, t = compare(args, build, probe, nb, npr)
    assert got.shape[0] > 100000


def test_composite_int_keys():
    _say("case: composite (int, int) keys")
    rng = np.random.default_rng(32)
    n = 40000
    cols = ["int", "int", "long"]
    build = pack(cols, [(int(a), int(b), int(c)) for a, b, c in
                        zip(rng.integers(0, 50, n), rng.integers(0, 50, n), rng.integers(0, 1 << 62, n))])
    probe = pack(cols, [(int(a), int(b), int(c)) for a, b, c in
                        zip(rng.integers(0, 60, n), rng.integers(0, 60, n), rng.integers(0, 1 << 62, n))])
    args = dict(build_cols=cols, probe_cols=cols, buildjattr=[0, 1], probejattr=[0, 1],
                projection=["B$0", "B$1", "B$2", "P$2"])
    got, _ = compare(args, build, probe, n, n)
    assert got.shape[0] > 100000


def test_twelve_byte_tuples():
    _say("case: 12-byte tuples (long key, int payload)")
    rng = np.random.default_rng(33)
    nb, npr = 50000, 200000
    keys = rng.choice(1 << 40, nb, replace=False)
    build = pack(["long", "int"], [(int(k), int(v)) for k, v in zip(keys, rng.integers(0, 1 << 31, nb))])
    fk = keys[rng.integers(0, nb, npr)]
    probe = pack(["long", "int"], [(int(k), int(v)) for k, v in zip(fk, rng.integers(0, 1 << 31, npr))])
    args = dict(build_cols=["long", "int"], probe_cols=["long", "int"], buildjattr=[0], probejattr=[0],
                projection=["B$0", "B$1", "P$1"])
    got, _ = compare(args, build, probe, nb, npr)
    assert got.shape[0] == npr


def test_empty_and_all_miss():
    _say("case: empty / all-miss")
    b, p = O.gen_pkfk(3, 1 << 40, 1000, 5000, 5000)
    e = np.zeros(0, np.uint8)
    assert run_join(P.make_desc(**HEADLINE), e, p.view(np.uint8), 0, 5000)[0].shape[0] == 0
    assert run_join(P.make_desc(**HEADLINE), b.view(np.uint8), e, 1000, 0)[0].shape[0] == 0
    assert run_join(P.make_desc(**HEADLINE), b.view(np.uint8), p.view(np.uint8), 1000, 5000)[0].shape[0] == 0


def test_heavy_duplicate_build_key_spills_lds():
    # one key 20000 times: one partition larger than an LDS table, joined by
    # splitting the build side; every probe of that key meets all copies
    _say("case: heavy duplicate build key")
    nb, npr = 20000, 12
    build = np.stack([np.full(nb, 77, np.int64), np.arange(nb, dtype=np.int64)], 1).reshape(-1)
    probe = np.stack([np.array([77] * 10 + [5, 6], np.int64), np.arange(npr, dtype=np.int64)], 1).reshape(-1)
    got, t = run_join(P.make_desc(**HEADLINE, timing=True), build.view(np.uint8), probe.view(np.uint8), nb, npr,
                      cap=nb * 10)
    assert got.shape[0] == nb * 10
    assert t.spilled_items > 0
    exp, _ = O.hashjoin(oracle_desc(HEADLINE), build.view(np.uint8), probe.view(np.uint8), out_cap=nb * 10)
    assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))


def test_output_capacity_error_reports_need():
    _say("case: capacity error")
    b, p = O.gen_pkfk(4, 1 << 40, 1000, 5000, 0)
    with pytest.raises(P.PJError) as e:
        run_join(P.make_desc(**HEADLINE), b.view(np.uint8), p.view(np.uint8), 1000, 5000, cap=100)
    assert e.value.code == P.PJ_ERR_CAPACITY and "5000" in str(e.value)


def test_checksum_device_matches_oracle():
    _say("case: device checksum")
    rng = np.random.default_rng(5)
    for w in (16, 24, 12):
        rows = rng.integers(0, 256, size=(100003, w), dtype=np.uint8)
        d = to_gpu(rows)
        assert P.checksum_device(d.data_ptr(), rows.shape[0], w) == O.checksum(rows, w)


def test_partition_exactrange_vs_oracle():
    # PartitionOp: per-partition multisets equal the restated PartitionOp's
    _say("case: exactrange partition")
    T = 128 * 1024
    rng = np.random.default_rng(6)
    vals = rng.integers(1, T + 1, size=2 * T)
    rows = np.stack([vals, vals ^ 0xABCDEF], axis=1).astype(np.int64)
    for threads in (1, 5, 15, 127):
        exp = O.partition(O.schema(["long", "long"]), 0, 1, T, [rows.view(np.uint8).reshape(-1)])
        exp_multi = O.partition(O.schema(["long", "long"]), 0, 1, T,
                                [s.view(np.uint8).reshape(-1) for s in np.array_split(rows, threads)])
        din = to_gpu(rows)
        dout = torch.empty_like(din)
        sizes = (C.c_uint64 * threads)()
        sch = P.make_schema(["long", "long"])
        P.check(P.lib().pj_partition_exactrange_device(C.byref(sch), 0, 1, T, threads, C.c_void_p(din.data_ptr()),
                                                       rows.shape[0], C.c_void_p(dout.data_ptr()), sizes, None))
        out = dout.cpu().numpy()
        o = 0
        for j in range(threads):
            part = out[o * 16:(o + sizes[j]) * 16].reshape(-1, 16)
            ref = exp_multi[j].reshape(-1, 16)
            assert np.array_equal(O.sorted_rows(part), O.sorted_rows(ref))
            o += sizes[j]
        assert o == rows.shape[0]
        del exp


def test_page_protocol_resumes_like_getnext():
    # HashJoinOp::getNext returns full pages and resumes (join.cpp:716-719):
    # staged build, batched probe, output drained a few tuples at a time
    _say("case: page protocol")
    b, p = O.gen_pkfk(8, 1 << 40, 30000, 100000, 5000)
    j = P.Join(P.make_desc(**HEADLINE))
    L = P.lib()
    bb = b.view(np.uint8)
    for chunk in np.array_split(bb.reshape(-1, 16), 7):
        c = np.ascontiguousarray(chunk)
        P.check(L.pj_join_build_append(j.h, c.ctypes.data_as(C.c_void_p), c.shape[0]))
    P.check(L.pj_join_build_finish(j.h))
    rows = []
    cur = P.Cursor()
    page = np.zeros(24 * 1000, np.uint8)
    n = C.c_uint64()
    for batch in np.array_split(p.view(np.uint8).reshape(-1, 16), 3):
        bt = np.ascontiguousarray(batch)
        P.check(L.pj_join_probe(j.h, bt.ctypes.data_as(C.c_void_p), bt.shape[0], C.byref(cur),
                                page.ctypes.data_as(C.c_void_p), 1000, C.byref(n)))
        rows.append(page[: n.value * 24].copy())
        while cur.pending:
            P.check(L.pj_join_probe(j.h, None, 0, C.byref(cur), page.ctypes.data_as(C.c_void_p), 1000,
                                    C.byref(n)))
            assert n.value <= 1000
            rows.append(page[: n.value * 24].copy())
    L.pj_cursor_release(C.byref(cur))
    j.close()
    got = np.concatenate(rows).reshape(-1, 24)
    exp, _ = O.hashjoin(oracle_desc(HEADLINE), bb, p.view(np.uint8))
    assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))


def _full_pkfk(seed, dom, nb, npr, nmiss):
    db = torch.empty(nb * 16, dtype=torch.uint8, device="cuda")
    dp = torch.empty(npr * 16, dtype=torch.uint8, device="cuda")
    P.check(P.lib().pj_gen_pkfk(seed, dom, nb, npr, nmiss, C.c_void_p(db.data_ptr()), C.c_void_p(dp.data_ptr()),
                                1, None))
    j = P.Join(P.make_desc(**HEADLINE, timing=True))
    out = torch.empty(npr * 24, dtype=torch.uint8, device="cuda")
    n = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, out.data_ptr(), npr)
    s, x = P.checksum_device(out.data_ptr(), n, 24)
    j.close()
    del db, dp, out
    torch.cuda.empty_cache()
    return n, s


def test_full_1m_x_16m_dense_keys():
    # queryhashjoin shape at 1M x 16M: keys 0..1M-1, every probe matches
    _say("case: 1M x 16M")
    nb, npr = 1 << 20, 1 << 24
    n, s = _full_pkfk(101, nb, nb, npr, 0)
    es, ec = O.expected_pkfk(101, nb, nb, npr, 0, threads=16)
    assert (n, s) == (ec, es)


def test_full_16m_x_256m_uniform():
    # 16M x 256M, keys from [0, 2^40), 10% of probes miss
    _say("case: 16M x 256M")
    nb, npr = 1 << 24, 1 << 28
    nmiss = round(0.1 * npr)
    n, s = _full_pkfk(202, 1 << 40, nb, npr, nmiss)
    es, ec = O.expected_pkfk(202, 1 << 40, nb, npr, nmiss, threads=16)
    assert ec == npr - nmiss
    assert (n, s) == (ec, es)


def test_gpu_equals_reference_hashjoin():
    # the device join against the reference HashJoinOp itself (oracle/_ref),
    # where that build travelled with the tree
    import reference_run as RR
    if not RR.available():
        pytest.skip("oracle/_ref not built")
    _say("case: GPU vs reference HashJoinOp")
    b, p = O.gen_pkfk(55, 1 << 40, 200000, 1000000, 100000)
    ref = RR.run_pkfk(b, p, threads=8)
    got, _ = run_join(P.make_desc(**HEADLINE), b.view(np.uint8), p.view(np.uint8), 200000, 1000000)
    assert got.shape == ref["rows"].shape
    assert np.array_equal(O.sorted_rows(got), O.sorted_rows(ref["rows"]))


def _rows_in_order(rows):
    """24-byte rows of three int64 columns in a canonical order for a multiset
    compare: by the third column when its values are distinct (the probe
    payloads here, a sort of seconds at 16M rows), else by all three."""
    c = np.ascontiguousarray(rows).view("<u8").reshape(-1, 3)
    o = c[np.argsort(c[:, 2])]
    if np.all(o[1:, 2] > o[:-1, 2]):
        return o
    return c[np.lexsort((c[:, 2], c[:, 1], c[:, 0]))]


def test_full_1m_x_16m_sorted_compare_with_reference(tmp_path):
    """Config 1 at full size: the device join's 16.8M rows against the rows of
    the reference HashJoinOp itself (oracle/_ref, run in a process of its own
    on the same host inputs), as sorted multisets."""
    import os
    import subprocess
    import reference_run as RR
    if not RR.available():
        pytest.skip("oracle/_ref not built")
    nb, npr = 1 << 20, 1 << 24
    _say("case: 1M x 16M sorted compare, generating")
    b, p = O.gen_pkfk(103, nb, nb, npr, 0)
    fb, fp, fo = tmp_path / "b.npy", tmp_path / "p.npy", tmp_path / "rows.npy"
    np.save(fb, b)
    np.save(fp, p)
    _say("case: 1M x 16M sorted compare, reference HashJoinOp")
    threads = min(16, os.cpu_count() or 1)
    r = subprocess.run([sys.executable, RR.__file__, "rows", str(fb), str(fp), str(threads), str(fo)],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    ref = np.load(fo)
    assert ref.shape == (npr, 24)
    _say("case: 1M x 16M sorted compare, device join")
    got, _ = run_join(P.make_desc(**HEADLINE), b.view(np.uint8), p.view(np.uint8), nb, npr)
    assert got.shape == ref.shape
    _say("case: 1M x 16M sorted compare, sorting")
    assert np.array_equal(_rows_in_order(got), _rows_in_order(ref))


COMPOSITE = dict(build_cols=["int", "int", "long"], probe_cols=["int", "int", "long"], buildjattr=[0, 1],
                 probejattr=[0, 1], projection=["B$0", "B$1", "B$2", "P$2"])


def test_zipf_generator_device_equals_host():
    _say("case: zipf pair generator device == host")
    n, m = 5000, 60000
    cdf = O.zipf_cdf(n, 1.5)
    hb, hp = O.gen_pairs_zipf(11, n, m, cdf)
    db = torch.empty(n * 16, dtype=torch.uint8, device="cuda")
    dp = torch.empty(m * 16, dtype=torch.uint8, device="cuda")
    P.check(P.lib().pj_gen_pairs_zipf(11, n, m, cdf.ctypes.data_as(C.c_void_p), C.c_void_p(db.data_ptr()),
                                      C.c_void_p(dp.data_ptr()), 1, None))
    assert np.array_equal(db.cpu().numpy(), hb) and np.array_equal(dp.cpu().numpy(), hp)


@pytest.mark.parametrize("z", [0.5, 1.0, 1.5])
def test_zipf_composite_vs_oracle(z):
    # heavy hitters: at z = 1.5 one key takes ~38% of the probe side
    _say("case: zipf composite z=%.1f" % z)
    n, m = 40000, 400000
    cdf = O.zipf_cdf(n, z)
    b, p = O.gen_pairs_zipf(12, n, m, cdf)
    got, t = compare(COMPOSITE, b, p, n, m, timing=True)
    assert got.shape[0] == m


def _full_zipf(z, n, m, seed):
    cdf = O.zipf_cdf(n, z)
    db = torch.empty(n * 16, dtype=torch.uint8, device="cuda")
    dp = torch.empty(m * 16, dtype=torch.uint8, device="cuda")
    P.check(P.lib().pj_gen_pairs_zipf(seed, n, m, cdf.ctypes.data_as(C.c_void_p), C.c_void_p(db.data_ptr()),
                                      C.c_void_p(dp.data_ptr()), 1, None))
    j = P.Join(P.make_desc(**COMPOSITE, timing=True))
    out = torch.empty(m * 24, dtype=torch.uint8, device="cuda")
    cnt = j.run_device(db.data_ptr(), n, dp.data_ptr(), m, out.data_ptr(), m)
    s, _ = P.checksum_device(out.data_ptr(), cnt, 24)
    j.close()
    del db, dp, out
    torch.cuda.empty_cache()
    es, ec = O.expected_pairs_zipf(seed, n, m, cdf, threads=16)
    return cnt, s, ec, es


@pytest.mark.parametrize("z", [0.5, 1.0, 1.5])
def test_full_8m_x_128m_zipf_composite(z):
    _say("case: 8M x 128M zipf composite z=%.1f" % z)
    cnt, s, ec, es = _full_zipf(z, 1 << 23, 1 << 27, 303)
    assert (cnt, s) == (ec, es)


def _clustered(seed, n, m, z=1.0, **kw):
    t = P.cluster_table(seed, n, **kw)
    cdf = O.zipf_cdf(n, z)
    db = torch.empty(n * 16, dtype=torch.uint8, device="cuda")
    dp = torch.empty(m * 16, dtype=torch.uint8, device="cuda")
    P.check(P.lib().pj_gen_clustered(t.ctypes.data_as(C.c_void_p), m, cdf.ctypes.data_as(C.c_void_p),
                                     C.c_void_p(db.data_ptr()), C.c_void_p(dp.data_ptr()), 1, None))
    return t, cdf, db, dp


_PARTS = ["radix", "learned"]


@pytest.mark.parametrize("partitioner", _PARTS)
def test_clustered_vs_oracle(partitioner):
    _say("case: clustered keys, %s" % partitioner)
    n, m = 300000, 1200000
    t, cdf, db, dp = _clustered(45, n, m, nclusters=64, log2_wmin=10, log2_wmax=17)
    b = db.cpu().numpy()
    p = dp.cpu().numpy()
    got, tm = compare(HEADLINE, b, p, n, m, partitioner=partitioner, timing=True)
    assert got.shape[0] == m


@pytest.mark.parametrize("partitioner", _PARTS)
def test_full_64m_x_512m_clustered(partitioner):
    _say("case: 64M x 512M clustered, %s" % partitioner)
    n, m = 1 << 26, 1 << 29
    t, cdf, db, dp = _clustered(404, n, m)
    j = P.Join(P.make_desc(**HEADLINE, partitioner=partitioner, timing=True))
    out = torch.empty(m * 24, dtype=torch.uint8, device="cuda")
    cnt = j.run_device(db.data_ptr(), n, dp.data_ptr(), m, out.data_ptr(), m)
    s, _ = P.checksum_device(out.data_ptr(), cnt, 24)
    used = j.timings().used_learned
    j.close()
    del db, dp, out
    torch.cuda.empty_cache()
    es, ec = O.expected_clustered(t, m, cdf, threads=16)
    assert (cnt, s) == (ec, es)
    if partitioner == "learned":
        assert used == 1


def test_full_64m_x_512m_clustered_learned_probe_of_another_layout():
    """The learned partition function is evaluated by one histogram kernel
    per record layout: with 16-byte build tuples and 20-byte probe tuples
    (long fk, int, long payload) the two sides run different kernels, which
    must put every key in the same partition (one partition apart on a few
    hundred keys when the compiler is free to fuse their multiplies and
    subtracts differently). Same rows as the 16-byte join: the closed form."""
    _say("case: 64M x 512M clustered, learned, 20-byte probe tuples")
    n, m = 1 << 26, 1 << 29
    t, cdf, db, dp = _clustered(404, n, m)
    p20 = torch.empty((m, 20), dtype=torch.uint8, device="cuda")
    src = dp.view(m, 16)
    p20[:, :8] = src[:, :8]
    p20[:, 8:12] = 7
    p20[:, 12:] = src[:, 8:]
    del dp, src
    torch.cuda.empty_cache()
    desc = P.make_desc(["long", "long"], ["long", "int", "long"], [0], [0], ["B$0", "B$1", "P$2"],
                       partitioner="learned", timing=True)
    j = P.Join(desc)
    out = torch.empty(m * 24, dtype=torch.uint8, device="cuda")
    cnt = j.run_device(db.data_ptr(), n, p20.data_ptr(), m, out.data_ptr(), m)
    s, _ = P.checksum_device(out.data_ptr(), cnt, 24)
    used = j.timings().used_learned
    j.close()
    del db, p20, out
    torch.cuda.empty_cache()
    es, ec = O.expected_clustered(t, m, cdf, threads=16)
    assert used == 1
    assert (cnt, s) == (ec, es)


def _fin64(x):
    m = (1 << 64) - 1
    x ^= x >> 30
    x = (x * 0xbf58476d1ce4e5b9) & m
    x ^= x >> 27
    x = (x * 0x94d049bb133111eb) & m
    return x ^ (x >> 31)


@pytest.mark.parametrize("bits", [0, 4, 11, 14])
def test_keys_equal_to_padding_and_extremes(bits):
    """Real tuples whose keys equal the partition pass's padding keys (0, and
    the first key whose mix lands outside partition 0), -1, INT64_MIN and
    INT64_MAX, on both sides and repeated, among ordinary keys: every
    partition count, padding included, still joins exactly."""
    _say("case: padding / extreme keys bits=%d" % bits)
    pk = next(k for k in range(1, 1 << 20) if _fin64(k) >> 63)
    special = np.array([0, pk, -1, -(1 << 63), (1 << 63) - 1], np.int64)
    rng = np.random.default_rng(90 + bits)
    bkeys = np.concatenate([special, special[:2], rng.integers(-(1 << 40), 1 << 40, 50000)])
    pkeys = np.concatenate([np.repeat(special, 300), rng.choice(bkeys, 200000), rng.integers(0, 1 << 40, 1000)])
    rng.shuffle(pkeys)
    build = np.stack([bkeys, np.arange(bkeys.size, dtype=np.int64)], 1).reshape(-1)
    probe = np.stack([pkeys, np.arange(pkeys.size, dtype=np.int64)], 1).reshape(-1)
    got, _ = compare(HEADLINE, build.view(np.uint8), probe.view(np.uint8), bkeys.size, pkeys.size, radix_bits=bits)
    assert got.shape[0] >= 300 * 5
