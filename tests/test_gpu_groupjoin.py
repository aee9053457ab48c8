"""GPU parity of pj_join_count_by_key: aggregate_count grouped by the build
key over a join (aggregatecount.cpp:37-52 over HashJoinOp's output), computed
without the join's rows, against the oracle's aggregate over the oracle's
join, and at full size against the device join followed by the device
aggregation."""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle as O
import pjoin as P
from _util import pack, to_gpu

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


def expected(bcols, pcols, bk, pk, build, probe):
    """Oracle: the join projected to the build key columns, then the count
    per key."""
    proj = ["B$%d" % k for k in bk]
    d = O.join_desc(bcols, pcols, bk, pk, proj, buckets=1 << 12)
    rows, _ = O.hashjoin(d, build, probe)
    return O.aggregate([bcols[k] for k in bk], list(range(len(bk))), "count", None, rows)


def fused(bcols, pcols, bk, pk, build, probe, nb, npr, cap=None, **kw):
    desc = P.make_desc(bcols, pcols, bk, pk, ["B$%d" % bk[0], "P$0"], **kw)
    j = P.Join(desc)
    db = to_gpu(build) if build.size else torch.zeros(16, dtype=torch.uint8, device="cuda")
    dp = to_gpu(probe) if probe.size else torch.zeros(16, dtype=torch.uint8, device="cuda")
    width = 4 * len(bk) + 8 if all(bcols[k] == "int" for k in bk) else 16
    cap = max(nb, 1) if cap is None else cap
    out = torch.empty(max(cap, 1) * width, dtype=torch.uint8, device="cuda")
    g = j.count_by_key(db.data_ptr(), nb, dp.data_ptr(), npr, out.data_ptr(), cap)
    j.close()
    return out[: g * width].cpu().numpy().reshape(g, width)


def check(bcols, pcols, bk, pk, build, probe, nb, npr, **kw):
    exp = expected(bcols, pcols, bk, pk, build, probe)
    got = fused(bcols, pcols, bk, pk, build, probe, nb, npr, **kw)
    assert got.shape == exp.shape, (got.shape, exp.shape)
    assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))
    return got


@pytest.mark.parametrize("nb,npr,bits", [(1, 1, -1), (3, 1000, -1), (1 << 16, 1 << 20, -1), (100000, 300000, 14),
                                         (9000, 50000, 0), (1 << 20, 3 << 20, 11)])
def test_pkfk(nb, npr, bits):
    _say("case: count_by_key pkfk nb=%d np=%d bits=%d" % (nb, npr, bits))
    b, p = O.gen_pkfk(nb + npr + 11, 1 << 40, nb, npr, npr // 10)
    got = check(["long", "long"], ["long", "long"], [0], [0], b.view(np.uint8), p.view(np.uint8), nb, npr,
                radix_bits=bits)
    assert got.view(np.int64).reshape(-1, 2)[:, 1].sum() == npr - npr // 10


def test_many_to_many_cross_width_keys():
    """Repeated keys on both sides (the merge of a key's records), an INTEGER
    build key against a LONG probe key, negative keys."""
    _say("case: count_by_key many-to-many")
    rng = np.random.default_rng(51)
    nb, npr = 3000, 20000
    build = pack(["int", "long"], [(int(k), int(v)) for k, v in
                                   zip(rng.integers(-300, 300, nb), rng.integers(0, 1 << 62, nb))])
    probe = pack(["long", "int"], [(int(k), int(v)) for k, v in
                                   zip(rng.integers(-300, 300, npr), rng.integers(0, 1 << 31, npr))])
    got = check(["int", "long"], ["long", "int"], [0], [0], build, probe, nb, npr)
    assert got.shape[1] == 12


def test_one_key_split_over_tables():
    """One build key 20,000 times: its partition is split over several LDS
    tables, whose counts the merge adds up."""
    _say("case: count_by_key split partition")
    nb = 20000
    build = np.stack([np.full(nb, 77, np.int64), np.arange(nb, dtype=np.int64)], 1).reshape(-1)
    probe = np.stack([np.array([77] * 10 + [5, 6, 77], np.int64), np.arange(13, dtype=np.int64)], 1).reshape(-1)
    got = check(["long", "long"], ["long", "long"], [0], [0], build.view(np.uint8), probe.view(np.uint8), nb, 13)
    assert got.view(np.int64).reshape(-1).tolist() == [77, 11 * nb]


def test_composite_and_date_keys():
    _say("case: count_by_key composite / date keys")
    rng = np.random.default_rng(52)
    n = 40000
    cols = ["int", "int", "long"]
    build = pack(cols, [(int(a), int(b), int(c)) for a, b, c in
                        zip(rng.integers(-20, 20, n), rng.integers(0, 50, n), rng.integers(0, 1 << 62, n))])
    probe = pack(cols, [(int(a), int(b), int(c)) for a, b, c in
                        zip(rng.integers(-30, 30, n), rng.integers(0, 60, n), rng.integers(0, 1 << 62, n))])
    check(cols, cols, [0, 1], [0, 1], build, probe, n, n)
    dcols = ["date", "long"]
    keys = rng.choice(1 << 30, 50000, replace=False)
    build = pack(dcols, [(int(k), int(v)) for k, v in zip(keys, rng.integers(0, 1 << 62, 50000))])
    fk = keys[rng.integers(0, 50000, 200000)]
    probe = pack(dcols, [(int(k), int(v)) for k, v in zip(fk, rng.integers(0, 1 << 62, 200000))])
    check(dcols, dcols, [0], [0], build, probe, 50000, 200000)


def test_empty_and_capacity():
    _say("case: count_by_key empty / capacity")
    b, p = O.gen_pkfk(5, 1 << 40, 1000, 5000, 0)
    e = np.zeros(0, np.uint8)
    cols = ["long", "long"]
    assert fused(cols, cols, [0], [0], e, p.view(np.uint8), 0, 5000).shape[0] == 0
    assert fused(cols, cols, [0], [0], b.view(np.uint8), e, 1000, 0).shape[0] == 0
    with pytest.raises(P.PJError) as ex:
        fused(cols, cols, [0], [0], b.view(np.uint8), p.view(np.uint8), 1000, 5000, cap=10)
    assert ex.value.code == P.PJ_ERR_CAPACITY


def test_full_16m_x_256m_equals_join_then_aggregate():
    _say("case: count_by_key 16M x 256M")
    nb, npr = 1 << 24, 1 << 28
    db = torch.empty(nb * 16, dtype=torch.uint8, device="cuda")
    dp = torch.empty(npr * 16, dtype=torch.uint8, device="cuda")
    P.check(P.lib().pj_gen_pkfk(202, 1 << 40, nb, npr, round(0.1 * npr), C.c_void_p(db.data_ptr()),
                                C.c_void_p(dp.data_ptr()), 1, None))
    j = P.Join(P.make_desc(["long", "long"], ["long", "long"], [0], [0], ["B$0", "B$1", "P$1"]))
    fused_out = torch.empty(nb * 16, dtype=torch.uint8, device="cuda")
    g1 = j.count_by_key(db.data_ptr(), nb, dp.data_ptr(), npr, fused_out.data_ptr(), nb)
    rows = torch.empty(npr * 24, dtype=torch.uint8, device="cuda")
    n = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, rows.data_ptr(), npr)
    agg = P.Aggregate(["long"] * 3, [0], "count", None, max_groups=nb)
    unfused = torch.empty(nb * 16, dtype=torch.uint8, device="cuda")
    g2 = agg.run_device(rows.data_ptr(), n, unfused.data_ptr(), nb)
    agg.close()
    j.close()
    assert g1 == g2
    a = fused_out[: g1 * 16].view(torch.int64).view(-1, 2)
    b = unfused[: g2 * 16].view(torch.int64).view(-1, 2)
    a = a[torch.argsort(a[:, 0])]
    b = b[torch.argsort(b[:, 0])]
    assert torch.equal(a, b)
    assert int(a[:, 1].sum()) == n


def test_padding_and_extreme_keys():
    _say("case: count_by_key padding / extreme keys")
    from test_gpu_join import _fin64
    pk = next(k for k in range(1, 1 << 20) if _fin64(k) >> 63)
    special = np.array([0, pk, -1, -(1 << 63), (1 << 63) - 1], np.int64)
    rng = np.random.default_rng(61)
    bkeys = np.concatenate([special, rng.choice(1 << 40, 60000, replace=False)])
    pkeys = np.concatenate([np.repeat(special, 400), rng.choice(bkeys, 300000)])
    rng.shuffle(pkeys)
    build = np.stack([bkeys, np.arange(bkeys.size, dtype=np.int64)], 1).reshape(-1)
    probe = np.stack([pkeys, np.arange(pkeys.size, dtype=np.int64)], 1).reshape(-1)
    cols = ["long", "long"]
    for bits in (-1, 11, 14):
        check(cols, cols, [0], [0], build.view(np.uint8), probe.view(np.uint8), bkeys.size, pkeys.size,
              radix_bits=bits)


@pytest.mark.parametrize("bits", [1, 6, 11])
def test_appended_probe_side(bits):
    """The probe side appended (PJ_FLAG_LAYOUT_APPEND): the count per key
    reads its probe records through the chunk list; padding and extreme keys,
    a hot key and a key held by many build records."""
    _say("case: count_by_key, appended probe side, bits=%d" % bits)
    from test_gpu_join import _fin64
    pk = next(k for k in range(1, 1 << 20) if _fin64(k) >> 63)
    special = np.array([0, pk, -1, -(1 << 63), (1 << 63) - 1], np.int64)
    rng = np.random.default_rng(70 + bits)
    bkeys = np.concatenate([special, np.full(300, 12345), rng.choice(1 << 40, 40000, replace=False)])
    pkeys = np.concatenate([np.repeat(special, 200), np.full(50000, bkeys[900]), np.full(100, 12345),
                            rng.choice(bkeys, 250000)])
    rng.shuffle(pkeys)
    build = np.stack([bkeys, np.arange(bkeys.size, dtype=np.int64)], 1).reshape(-1)
    probe = np.stack([pkeys, np.arange(pkeys.size, dtype=np.int64)], 1).reshape(-1)
    cols = ["long", "long"]
    check(cols, cols, [0], [0], build.view(np.uint8), probe.view(np.uint8), bkeys.size, pkeys.size,
          radix_bits=bits, flags=P.PJ_FLAG_LAYOUT_APPEND)

