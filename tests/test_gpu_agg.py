"""GPU tests of the device aggregation (pj_agg_*): aggregate_sum and
aggregate_count in global mode against the oracle (oracle.aggregate, pinned
against the reference's aggregates), alone and over the output of a device
join."""
import sys

import numpy as np
import pytest

import oracle as O
import pjoin as P
from test_aggregate_oracle import CASES, _table

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


def _device_agg(cols, fields, op, sumfield, rows, max_groups=1 << 20):
    agg = P.Aggregate(cols, fields, op, sumfield, max_groups=max_groups)
    n = rows.shape[0]
    d = torch.from_numpy(np.ascontiguousarray(rows).reshape(-1).copy()).cuda() if n else \
        torch.zeros(16, dtype=torch.uint8, device="cuda")
    cap = max(n, 1)
    out = torch.empty(cap * agg.out_width, dtype=torch.uint8, device="cuda")
    g = agg.run_device(d.data_ptr(), n, out.data_ptr(), cap)
    got = out[: g * agg.out_width].cpu().numpy().reshape(g, agg.out_width)
    agg.close()
    return got


@pytest.mark.parametrize("cols,fields,op,sumfield,groups", CASES)
@pytest.mark.parametrize("n", [1, 1000, 1 << 20])
def test_device_aggregate_equals_oracle(cols, fields, op, sumfield, groups, n):
    _say("case: agg %s fields=%s %s n=%d" % (cols, fields, op, n))
    rng = np.random.default_rng(n)
    _, rows = _table(cols, n, rng, groups)
    exp = O.aggregate(cols, fields, op, sumfield, rows)
    got = _device_agg(cols, fields, op, sumfield, rows)
    assert got.shape == exp.shape
    assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))


@pytest.mark.parametrize("op", ["sum", "count"])
def test_device_aggregate_many_groups(op):
    """More distinct groups than a block's LDS table holds: most tuples fold
    straight into the global table."""
    _say("case: agg many groups %s" % op)
    rng = np.random.default_rng(8)
    cols = ["long", "int"]
    _, rows = _table(cols, 600000, rng, {0: 300000})
    exp = O.aggregate(cols, [0], op, 1, rows)
    got = _device_agg(cols, [0], op, 1, rows, max_groups=300000)
    assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))


def test_device_aggregate_key_equal_to_the_empty_slot_marker():
    """A LONG group key and a composite key whose bits equal the table's
    empty-slot marker are groups like any other."""
    _say("case: agg marker keys")
    marker = np.uint64(0xA5A5A5A5A5A5A5A5).view(np.int64)
    t = np.zeros(7, np.dtype([("c0", "<i8"), ("c1", "<i8")]))
    t["c0"] = [marker, 1, marker, 2, 1, marker, 0]
    t["c1"] = [5, 6, 7, 8, 9, 10, 11]
    rows = t.view(np.uint8).reshape(7, 16)
    for op in ("sum", "count"):
        exp = O.aggregate(["long", "long"], [0], op, 1, rows)
        got = _device_agg(["long", "long"], [0], op, 1, rows)
        assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))
    half = np.uint32(0xA5A5A5A5).view(np.int32)
    u = np.zeros(4, np.dtype([("a", "<i4"), ("b", "<i4"), ("v", "<i4")]))
    u["a"] = [half, half, 3, half]
    u["b"] = [half, half, 4, 0]
    u["v"] = [1, 2, 3, 4]
    rows = u.view(np.uint8).reshape(4, 12)
    exp = O.aggregate(["int", "int", "int"], [0, 1], "sum", 2, rows)
    got = _device_agg(["int", "int", "int"], [0, 1], "sum", 2, rows)
    assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))


def test_device_aggregate_empty_and_capacity():
    _say("case: agg empty / capacity")
    got = _device_agg(["long", "long"], [0], "sum", 1, np.zeros((0, 16), np.uint8))
    assert got.shape[0] == 0
    got = _device_agg(["long", "long"], [], "count", None, np.zeros((0, 16), np.uint8))
    assert got.shape[0] == 0
    rng = np.random.default_rng(9)
    _, rows = _table(["long", "long"], 5000, rng, {0: 4000})
    with pytest.raises(P.PJError) as e:    # more groups than the table's 1024 slots
        _device_agg(["long", "long"], [0], "count", None, rows, max_groups=16)
    assert e.value.code == P.PJ_ERR_CAPACITY
    agg = P.Aggregate(["long", "long"], [0], "count", None)
    d = torch.from_numpy(rows.reshape(-1).copy()).cuda()
    out = torch.empty(10 * agg.out_width, dtype=torch.uint8, device="cuda")
    with pytest.raises(P.PJError) as e:    # more groups than the output holds
        agg.run_device(d.data_ptr(), 5000, out.data_ptr(), 10)
    assert e.value.code == P.PJ_ERR_CAPACITY
    agg.close()


@pytest.mark.parametrize("fields,op,sumfield", [([0], "count", None), ([], "sum", 2), ([1], "sum", 2)])
def test_join_then_aggregate_on_the_device(fields, op, sumfield):
    """aggregate over hashjoin with the join's output left in HBM: count per
    build key, sum of the probe payload, sum per build payload."""
    _say("case: join -> agg %s %s" % (fields, op))
    nb, npr = 1 << 16, 1 << 20
    b, p = O.gen_pkfk(57, 1 << 40, nb, npr, npr // 10)
    proj = ["B$0", "B$1", "P$1"]
    d = O.join_desc(["long", "long"], ["long", "long"], [0], [0], proj, buckets=1 << 16)
    exp_join, _ = O.hashjoin(d, b.view(np.uint8), p.view(np.uint8))
    cols = ["long", "long", "long"]
    exp = O.aggregate(cols, fields, op, sumfield, exp_join)
    db = torch.from_numpy(b.view(np.uint8).copy()).cuda()
    dp = torch.from_numpy(p.view(np.uint8).copy()).cuda()
    j = P.Join(P.make_desc(["long", "long"], ["long", "long"], [0], [0], proj))
    out = torch.empty(npr * j.out_width, dtype=torch.uint8, device="cuda")
    n = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, out.data_ptr(), npr)
    agg = P.Aggregate(cols, fields, op, sumfield, max_groups=nb)
    groups = torch.empty(nb * agg.out_width, dtype=torch.uint8, device="cuda")
    g = agg.run_device(out.data_ptr(), n, groups.data_ptr(), nb)
    got = groups[: g * agg.out_width].cpu().numpy().reshape(g, agg.out_width)
    assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))
    j.close()
    agg.close()


# The partitioned path: at least 2^16 groups expected and 2^20 tuples.
PART_CASES = [
    (["long", "long"], [0], "count", None, {0: 200000}, 1 << 18),
    (["long", "long"], [0], "sum", 1, {0: 200000}, 1 << 18),
    (["int", "int", "int"], [0, 1], "sum", 2, {0: 400, 1: 500}, 1 << 17),   # INTEGER sums that wrap
    (["int", "long"], [0], "sum", 1, {0: 100000}, 1 << 17),
    (["long", "int", "date"], [2], "count", None, {2: 70000}, 1 << 16),
]


@pytest.mark.parametrize("cols,fields,op,sumfield,groups,max_groups", PART_CASES)
def test_partitioned_aggregate_equals_oracle(cols, fields, op, sumfield, groups, max_groups):
    _say("case: partitioned agg %s fields=%s %s max_groups=%d" % (cols, fields, op, max_groups))
    rng = np.random.default_rng(max_groups + len(cols))
    _, rows = _table(cols, 1 << 21, rng, groups)
    exp = O.aggregate(cols, fields, op, sumfield, rows)
    got = _device_agg(cols, fields, op, sumfield, rows, max_groups=max_groups)
    assert got.shape == exp.shape
    assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))


def test_aggregate_handle_reuse_across_paths():
    """One handle, runs in a row that take both paths: the global table's
    state carries over (whole-table clear, clear of the listed claims, or
    none), partitions with more groups than their LDS table send the rest to
    the global table, more claims than the claim list holds fall back to a
    whole-table scan, and the empty-slot marker key appears among many."""
    _say("case: agg handle reuse")
    cols = ["long", "long"]
    agg = P.Aggregate(cols, [0], "sum", 1, max_groups=1 << 16)   # 16 partitions of 8192 LDS slots
    cap = 1 << 18
    out = torch.empty(cap * agg.out_width, dtype=torch.uint8, device="cuda")
    marker = np.uint64(0xA5A5A5A5A5A5A5A5).view(np.int64)
    runs = [(120000, 600000, False),     # global table, more claims than listed
            (120000, 1 << 21, False),    # partitioned, overflowing partitions
            (30000, 1 << 20, False),     # partitioned, clears the listed claims
            (120000, 1 << 21, True),     # partitioned, marker key
            (60000, 600000, False),      # global table, every claim listed
            (50000, 1 << 20, False),     # partitioned after it
            (10, 1000, True)]            # small, marker key
    for run, (g, n, mark) in enumerate(runs):
        rng = np.random.default_rng(100 + run)
        t, rows = _table(cols, n, rng, {0: g})
        if mark:
            t["c0"][::97] = marker
        exp = O.aggregate(cols, [0], "sum", 1, rows)
        d = torch.from_numpy(rows.reshape(-1).copy()).cuda()
        ng = agg.run_device(d.data_ptr(), n, out.data_ptr(), cap)
        got = out[: ng * agg.out_width].cpu().numpy().reshape(ng, agg.out_width)
        assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp)), "run %d" % run
    agg.close()


def test_join_then_partitioned_count_by_key():
    """count per build key over a 2M-probe join's output in HBM: the
    partitioned path (2^17 expected groups)."""
    _say("case: join -> partitioned count by key")
    nb, npr = 1 << 17, 1 << 21
    b, p = O.gen_pkfk(58, 1 << 40, nb, npr, npr // 10)
    proj = ["B$0", "B$1", "P$1"]
    d = O.join_desc(["long", "long"], ["long", "long"], [0], [0], proj, buckets=1 << 17)
    exp_join, _ = O.hashjoin(d, b.view(np.uint8), p.view(np.uint8))
    cols = ["long", "long", "long"]
    exp = O.aggregate(cols, [0], "count", None, exp_join)
    db = torch.from_numpy(b.view(np.uint8).copy()).cuda()
    dp = torch.from_numpy(p.view(np.uint8).copy()).cuda()
    j = P.Join(P.make_desc(["long", "long"], ["long", "long"], [0], [0], proj))
    out = torch.empty(npr * j.out_width, dtype=torch.uint8, device="cuda")
    n = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, out.data_ptr(), npr)
    agg = P.Aggregate(cols, [0], "count", None, max_groups=nb)
    groups = torch.empty(nb * agg.out_width, dtype=torch.uint8, device="cuda")
    g = agg.run_device(out.data_ptr(), n, groups.data_ptr(), nb)
    got = groups[: g * agg.out_width].cpu().numpy().reshape(g, agg.out_width)
    assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))
    j.close()
    agg.close()


def test_partitioned_aggregate_padding_and_extreme_keys():
    """Group keys equal to the partition pass's padding keys (0, and the
    first key whose mix lands outside partition 0), -1, INT64_MIN and
    INT64_MAX among many groups: the padding is skipped, the real groups
    counted."""
    _say("case: partitioned agg padding / extreme keys")
    from test_gpu_join import _fin64
    pk = next(k for k in range(1, 1 << 20) if _fin64(k) >> 63)
    rng = np.random.default_rng(12)
    cols = ["long", "long"]
    t, rows = _table(cols, 1 << 21, rng, {0: 150000})
    special = np.array([0, pk, -1, -(1 << 63), (1 << 63) - 1], np.int64)
    t["c0"][: 5 * 1000] = np.repeat(special, 1000)
    for op in ("sum", "count"):
        exp = O.aggregate(cols, [0], op, 1, rows)
        got = _device_agg(cols, [0], op, 1, rows, max_groups=1 << 18)
        assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))
