"""Pins the aggregation oracle (oracle.aggregate): against the reference's
AggregateSum / AggregateCount in global mode, compiled from its sources into
oracle/_ref, and against properties any correct aggregation has (the counts
add up to the input, the groups are the distinct key values, the sums of the
groups add up to the sum of the input, INTEGER sums wrap at 32 bits)."""
import numpy as np
import pytest

import oracle as O
import reference_run as RR

ref_only = pytest.mark.skipif(not RR.available(), reason="oracle/_ref not built (reference tree absent)")


def _table(cols, n, rng, groups):
    dt = np.dtype([("c%d" % i, O._NP[c]) for i, c in enumerate(cols)])
    t = np.zeros(n, dt)
    for i, c in enumerate(cols):
        if c == "int":
            t["c%d" % i] = rng.integers(-2 ** 31, 2 ** 31, n, dtype=np.int64).astype(np.int32)
        else:
            t["c%d" % i] = rng.integers(-2 ** 63, 2 ** 63 - 1, n, dtype=np.int64)
    for f, g in groups.items():   # grouping columns drawn from g distinct values
        vals = rng.integers(-2 ** 31 if cols[f] == "int" else -2 ** 62, 2 ** 31 if cols[f] == "int" else 2 ** 62, g)
        t["c%d" % f] = vals[rng.integers(0, g, n)]
    return t, t.view(np.uint8).reshape(n, dt.itemsize)


CASES = [
    (["long", "long"], [0], "sum", 1, {0: 300}),
    (["int", "int"], [0], "sum", 1, {0: 50}),           # INTEGER sums that wrap
    (["int", "long", "long"], [0], "count", None, {0: 1000}),
    (["long", "int"], [], "sum", 1, {}),
    (["long", "int"], [], "count", None, {}),
    (["int", "int", "long"], [0, 1], "sum", 2, {0: 7, 1: 11}),
    (["date", "long"], [0], "sum", 1, {0: 64}),
]


@ref_only
@pytest.mark.parametrize("cols,fields,op,sumfield,groups", CASES)
@pytest.mark.parametrize("threads", [1, 4])
def test_oracle_aggregate_equals_reference(cols, fields, op, sumfield, groups, threads):
    rng = np.random.default_rng(71 + threads)
    t, rows = _table(cols, 20000, rng, groups)
    exp = O.aggregate(cols, fields, op, sumfield, rows)
    fn = "bytes" if "date" in [cols[f] for f in fields] else "modulo"   # the modulo hasher takes no DATE
    got, _ = RR.run_aggregate(cols, fields, op, sumfield, rows, threads=threads, fn=fn, buckets=256)
    assert got.shape == exp.shape
    assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))


@ref_only
def test_reference_aggregate_of_nothing_is_empty():
    rows = np.zeros((0, 16), np.uint8)
    got, _ = RR.run_aggregate(["long", "long"], [0], "sum", 1, rows, threads=2)
    assert got.shape[0] == 0
    assert O.aggregate(["long", "long"], [0], "sum", 1, rows).shape[0] == 0


@pytest.mark.parametrize("cols,fields,op,sumfield,groups", CASES)
def test_oracle_aggregate_properties(cols, fields, op, sumfield, groups):
    rng = np.random.default_rng(5)
    n = 5000
    t, rows = _table(cols, n, rng, groups)
    out = O.aggregate(cols, fields, op, sumfield, rows)
    agg_t = "<i8" if op == "count" else O._NP[cols[sumfield]]
    odt = np.dtype([("g%d" % k, O._NP[cols[f]]) for k, f in enumerate(fields)] + [("a", agg_t)])
    o = out.view(odt).reshape(-1)
    # the groups are the distinct values of the grouping columns, each once
    if fields:
        keyset = {tuple(int(t["c%d" % f][i]) for f in fields) for i in range(n)}
        got = [tuple(int(o["g%d" % k][i]) for k in range(len(fields))) for i in range(o.size)]
        assert len(got) == len(set(got)) == len(keyset) and set(got) == keyset
    else:
        assert o.size == 1
    if op == "count":
        assert int(o["a"].astype(np.int64).sum()) == n
    else:
        total = int(t["c%d" % sumfield].astype(object).sum())
        bits = 32 if agg_t == "<i4" else 64
        s = int(o["a"].astype(object).sum())
        assert (s - total) % (1 << bits) == 0   # sums of the groups = sum of the input, modulo the wrap


def test_oracle_integer_sum_wraps_at_32_bits():
    t = np.zeros(3, np.dtype([("c0", "<i4"), ("c1", "<i4")]))
    t["c1"] = [2 ** 31 - 1, 2 ** 31 - 1, 5]
    out = O.aggregate(["int", "int"], [0], "sum", 1, t.view(np.uint8).reshape(3, 8))
    assert out.view(np.dtype([("g", "<i4"), ("a", "<i4")]))["a"][0] == np.int32(2 ** 32 + 3 - 2 ** 32)   # 3
