"""aggregate_sum / aggregate_count plan nodes through the C++ operators
(AggregateOp over the device aggregation), driven by plan text as the
reference's sample queries drive GenericAggregate
(drivers/sample_queries/q1/query1.conf:91-103). Plans rejected in init or
before their threads start are checked on the CPU; plans that run need the
GPU."""
import numpy as np
import pytest

import oracle as O
import pjoin as P

AGG_OVER_SCAN = """
buffsize = 4096;
src: { type = "memscan"; schema = ( %(schema)s ); };
agg: { type = "%(type)s"; %(fields)s %(sum)s hash: { fn = "modulo"; buckets = 1024; field = 0; }; %(global)s };
merge: { type = "merge"; threads = %(threads)d; };
treeroot: { name = "merge"; input: { name = "agg"; input: { name = "src"; }; }; };
"""


def plan(cols, fields, op, sumfield=None, threads=4, global_threads=None):
    schema = ", ".join('"%s"' % c for c in cols)
    f = ("field = %d;" % fields[0]) if len(fields) == 1 else ("fields = [ %s ];" % ", ".join(map(str, fields)))
    s = ("sumfield = %d;" % sumfield) if op == "sum" else ""
    g = ("global = \"yes\"; threads = %d;" % global_threads) if global_threads else ""
    return AGG_OVER_SCAN % dict(schema=schema, type="aggregate_" + op, fields=f, sum=s, threads=threads, **{"global": g})


def _rows(cols, n, seed, groups):
    rng = np.random.default_rng(seed)
    dt = np.dtype([("c%d" % i, O._NP[c]) for i, c in enumerate(cols)])
    t = np.zeros(n, dt)
    for i, c in enumerate(cols):
        hi = 2 ** 31 if c == "int" else 2 ** 40
        t["c%d" % i] = rng.integers(-hi, hi, n)
    t["c0"] = rng.integers(-groups, groups, n)
    return t.view(np.uint8).reshape(n, dt.itemsize)


def test_missing_field_is_rejected():
    p = plan(["long", "long"], [0], "count").replace("field = 0;", "")
    with pytest.raises(P.PJError) as e:
        P.run_plan(p.replace("hash: { fn = \"modulo\"; buckets = 1024; field = 0; };", ""), {"src": (np.zeros(16, np.uint8), 1)})
    assert "field" in str(e.value)


def test_decimal_sum_is_unsupported():
    with pytest.raises(P.PJError) as e:
        P.run_plan(plan(["long", "dec"], [0], "sum", 1), {"src": (np.zeros(16, np.uint8), 1)})
    assert e.value.code == P.PJ_ERR_UNSUPPORTED


def test_global_threads_must_match_the_merge():
    with pytest.raises(P.PJError) as e:
        P.run_plan(plan(["long", "long"], [0], "count", threads=4, global_threads=2),
                   {"src": (np.zeros(16, np.uint8), 1)})
    assert e.value.code == P.PJ_ERR_INVALID and "threads" in str(e.value)


@pytest.mark.gpu
@pytest.mark.parametrize("cols,fields,op,sumfield", [
    (["long", "long"], [0], "count", None),
    (["long", "int"], [0], "sum", 1),             # INTEGER sums wrap at 32 bits
    (["int", "int", "long"], [0, 1], "sum", 2),
])
def test_global_aggregation_plan_equals_oracle(cols, fields, op, sumfield):
    rows = _rows(cols, 300000, 7, 5000)
    exp = O.aggregate(cols, fields, op, sumfield, rows)
    got = P.run_plan(plan(cols, fields, op, sumfield, threads=4, global_threads=4),
                     {"src": (rows.reshape(-1), rows.shape[0])})
    assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))


@pytest.mark.gpu
def test_thread_local_aggregation_partials_add_up():
    """Thread-local mode (no "global" key, as in the reference's q1 plan): a
    key may come once per thread; the partial counts add up to the global
    ones, and no thread repeats a key."""
    cols = ["long", "long"]
    rows = _rows(cols, 200000, 8, 3000)
    exp = O.aggregate(cols, [0], "count", None, rows).view(np.int64).reshape(-1, 2)
    got = P.run_plan(plan(cols, [0], "count", threads=4), {"src": (rows.reshape(-1), rows.shape[0])})
    g = got.view(np.int64).reshape(-1, 2)
    assert exp.shape[0] <= g.shape[0] <= 4 * exp.shape[0]
    keys, inv = np.unique(g[:, 0], return_inverse=True)
    sums = np.zeros(keys.size, np.int64)
    np.add.at(sums, inv.reshape(-1), g[:, 1])
    e = exp[np.argsort(exp[:, 0])]
    assert np.array_equal(keys, e[:, 0]) and np.array_equal(sums, e[:, 1])
    one = P.run_plan(plan(cols, [0], "count", threads=1), {"src": (rows.reshape(-1), rows.shape[0])})
    assert np.array_equal(O.sorted_rows(one), O.sorted_rows(exp.view(np.uint8).reshape(-1, 16)))


@pytest.mark.gpu
@pytest.mark.parametrize("field", [0, 1])
def test_count_per_key_over_a_join_plan(field):
    """aggregate_count (global) over a hashjoin, the shape of a grouped join
    query, against the oracle's aggregate over the oracle's join. Grouped by
    the build key (field 0) the count is taken inside the join; grouped by
    the probe payload (field 1) over the join's rows."""
    b, p = O.gen_pkfk(33, 1 << 40, 20000, 100000, 10000)
    if field == 1:   # few distinct payloads, so that groups gather many rows
        p.reshape(-1, 2)[:, 1] %= 97
    text = """
buffsize = 65536;
build: { type = "memscan"; schema = ( "long", "long" ); };
probe: { type = "memscan"; schema = ( "long", "long" ); };
join: { type = "hashjoin"; hash: { fn = "modulo"; buckets = 32768; }; threadgroups = ( [0, 1, 2] );
        tuplesperbucket = 2; allocpolicy = "local"; buildjattr = 0; probejattr = 0;
        projection = ( "B$0", "P$1" ); };
agg: { type = "aggregate_count"; field = %d; hash: { fn = "modulo"; buckets = 32768; field = %d; };
       global = "yes"; threads = 3; };
merge: { type = "merge"; threads = 3; };
treeroot: { name = "merge"; input: { name = "agg"; input: { name = "join"; build: { name = "build"; };
            probe: { name = "probe"; }; }; }; };
""" % (field, field)
    got = P.run_plan(text, {"build": (b.view(np.uint8), 20000), "probe": (p.view(np.uint8), 100000)})
    d = O.join_desc(["long", "long"], ["long", "long"], [0], [0], ["B$0", "P$1"], buckets=1 << 15)
    jrows, _ = O.hashjoin(d, b.view(np.uint8), p.view(np.uint8))
    exp = O.aggregate(["long", "long"], [field], "count", None, jrows)
    assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))
