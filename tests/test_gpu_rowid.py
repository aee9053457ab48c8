"""GPU parity of the row-id form: joins whose keys or projections the 16-byte
record cannot hold (two LONG key columns, three or more key columns, more
than 8 payload bytes a side, CHAR columns, tuple widths that are not a
multiple of 4).

Each case runs the HIP path through the C ABI (or a plan through the C++
operators) and compares the sorted output rows byte for byte with the
restated HashJoinOp (oracle/oracle.c) and, where oracle/_ref is built, with
the reference's own HashJoinOp on the same inputs.
"""
import sys

import numpy as np
import pytest

import oracle as O
import pjoin as P
import reference_run as RR
from _util import dtype_of, pack, to_gpu, unpack

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


def table(cols, n, rng, gen):
    """n packed tuples; gen[i](rng, n) gives column i (default: random bits)."""
    a = np.zeros(n, dtype=dtype_of(cols))
    for i, c in enumerate(cols):
        f = "f%d" % i
        if i in gen:
            a[f] = gen[i](rng, n)
        elif isinstance(c, tuple):
            raw = rng.integers(0, 256, (n, c[1]), dtype=np.uint8)
            raw[rng.random((n, c[1])) < 0.2] = 0   # NULs inside, with bytes after them
            a[f] = raw.view("S%d" % c[1]).reshape(n)
        elif c == "int":
            a[f] = rng.integers(-(1 << 31), 1 << 31, n)
        elif c == "dec":
            a[f] = rng.standard_normal(n)
        else:
            a[f] = rng.integers(-(1 << 63), (1 << 63) - 1, n, dtype=np.int64)
    return a.view(np.uint8).reshape(-1).copy()


def run_device(args, build, probe, nb, npr, cap=None, flags=0):
    d = P.make_desc(**args)
    d.flags |= flags
    j = P.Join(d)
    db = to_gpu(build) if build.size else torch.zeros(16, dtype=torch.uint8, device="cuda")
    dp = to_gpu(probe) if probe.size else torch.zeros(16, dtype=torch.uint8, device="cuda")
    cap = cap if cap is not None else max(npr, 1)
    out = torch.empty(max(cap, 1) * j.out_width + 16, dtype=torch.uint8, device="cuda")
    try:
        n = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, out.data_ptr(), cap,
                         torch.cuda.current_stream().cuda_stream)
        rows = out[: n * j.out_width].cpu().numpy().reshape(-1, j.out_width)
    finally:
        j.close()
    return rows


def oracle_rows(args, build, probe, fn="modulo"):
    d = O.join_desc(args["build_cols"], args["probe_cols"], args["buildjattr"], args["probejattr"],
                    args["projection"], fn=fn, buckets=1 << 12)
    rows, _ = O.hashjoin(d, build, probe)
    return rows


def check(args, build, probe, nb, npr, flags=0, ref=True, min_rows=1):
    exp = oracle_rows(args, build, probe)
    got = run_device(args, build, probe, nb, npr, cap=max(exp.shape[0], 1), flags=flags)
    assert got.shape == exp.shape, (got.shape, exp.shape)
    assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))
    assert exp.shape[0] >= min_rows
    if ref and RR.available():
        r, _, _ = RR.run_join(args["build_cols"], args["probe_cols"], args["buildjattr"], args["probejattr"],
                              args["projection"], build, probe, threads=4, fn="modulo", buckets=1 << 12)
        assert np.array_equal(O.sorted_rows(got), O.sorted_rows(r))
    return got


# unit_tests/queryhashjoincomposite.cpp:98-200: (long, long, long) against
# (long, long, dec), keys (0, 1), projection B$2, P$2; only (i, 2) matches.
COMPOSITE_PLAN = """
buffsize = 32;
build: { type = "memscan"; schema = ( "long", "long", "long" ); };
probe: { type = "memscan"; schema = ( "long", "long", "dec" ); };
join:
{
  type = "hashjoin";
  hash: { fn = "%(fn)s"; range = [ 1, 20 ]; buckets = 16; };
  threadgroups = ( [ 0, 1, 2, 3 ] );
  tuplesperbucket = 2;
  allocpolicy = "local";
  buildjattr = [ 0, 1 ];
  probejattr = [ 0, 1 ];
  projection = ( "B$2", "P$2" );
};
merge: { type = "merge"; threads = 4; };
treeroot: { name: "merge"; input: { name: "join"; build: { name: "build"; }; probe: { name: "probe"; }; }; };
"""


@pytest.mark.parametrize("fn", ["modulo", "bytes"])
def test_reference_composite_plan_long_long_keys(fn):
    _say("case: queryhashjoincomposite plan, fn=%s" % fn)
    T = 20
    build = pack(["long", "long", "long"], [(i, k, i) for i in range(1, T + 1) for k in (1, 2, 3, 4)])
    probe = pack(["long", "long", "dec"], [(i, k, i + 0.1) for i in range(1, T + 1) for k in (5, 2, 6, 7)])
    rows = P.run_plan(COMPOSITE_PLAN % dict(fn=fn), {"build": (build, 4 * T), "probe": (probe, 4 * T)})
    got = sorted((int(v), float(d)) for v, d in unpack(["long", "dec"], rows))
    assert got == [(i, i + 0.1) for i in range(1, T + 1)]   # the test's own check (queryhashjoincomposite.cpp:82-92)
    if RR.available():
        ref, _, _ = RR.run_join(["long", "long", "long"], ["long", "long", "dec"], [0, 1], [0, 1], ["B$2", "P$2"],
                                build, probe, threads=4, fn=fn, buckets=16, range_min=1, range_max=T, buffsize=32)
        assert np.array_equal(O.sorted_rows(rows), O.sorted_rows(ref))


LL = dict(build_cols=["long", "long", "long"], probe_cols=["long", "long", "dec"], buildjattr=[0, 1],
          probejattr=[0, 1], projection=["B$2", "P$2", "B$0", "P$1"])


def _ll_tables(nb, npr, seed):
    rng = np.random.default_rng(seed)
    spread = np.int64(0x100000001)
    a = lambda r, n, hi: (r.integers(0, hi, n) * spread - (1 << 40)).astype(np.int64)
    build = table(LL["build_cols"], nb, rng, {0: lambda r, n: a(r, n, nb // 2), 1: lambda r, n: r.integers(-2, 2, n)})
    probe = table(LL["probe_cols"], npr, rng, {0: lambda r, n: a(r, n, nb // 2 + nb // 20),
                                               1: lambda r, n: r.integers(-2, 2, n)})
    return build, probe


def test_long_long_composite_40k_x_400k():
    _say("case: (long, long) composite 40K x 400K, many-to-many")
    nb, npr = 40000, 400000
    build, probe = _ll_tables(nb, npr, 61)
    check(LL, build, probe, nb, npr, min_rows=150000)


def test_three_key_columns_across_widths():
    # IntLongEqual / LongIntEqual on every column: INTEGER values widened,
    # LONG values outside the INTEGER range never equal an INTEGER
    _say("case: three key columns, INTEGER against LONG")
    args = dict(build_cols=["int", "long", "int", "long"], probe_cols=["long", "int", "long", "long"],
                buildjattr=[0, 1, 2], probejattr=[0, 1, 2], projection=["B$0", "B$3", "P$3", "P$1", "B$2"])
    rng = np.random.default_rng(62)
    nb, npr = 20000, 100000
    small = lambda r, n: r.integers(-40, 40, n)
    wide = lambda r, n: np.where(r.random(n) < 0.1, r.integers(1 << 32, 1 << 40, n), r.integers(-40, 40, n))
    build = table(args["build_cols"], nb, rng, {0: small, 1: small, 2: lambda r, n: r.integers(-3, 3, n)})
    probe = table(args["probe_cols"], npr, rng, {0: wide, 1: small, 2: lambda r, n: r.integers(-3, 3, n)})
    check(args, build, probe, nb, npr, min_rows=10000)


def test_24_byte_payloads_48_byte_rows():
    # constructOutputTuple over three payload columns a side (join.cpp:648-679)
    _say("case: 24-byte payloads a side, 48-byte rows")
    cols = ["long", "long", "dec", "long"]
    args = dict(build_cols=cols, probe_cols=cols, buildjattr=[0], probejattr=[0],
                projection=["B$1", "B$2", "B$3", "P$1", "P$2", "P$3"])
    rng = np.random.default_rng(63)
    nb, npr = 50000, 500000
    build = table(cols, nb, rng, {0: lambda r, n: r.permutation(n).astype(np.int64) * 7919 - 99})
    probe = table(cols, npr, rng, {0: lambda r, n: r.integers(0, nb + nb // 10, n) * 7919 - 99})
    got = check(args, build, probe, nb, npr, min_rows=400000)
    assert got.shape[1] == 48


def test_char_columns_and_odd_tuple_widths():
    # CHAR columns copied as writeData's strncpy (schema.inl:75-82), 17- and
    # 7-byte tuples, an INTEGER probe key against a LONG build key
    _say("case: CHAR projection, odd tuple widths")
    args = dict(build_cols=[("char", 5), "long", "int"], probe_cols=["int", ("char", 3)], buildjattr=[1],
                probejattr=[0], projection=["B$0", "P$1", "B$2", "B$1", "P$0"])
    rng = np.random.default_rng(64)
    nb, npr = 10000, 100000
    build = table(args["build_cols"], nb, rng, {1: lambda r, n: r.integers(-6000, 6000, n)})
    probe = table(args["probe_cols"], npr, rng, {0: lambda r, n: r.integers(-7000, 7000, n)})
    got = check(args, build, probe, nb, npr, min_rows=50000)
    assert got.shape[1] == 24


def test_unequal_keys_with_equal_hashes_are_dropped():
    # PJ_FLAG_NARROW_ROW_HASH keeps 10 bits of the key hash: most pairs the
    # join finds have unequal keys, and the key check on the tuples decides
    _say("case: 10-bit row hash, key check decides")
    nb, npr = 6000, 60000
    build, probe = _ll_tables(nb, npr, 65)
    got = check(LL, build, probe, nb, npr, flags=P.PJ_FLAG_NARROW_ROW_HASH, ref=False, min_rows=10000)
    full = run_device(LL, build, probe, nb, npr, cap=got.shape[0])
    assert np.array_equal(O.sorted_rows(got), O.sorted_rows(full))


@pytest.mark.parametrize("flags", [0, P.PJ_FLAG_NARROW_ROW_HASH])
def test_capacity_error_reports_the_exact_row_count(flags):
    _say("case: row-id capacity error, flags=%d" % flags)
    nb, npr = 5000, 50000
    build, probe = _ll_tables(nb, npr, 66)
    n = oracle_rows(LL, build, probe).shape[0]
    with pytest.raises(P.PJError) as e:
        run_device(LL, build, probe, nb, npr, cap=n // 3, flags=flags)
    assert e.value.code == P.PJ_ERR_CAPACITY and ("need %d output" % n) in str(e.value)
    got = run_device(LL, build, probe, nb, npr, cap=n, flags=flags)
    assert got.shape[0] == n


def test_one_key_on_20000_build_tuples_splits_tables():
    # a partition larger than an LDS table, in the row-id form
    _say("case: row-id, one key 20,000 times")
    nb, npr = 24000, 3000
    rng = np.random.default_rng(67)
    build = table(LL["build_cols"], nb, rng, {0: lambda r, n: np.where(np.arange(n) < 20000, 5, r.integers(0, 99, n)),
                                              1: lambda r, n: np.zeros(n, np.int64)})
    probe = table(LL["probe_cols"], npr, rng, {0: lambda r, n: np.where(np.arange(n) < 3, 5, r.integers(100, 200, n)),
                                               1: lambda r, n: np.zeros(n, np.int64)})
    check(LL, build, probe, nb, npr, ref=False, min_rows=60000)


@pytest.mark.parametrize("nb,npr", [(0, 1000), (1000, 0), (0, 0)])
def test_row_id_empty_sides(nb, npr):
    _say("case: row-id empty sides %d x %d" % (nb, npr))
    build, probe = _ll_tables(max(nb, 2), max(npr, 2), 68)
    build = build[: nb * 24]
    probe = probe[: npr * 24]
    assert run_device(LL, build, probe, nb, npr).shape[0] == 0


def test_row_id_plan_pages_of_one_tuple():
    # HashJoinOp::getNext resuming after every output tuple (join.cpp:716-719)
    _say("case: row-id plan, one-tuple pages")
    nb, npr = 3000, 30000
    build, probe = _ll_tables(nb, npr, 69)
    plan = (COMPOSITE_PLAN % dict(fn="modulo")).replace("buffsize = 32;", "buffsize = 16;")
    rows = P.run_plan(plan, {"build": (build, nb), "probe": (probe, npr)}, out_cap_bytes=npr * 64)
    args = dict(LL, projection=["B$2", "P$2"])
    exp = oracle_rows(args, build, probe)
    assert rows.shape == exp.shape
    assert np.array_equal(O.sorted_rows(rows), O.sorted_rows(exp))
