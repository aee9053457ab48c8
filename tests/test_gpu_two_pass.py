"""GPU parity of the two-pass probe side (PJ_FLAG_LAYOUT_TWO_PASS): where
the build side has 2 to 8 times as many radix partitions as one
write-combining pass gives the probe side (radix bits 12 to 14 against 11),
the probe side is split again on the next bits within each partition, so
that the join reads each probe record once. Every case compares the rows
with the CPU oracle's HashJoinOp (sorted, bit for bit) and with the one-pass
layout (PJ_FLAG_LAYOUT_EXACT) on the same inputs, and checks which layout
ran.
"""
import sys

import numpy as np
import pytest

import oracle as O
import pjoin as P
from _util import to_gpu

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

HEADLINE = dict(build_cols=["long", "long"], probe_cols=["long", "long"], buildjattr=[0], probejattr=[0],
                projection=["B$0", "B$1", "P$1"])
TWO, EXACT = P.PJ_FLAG_LAYOUT_TWO_PASS, P.PJ_FLAG_LAYOUT_EXACT


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    P.lib()
    yield


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(args, build, probe, nb, npr, cap, **kw):
    """Output rows and the number of probe passes that ran."""
    j = P.Join(P.make_desc(**args, timing=True, **kw))
    db = to_gpu(build) if build.size else torch.zeros(16, dtype=torch.uint8, device="cuda")
    dp = to_gpu(probe) if probe.size else torch.zeros(16, dtype=torch.uint8, device="cuda")
    out = torch.empty(max(cap, 1) * j.out_width, dtype=torch.uint8, device="cuda")
    n = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, out.data_ptr(), max(cap, 1),
                     torch.cuda.current_stream().cuda_stream)
    rows = out[: n * j.out_width].cpu().numpy().reshape(-1, j.out_width)
    passes = j.timings().probe_passes
    j.close()
    return rows, passes


def _both(args, build, probe, nb, npr, two_expected=True, **kw):
    d = O.join_desc(args["build_cols"], args["probe_cols"], args["buildjattr"], args["probejattr"],
                    args["projection"], buckets=1 << 12)
    exp, _ = O.hashjoin(d, build, probe)
    exp = O.sorted_rows(exp)
    for flags in (TWO, EXACT):
        got, passes = _run(args, build, probe, nb, npr, cap=max(exp.shape[0], 1), flags=flags, **kw)
        assert got.shape == exp.shape, (flags, got.shape, exp.shape)
        assert np.array_equal(O.sorted_rows(got), exp), flags
        if npr:
            assert passes == (2 if flags == TWO and two_expected else 1), (flags, passes)
    return exp


def _tuples(keys, pay0=0):
    return np.stack([np.asarray(keys, np.int64), np.arange(pay0, pay0 + len(keys), dtype=np.int64)],
                    1).reshape(-1).view(np.uint8)


@pytest.mark.parametrize("nb,npr,bits", [(100000, 300000, 12), (100000, 300000, 13), (100000, 300000, 14),
                                         (3, 1000, 13), (9000, 50000, 12), (1 << 20, 3 << 20, 14),
                                         (100000, 300000, 11)])
def test_pkfk(nb, npr, bits):
    _say("case: two passes pkfk nb=%d np=%d bits=%d" % (nb, npr, bits))
    b, p = O.gen_pkfk(nb * 5 + npr + bits, 1 << 40, nb, npr, npr // 10)
    exp = _both(HEADLINE, b.view(np.uint8), p.view(np.uint8), nb, npr, two_expected=bits > 11, radix_bits=bits)
    assert exp.shape[0] == npr - npr // 10


def test_many_to_many():
    _say("case: two passes many-to-many")
    rng = np.random.default_rng(51)
    nb, npr = 4000, 60000
    exp = _both(HEADLINE, _tuples(rng.integers(-500, 500, nb)), _tuples(rng.integers(-520, 520, npr), 1 << 40),
                nb, npr, radix_bits=13)
    assert exp.shape[0] > 200000


@pytest.mark.parametrize("bits", [12, 14])
def test_one_probe_key_fills_its_partition(bits):
    _say("case: two passes, one hot probe key, bits=%d" % bits)
    rng = np.random.default_rng(52 + bits)
    bkeys = rng.integers(0, 1 << 40, 30000)
    pkeys = np.concatenate([np.full(600000, bkeys[7]), rng.choice(bkeys, 40000)])
    rng.shuffle(pkeys)
    exp = _both(HEADLINE, _tuples(bkeys), _tuples(pkeys, 1 << 33), bkeys.size, pkeys.size, radix_bits=bits)
    assert exp.shape[0] == pkeys.size


def _fin64(x):
    m = (1 << 64) - 1
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9 & m
    x ^= x >> 27
    x = (x * 0x94d049bb133111eb) & m
    return x ^ (x >> 31)


@pytest.mark.parametrize("bits", [12, 13])
def test_padding_and_extreme_keys(bits):
    """Keys equal to the first pass's padding keys, -1, INT64_MIN and
    INT64_MAX on both sides: the second pass drops only the padding records."""
    _say("case: two passes, padding / extreme keys bits=%d" % bits)
    pk = next(k for k in range(1, 1 << 20) if _fin64(k) >> 63)
    special = np.array([0, pk, -1, -(1 << 63), (1 << 63) - 1], np.int64)
    rng = np.random.default_rng(290 + bits)
    bkeys = np.concatenate([special, special[:2], rng.integers(-(1 << 40), 1 << 40, 50000)])
    pkeys = np.concatenate([np.repeat(special, 300), rng.choice(bkeys, 200000), rng.integers(0, 1 << 40, 1000)])
    rng.shuffle(pkeys)
    exp = _both(HEADLINE, _tuples(bkeys), _tuples(pkeys), bkeys.size, pkeys.size, radix_bits=bits)
    assert exp.shape[0] >= 300 * 5


def test_row_id_form():
    _say("case: two passes, row-id form")
    rng = np.random.default_rng(53)
    nb, npr = 20000, 200000
    ba = rng.integers(0, 3000, nb)
    bb = rng.integers(0, 3, nb)
    pi = rng.integers(0, nb, npr)
    bt = np.stack([ba, bb, np.arange(nb)], 1).astype(np.int64).reshape(-1).view(np.uint8)
    pt = np.stack([ba[pi], bb[pi] + (np.arange(npr) % 7 == 0), np.arange(npr)], 1).astype(np.int64)
    args = dict(build_cols=["long", "long", "long"], probe_cols=["long", "long", "long"], buildjattr=[0, 1],
                probejattr=[0, 1], projection=["B$0", "B$1", "B$2", "P$2"])
    exp = _both(args, bt, pt.reshape(-1).view(np.uint8), nb, npr, radix_bits=13)
    assert exp.shape[0] > npr // 2


def test_word_records():
    # (int, int) keys with an int64 payload: records assembled word by word
    _say("case: two passes, composite INTEGER keys")
    rng = np.random.default_rng(54)
    nb, npr = 30000, 150000
    bk = rng.integers(-1000, 1000, (nb, 2)).astype(np.int32)
    pidx = rng.integers(0, nb, npr)
    pk = bk[pidx].copy()
    pk[::9, 1] += 1
    bt = np.zeros(nb, dtype=[("a", "<i4"), ("b", "<i4"), ("c", "<i8")])
    bt["a"], bt["b"], bt["c"] = bk[:, 0], bk[:, 1], np.arange(nb)
    pt = np.zeros(npr, dtype=[("a", "<i4"), ("b", "<i4"), ("c", "<i8")])
    pt["a"], pt["b"], pt["c"] = pk[:, 0], pk[:, 1], np.arange(npr)
    args = dict(build_cols=["int", "int", "long"], probe_cols=["int", "int", "long"], buildjattr=[0, 1],
                probejattr=[0, 1], projection=["B$0", "B$1", "B$2", "P$2"])
    exp = _both(args, bt.view(np.uint8), pt.view(np.uint8), nb, npr, radix_bits=12)
    assert exp.shape[0] > 0


def test_empty_and_reused_handle():
    _say("case: two passes, empty sides and one handle over several sizes")
    b, p = O.gen_pkfk(7, 1 << 40, 1000, 5000, 500)
    e = np.zeros(0, np.uint8)
    rows, _ = _run(HEADLINE, e, p.view(np.uint8), 0, 5000, 5000, flags=TWO, radix_bits=13)
    assert rows.shape[0] == 0
    rows, _ = _run(HEADLINE, b.view(np.uint8), e, 1000, 0, 1, flags=TWO, radix_bits=13)
    assert rows.shape[0] == 0
    j = P.Join(P.make_desc(**HEADLINE, flags=TWO, radix_bits=12))
    d = O.join_desc(["long", "long"], ["long", "long"], [0], [0], ["B$0", "B$1", "P$1"], buckets=1 << 12)
    for i, (nb, npr) in enumerate([(50000, 400000), (3000, 9000), (80000, 700000), (1, 100)]):
        b, p = O.gen_pkfk(400 + i, 1 << 40, nb, npr, npr // 5)
        exp, _ = O.hashjoin(d, b.view(np.uint8), p.view(np.uint8))
        db, dp = to_gpu(b.view(np.uint8)), to_gpu(p.view(np.uint8))
        out = torch.empty(npr * j.out_width, dtype=torch.uint8, device="cuda")
        n = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, out.data_ptr(), npr,
                         torch.cuda.current_stream().cuda_stream)
        got = out[: n * j.out_width].cpu().numpy().reshape(-1, j.out_width)
        assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp)), (nb, npr)
    j.close()
