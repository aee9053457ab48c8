"""GPU parity of the appended probe layout (PJ_FLAG_LAYOUT_APPEND): the
probe side partitioned without a histogram pass, each scatter block
appending whole 8-record chunks, the chunks then listed partition by
partition and the join reading its probe records through that list. Every
case compares the rows with the CPU oracle's HashJoinOp (sorted, bit for
bit) and with the contiguous layout (PJ_FLAG_LAYOUT_EXACT) on the same
inputs. Cases the appended layout does not take (one partition, build
partitions finer than the probe partitions at 14 radix bits) run the
contiguous layout under either flag and are compared all the same.
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
APPEND, EXACT = P.PJ_FLAG_LAYOUT_APPEND, P.PJ_FLAG_LAYOUT_EXACT


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    P.lib()
    yield


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(args, build, probe, nb, npr, cap, **kw):
    j = P.Join(P.make_desc(**args, **kw))
    db = to_gpu(build) if build.size else torch.zeros(16, dtype=torch.uint8, device="cuda")
    dp = to_gpu(probe) if probe.size else torch.zeros(16, dtype=torch.uint8, device="cuda")
    out = torch.empty(max(cap, 1) * j.out_width, dtype=torch.uint8, device="cuda")
    n = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, out.data_ptr(), max(cap, 1),
                     torch.cuda.current_stream().cuda_stream)
    rows = out[: n * j.out_width].cpu().numpy().reshape(-1, j.out_width)
    j.close()
    return rows


def _both(args, build, probe, nb, npr, **kw):
    """The appended and the contiguous layout against the oracle."""
    d = O.join_desc(args["build_cols"], args["probe_cols"], args["buildjattr"], args["probejattr"],
                    args["projection"], buckets=1 << 12)
    exp, _ = O.hashjoin(d, build, probe)
    exp = O.sorted_rows(exp)
    for flags in (APPEND, EXACT):
        got = _run(args, build, probe, nb, npr, cap=max(exp.shape[0], 1), flags=flags, **kw)
        assert got.shape == exp.shape, (flags, got.shape, exp.shape)
        assert np.array_equal(O.sorted_rows(got), exp), flags
    return exp


def _tuples(keys, pay0=0):
    return np.stack([np.asarray(keys, np.int64), np.arange(pay0, pay0 + len(keys), dtype=np.int64)],
                    1).reshape(-1).view(np.uint8)


@pytest.mark.parametrize("nb,npr,bits", [(1 << 16, 1 << 20, -1), (100000, 300000, 0), (100000, 300000, 1),
                                         (100000, 300000, 4), (100000, 300000, 10), (100000, 300000, 11),
                                         (100000, 300000, 14), (3, 1000, -1), (3, 1000, 11), (9000, 50000, 10),
                                         (1 << 20, 3 << 20, 11)])
def test_pkfk(nb, npr, bits):
    _say("case: appended layout pkfk nb=%d np=%d bits=%d" % (nb, npr, bits))
    b, p = O.gen_pkfk(nb * 7 + npr, 1 << 40, nb, npr, npr // 10)
    exp = _both(HEADLINE, b.view(np.uint8), p.view(np.uint8), nb, npr, radix_bits=bits)
    assert exp.shape[0] == npr - npr // 10


def test_many_to_many():
    _say("case: appended layout many-to-many")
    rng = np.random.default_rng(41)
    nb, npr = 4000, 60000
    exp = _both(HEADLINE, _tuples(rng.integers(-500, 500, nb)), _tuples(rng.integers(-520, 520, npr), 1 << 40),
                nb, npr, radix_bits=6)
    assert exp.shape[0] > 200000


@pytest.mark.parametrize("bits", [1, 6, 11])
def test_one_probe_key_fills_its_partition(bits):
    # 600K probes of one key and 40K of others: one partition's buffer
    # overflows by whole chunks every round (records that leave at slots of
    # their own, the rest refilling the buffer)
    _say("case: appended layout, one hot probe key, bits=%d" % bits)
    rng = np.random.default_rng(42 + bits)
    bkeys = rng.integers(0, 1 << 40, 30000)
    pkeys = np.concatenate([np.full(600000, bkeys[7]), rng.choice(bkeys, 40000)])
    rng.shuffle(pkeys)
    exp = _both(HEADLINE, _tuples(bkeys), _tuples(pkeys, 1 << 33), bkeys.size, pkeys.size, radix_bits=bits)
    assert exp.shape[0] == pkeys.size


@pytest.mark.parametrize("bits", [10, 11])
def test_hot_and_cold_rounds(bits):
    # the probe side alternates runs of 6000 uniform probes with runs of
    # 6000 probes of which a third are four hot keys: the scatter's rounds
    # with few whole chunks of overflow and those with many (stored by two
    # different loops) follow each other within every block
    _say("case: appended layout, hot and cold rounds, bits=%d" % bits)
    rng = np.random.default_rng(60 + bits)
    bkeys = rng.choice(1 << 40, 200000, replace=False)
    runs = []
    for r in range(300):
        run = rng.choice(bkeys, 6000)
        if r % 2:
            hot = rng.random(6000) < 1 / 3
            run[hot] = bkeys[rng.integers(0, 4, int(hot.sum()))]
        runs.append(run)
    pkeys = np.concatenate(runs)
    exp = _both(HEADLINE, _tuples(bkeys), _tuples(pkeys, 1 << 35), bkeys.size, pkeys.size, radix_bits=bits)
    assert exp.shape[0] == pkeys.size


def _fin64(x):
    m = (1 << 64) - 1
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9 & m
    x ^= x >> 27
    x = (x * 0x94d049bb133111eb) & m
    return x ^ (x >> 31)


@pytest.mark.parametrize("bits", [1, 4, 11])
def test_padding_and_extreme_keys(bits):
    """Keys equal to the scatter's padding keys, -1, INT64_MIN and INT64_MAX
    on both sides: the padding records of each block's last chunks never
    match."""
    _say("case: appended layout, padding / extreme keys bits=%d" % bits)
    pk = next(k for k in range(1, 1 << 20) if _fin64(k) >> 63)
    special = np.array([0, pk, -1, -(1 << 63), (1 << 63) - 1], np.int64)
    rng = np.random.default_rng(190 + bits)
    bkeys = np.concatenate([special, special[:2], rng.integers(-(1 << 40), 1 << 40, 50000)])
    pkeys = np.concatenate([np.repeat(special, 300), rng.choice(bkeys, 200000), rng.integers(0, 1 << 40, 1000)])
    rng.shuffle(pkeys)
    exp = _both(HEADLINE, _tuples(bkeys), _tuples(pkeys), bkeys.size, pkeys.size, radix_bits=bits)
    assert exp.shape[0] >= 300 * 5


def test_heavy_duplicate_build_key():
    # one key 20000 times on the build side: tables split, every probe of
    # the key meets all copies through the listed chunks
    _say("case: appended layout, heavy duplicate build key")
    nb = 20000
    exp = _both(HEADLINE, _tuples(np.full(nb, 77)), _tuples(np.array([77] * 10 + [5, 6] * 500)), nb, 1010,
                radix_bits=3)
    assert exp.shape[0] == nb * 10


def test_row_id_form():
    # two LONG key columns take the row-id form; its records carry the
    # probe row, which the listed chunks keep
    _say("case: appended layout, row-id form")
    rng = np.random.default_rng(43)
    nb, npr = 20000, 200000
    ba = rng.integers(0, 3000, nb)
    bb = rng.integers(0, 3, nb)
    pi = rng.integers(0, nb, npr)
    bt = np.stack([ba, bb, np.arange(nb)], 1).astype(np.int64).reshape(-1).view(np.uint8)
    pt = np.stack([ba[pi], bb[pi] + (np.arange(npr) % 7 == 0), np.arange(npr)], 1).astype(np.int64)
    args = dict(build_cols=["long", "long", "long"], probe_cols=["long", "long", "long"], buildjattr=[0, 1],
                probejattr=[0, 1], projection=["B$0", "B$1", "B$2", "P$2"])
    exp = _both(args, bt, pt.reshape(-1).view(np.uint8), nb, npr, radix_bits=5)
    assert exp.shape[0] > npr // 2


def test_empty_sides():
    _say("case: appended layout, empty sides")
    b, p = O.gen_pkfk(5, 1 << 40, 1000, 5000, 5000)
    e = np.zeros(0, np.uint8)
    for flags in (APPEND, EXACT):
        assert _run(HEADLINE, e, p.view(np.uint8), 0, 5000, 5000, flags=flags).shape[0] == 0
        assert _run(HEADLINE, b.view(np.uint8), e, 1000, 0, 1, flags=flags).shape[0] == 0
        assert _run(HEADLINE, b.view(np.uint8), p.view(np.uint8), 1000, 5000, 5000, flags=flags).shape[0] == 0


def test_handle_reused_across_sizes():
    # one handle, probe sides of shrinking and growing sizes in turn: the
    # appended layout's buffers and chunk list are rebuilt every run
    _say("case: appended layout, one handle over several runs")
    j = P.Join(P.make_desc(**HEADLINE, flags=APPEND, radix_bits=7))
    d = O.join_desc(["long", "long"], ["long", "long"], [0], [0], ["B$0", "B$1", "P$1"], buckets=1 << 12)
    for i, (nb, npr) in enumerate([(50000, 400000), (3000, 9000), (80000, 700000), (1, 100)]):
        b, p = O.gen_pkfk(300 + i, 1 << 40, nb, npr, npr // 5)
        exp, _ = O.hashjoin(d, b.view(np.uint8), p.view(np.uint8))
        db, dp = to_gpu(b.view(np.uint8)), to_gpu(p.view(np.uint8))
        out = torch.empty(npr * j.out_width, dtype=torch.uint8, device="cuda")
        n = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, out.data_ptr(), npr,
                         torch.cuda.current_stream().cuda_stream)
        got = out[: n * j.out_width].cpu().numpy().reshape(-1, j.out_width)
        assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp)), (nb, npr)
    j.close()


@pytest.mark.parametrize("bits", [3, 9, 11])
def test_word_by_word_records(bits):
    """12-byte (int64 key, int32 payload) tuples, read word by word into the
    canonical records, and a probe side whose buffer is not 16-byte aligned."""
    _say("case: appended layout, 12-byte tuples, bits=%d" % bits)
    rng = np.random.default_rng(80 + bits)
    nb, npr = 60000, 500000
    bkeys = rng.choice(1 << 40, nb, replace=False)
    pkeys = np.where(rng.random(npr) < 0.8, rng.choice(bkeys, npr), rng.integers(0, 1 << 40, npr))
    bt = np.zeros(nb, dtype=[("k", "<i8"), ("v", "<i4")])
    bt["k"], bt["v"] = bkeys, np.arange(nb)
    pt = np.zeros(npr, dtype=[("k", "<i8"), ("v", "<i4")])
    pt["k"], pt["v"] = pkeys, np.arange(npr)
    args = dict(build_cols=["long", "int"], probe_cols=["long", "int"], buildjattr=[0], probejattr=[0],
                projection=["B$0", "B$1", "P$1"])
    exp = _both(args, bt.view(np.uint8), pt.view(np.uint8), nb, npr, radix_bits=bits)
    assert exp.shape[0] > npr // 2

