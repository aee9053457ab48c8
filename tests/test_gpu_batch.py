"""GPU tests of the batched joins (pj_batch_*): every join of a batch
against the generator's closed form (count and 64-bit checksum) or the
CPU oracle's HashJoinOp (sorted rows), in the modes that run each join as a
whole single join (one after another, on four streams, or those streams
captured into one graph). The batched pipeline's tests (its kernels
launched one by one, or captured into one graph) are skipped while that
pipeline is switched off.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle as O
import pjoin as P

pytestmark = pytest.mark.gpu

PIPELINE_OFF = pytest.mark.skip(reason="the batched pipeline is switched off: its first device run (a three-join "
                                       "batch) did not finish and the cause has not been found")

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    P.lib()
    yield


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


DESC = dict(build_cols=["long", "int"], probe_cols=["long", "int"], buildjattr=[0], probejattr=[0],
            projection=["B$0", "B$1", "P$1"])


def _make(seed, sizes, probe_factor=4):
    """Device buffers of a generated batch; returns (joins, keepalive)."""
    joins, keep = [], []
    for j, nb in enumerate(sizes):
        nb = int(nb)
        npr = probe_factor * nb
        b = torch.empty(max(nb, 1) * 12, dtype=torch.uint8, device="cuda")
        p = torch.empty(max(npr, 1) * 12, dtype=torch.uint8, device="cuda")
        o = torch.empty(max(npr, 1) * 16, dtype=torch.uint8, device="cuda")
        P.check(P.lib().pj_gen_batch_join(seed, j, nb, npr, C.c_void_p(b.data_ptr()), C.c_void_p(p.data_ptr()),
                                          1, None))
        joins.append((b.data_ptr(), nb, p.data_ptr(), npr, o.data_ptr(), max(npr, 1)))
        keep += [b, p, o]
    torch.cuda.synchronize()
    return joins, keep


def _check(seed, joins, cnt):
    for j, (b, nb, p, npr, o, cap) in enumerate(joins):
        es, ec = O.expected_batch_join(seed, j, nb, npr, threads=16)
        s, _ = P.checksum_device(o, int(cnt[j]), 16)
        assert (int(cnt[j]), s) == (ec, es), "join %d (%d x %d)" % (j, nb, npr)


@PIPELINE_OFF
def test_batch_pipeline_three_joins():
    _say("case: pipeline, three joins")
    seed = 20
    joins, keep = _make(seed, [5000, 300, 70000])
    bt = P.Batch(P.make_desc(**DESC))
    bt.prepare(joins)
    for mode in (P.Batch.PLAIN, P.Batch.GRAPH):
        for j in range(len(joins)):
            keep[3 * j + 2].fill_(0xAB)
        bt.run(mode)
        cnt, done = bt.results()
        _check(seed, joins, cnt)
        assert np.all(done > 0)
    bt.close()


@PIPELINE_OFF
def test_batch_small_joins_graph_and_plain():
    _say("case: batch of 40 joins, graph replays and plain launches")
    seed = 21
    sizes = list(P.batch_sizes(seed, 38, 1024, 1 << 16)) + [1, 3]
    joins, keep = _make(seed, sizes)
    bt = P.Batch(P.make_desc(**DESC))
    bt.prepare(joins)
    prev = None
    for rep in range(3):
        bt.run(P.Batch.GRAPH)
        torch.cuda.synchronize()
        x = torch.ones(1 << 20, device="cuda").sum().item()   # eager work between replays
        assert x == float(1 << 20)
        cnt, done = bt.results()
        _check(seed, joins, cnt)
        assert np.all(done > 0)
        if prev is not None:
            assert np.array_equal(cnt, prev)
        prev = cnt
    bt.run(P.Batch.PLAIN)
    cnt, _ = bt.results()
    _check(seed, joins, cnt)
    bt.close()


@pytest.mark.parametrize("bits,lf,piece", [(4, 0.0, 0), (8, 0.0, 32768), (12, 0.0, 0), (0, 0.5, 0),
                                           (0, 0.9, 262144), (9, 0.0, 4096), (1, 0.0, 0)])
@PIPELINE_OFF
def test_batch_forced_configs(bits, lf, piece):
    _say("case: pipeline config bits=%d lf=%.1f piece=%d" % (bits, lf, piece))
    seed = 22
    sizes = P.batch_sizes(seed, 12, 1024, 1 << 20)
    joins, keep = _make(seed, sizes)
    bt = P.Batch(P.make_desc(**DESC))
    for c in range(32):
        bt.set_config(c, bits, 1, lf, piece=piece)
    bt.prepare(joins)
    bt.run(P.Batch.GRAPH)
    cnt, _ = bt.results()
    _check(seed, joins, cnt)
    bt.close()


MODES = pytest.mark.parametrize("mode", [P.Batch.SEQUENTIAL, P.Batch.CONCURRENT], ids=["sequential", "concurrent"])
ALL_MODES = pytest.mark.parametrize("mode", [pytest.param(P.Batch.PLAIN, marks=PIPELINE_OFF),
                                             pytest.param(P.Batch.GRAPH, marks=PIPELINE_OFF),
                                             P.Batch.SEQUENTIAL, P.Batch.CONCURRENT],
                                    ids=["plain", "graph", "sequential", "concurrent"])


@MODES
def test_batch_sequential_small_joins_repeat(mode):
    _say("case: batch of 40 joins, mode %d, two runs" % mode)
    seed = 21
    sizes = list(P.batch_sizes(seed, 38, 1024, 1 << 16)) + [1, 3]
    joins, keep = _make(seed, sizes)
    bt = P.Batch(P.make_desc(**DESC))
    bt.prepare(joins)
    prev = None
    for rep in range(2):
        bt.run(mode)
        cnt, done = bt.results()
        _check(seed, joins, cnt)
        own = bt.run_times()
        assert np.all(done > 0) and np.all(own > 0) and np.all(own <= done + 1e-3)
        if mode == P.Batch.SEQUENTIAL:
            assert np.all(np.diff(done) > 0)
        if prev is not None:
            assert np.array_equal(cnt, prev)
        prev = cnt
    bt.close()


def test_batch_concurrent_graph_replays():
    """The concurrent mode captured into one graph: three replays with eager
    work in between (the outputs are overwritten before each replay), every
    replay against the closed form; then an eager run and a fresh capture."""
    _say("case: batch graph replays")
    seed = 27
    sizes = list(P.batch_sizes(seed, 37, 1024, 1 << 16)) + [1, 3, 0]
    joins, keep = _make(seed, sizes)
    bt = P.Batch(P.make_desc(**DESC))
    bt.prepare(joins)
    for rep in range(3):
        for j in range(len(joins)):
            keep[3 * j + 2].fill_(0xAB)
        x = torch.ones(1 << 20, device="cuda").sum().item()   # eager work between replays
        assert x == float(1 << 20)
        torch.cuda.synchronize()
        bt.run(P.Batch.CONCURRENT_GRAPH)
        cnt, done = bt.results()
        _check(seed, joins, cnt)
        assert np.all(done > 0)
    bt.run(P.Batch.CONCURRENT)
    cnt, _ = bt.results()
    _check(seed, joins, cnt)
    for j in range(len(joins)):
        keep[3 * j + 2].fill_(0xCD)
    bt.run(P.Batch.CONCURRENT_GRAPH)
    cnt, _ = bt.results()
    _check(seed, joins, cnt)
    bt.close()


@ALL_MODES
@pytest.mark.parametrize("bits", [4, 9, 12])
def test_batch_sequential_forced_bits(bits, mode):
    _say("case: batch mode %d, bits=%d" % (mode, bits))
    seed = 22
    sizes = P.batch_sizes(seed, 12, 1024, 1 << 20)
    joins, keep = _make(seed, sizes)
    bt = P.Batch(P.make_desc(**DESC))
    for c in range(32):
        bt.set_config(c, bits, 1, 0.7)
    bt.prepare(joins)
    bt.run(mode)
    cnt, _ = bt.results()
    _check(seed, joins, cnt)
    bt.close()


@ALL_MODES
def test_batch_sequential_empty_sides(mode):
    """Joins with no build tuples, no probe tuples or neither, between normal
    ones: zero rows for the empty ones, the closed form for the others."""
    _say("case: batch sequential, empty sides")
    seed = 26
    joins, keep = _make(seed, [3000, 0, 700, 5000])
    b0, nb0, p0, np0, o0, c0 = joins[0]
    b2, nb2, p2, np2, o2, c2 = joins[2]
    joins[1] = (b2, 0, p0, np0, joins[1][4], 1)      # probe tuples, no build tuples
    joins[2] = (b2, nb2, p2, 0, o2, c2)                 # build tuples, no probe tuples
    joins.append((b0, 0, p0, 0, o0, 1))                 # neither
    bt = P.Batch(P.make_desc(**DESC))
    bt.prepare(joins)
    bt.run(mode)
    cnt, _ = bt.results()
    assert [int(x) for x in cnt[[1, 2, 4]]] == [0, 0, 0]
    for j in (0, 3):
        b, nb, p, npr, o, cap = joins[j]
        es, ec = O.expected_batch_join(seed, j, nb, npr, threads=16)
        s_, _ = P.checksum_device(o, int(cnt[j]), 16)
        assert (int(cnt[j]), s_) == (ec, es), j
    bt.close()


@ALL_MODES
def test_batch_tuner_choices_and_results(mode):
    _say("case: batch tuner")
    seed = 23
    sizes = P.batch_sizes(seed, 24, 1024, 1 << 20)
    joins, keep = _make(seed, sizes)
    bt = P.Batch(P.make_desc(**DESC))
    bt.tune(joins, trial=2)
    for nb in sizes:
        c = int(np.floor(np.log2(int(nb))))
        bits, passes, lf, piece = bt.config(c)
        assert 4 <= bits <= 12 and passes == 1 and piece in (0, 32768, 262144)
    bt.prepare(joins)
    bt.run(mode)
    cnt, _ = bt.results()
    _check(seed, joins, cnt)
    bt.close()


@ALL_MODES
def test_batch_repeated_keys_match_oracle(mode):
    """Joins with repeated build keys (many-to-many) against HashJoinOp."""
    _say("case: batch, repeated keys")
    rng = np.random.default_rng(24)
    desc = O.join_desc(["long", "int"], ["long", "int"], [0], [0], ["B$0", "B$1", "P$1"], buckets=1 << 10)
    joins, keep, host = [], [], []
    for j, (nb, reps) in enumerate([(3000, 3), (50000, 2), (700, 7), (20000, 1)]):
        keys = np.repeat(rng.integers(-2 ** 40, 2 ** 40, nb // reps), reps)
        rng.shuffle(keys)
        build = np.zeros(keys.size, dtype=[("k", "<i8"), ("p", "<i4")])
        build["k"] = keys
        build["p"] = rng.integers(-2 ** 31, 2 ** 31 - 1, keys.size)
        probe = np.zeros(4 * keys.size, dtype=[("k", "<i8"), ("p", "<i4")])
        probe["k"] = np.concatenate([rng.choice(keys, 3 * keys.size), rng.integers(-2 ** 40, 2 ** 40, keys.size)])
        probe["p"] = np.arange(probe.size)
        exp, _ = O.hashjoin(desc, build.view(np.uint8), probe.view(np.uint8))
        b = torch.from_numpy(build.view(np.uint8).copy()).cuda()
        p = torch.from_numpy(probe.view(np.uint8).copy()).cuda()
        o = torch.empty(max(exp.shape[0], 1) * 16, dtype=torch.uint8, device="cuda")
        joins.append((b.data_ptr(), keys.size, p.data_ptr(), probe.size, o.data_ptr(), max(exp.shape[0], 1)))
        keep += [b, p, o]
        host.append(exp)
    bt = P.Batch(P.make_desc(**DESC))
    bt.prepare(joins)
    bt.run(mode)
    cnt, _ = bt.results()
    for j, exp in enumerate(host):
        got = keep[3 * j + 2][: int(cnt[j]) * 16].cpu().numpy().reshape(-1, 16)
        assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp)), j
    bt.close()


@ALL_MODES
def test_batch_output_overflow_reported(mode):
    _say("case: batch, output too small")
    seed = 25
    joins, keep = _make(seed, [5000, 6000])
    b, nb, p, npr, o, cap = joins[1]
    joins[1] = (b, nb, p, npr, o, 100)
    bt = P.Batch(P.make_desc(**DESC))
    bt.prepare(joins)
    bt.run(mode)
    with pytest.raises(P.PJError) as e:
        bt.results()
    assert e.value.code == P.PJ_ERR_CAPACITY
    bt.close()


def test_full_batch_512_joins_tuned_graph():
    """The full 512-join batch (build sizes log-uniform in [1K, 4M], seed
    505, probe 4x build): tuned radix bits, the four streams captured into
    one graph and replayed twice, then launched directly; every join of every
    run against the generator's closed form (count and checksum)."""
    _say("case: full batch of 512 joins")
    seed = 505
    sizes = [int(x) for x in P.batch_sizes(seed, 512, 1 << 10, 1 << 22)]
    joins, keep = _make(seed, sizes)
    _say("generated %d joins, %d build tuples" % (len(sizes), sum(sizes)))
    bt = P.Batch(P.make_desc(**DESC))
    bt.tune(joins, trial=4)
    bt.prepare(joins)
    for mode in (P.Batch.CONCURRENT_GRAPH, P.Batch.CONCURRENT_GRAPH, P.Batch.CONCURRENT):
        for j in range(len(joins)):
            keep[3 * j + 2].fill_(0xAB)
        torch.cuda.synchronize()
        bt.run(mode)
        cnt, _ = bt.results()
        _check(seed, joins, cnt)
        _say("mode %d: 512 joins match, %d output tuples" % (mode, int(cnt.sum())))
    bt.close()
    del keep


def test_batch_of_16_byte_joins_large_and_small_probes_in_one_lane():
    """(long, long) tuples, lanes holding a probe side of 4M tuples and a
    smaller one at the same radix bits: every mode, the graph one included,
    gives each join its closed-form count and the same rows."""
    _say("case: batch of 16-byte joins, large and small probe sides")
    sizes = [1 << 20, 1 << 18, 70000, 3000, 600000, 1 << 17, 50000, 2000]
    keep, joins = [], []
    for j, nb in enumerate(sizes):
        npr = 4 * nb
        nmiss = npr // 8
        b = torch.empty(nb * 16, dtype=torch.uint8, device="cuda")
        p = torch.empty(npr * 16, dtype=torch.uint8, device="cuda")
        o = torch.empty(npr * 24, dtype=torch.uint8, device="cuda")
        P.check(P.lib().pj_gen_pkfk(700 + j, 1 << 40, nb, npr, nmiss, C.c_void_p(b.data_ptr()),
                                    C.c_void_p(p.data_ptr()), 1, None))
        joins.append((b.data_ptr(), nb, p.data_ptr(), npr, o.data_ptr(), npr))
        keep += [b, p, o]
    torch.cuda.synchronize()
    bt = P.Batch(P.make_desc(["long", "long"], ["long", "long"], [0], [0], ["B$0", "B$1", "P$1"]))
    bt.prepare(joins)
    sums = None
    for mode in (P.Batch.SEQUENTIAL, P.Batch.CONCURRENT, P.Batch.CONCURRENT_GRAPH, P.Batch.CONCURRENT_GRAPH):
        for j in range(len(joins)):
            keep[3 * j + 2].fill_(0xAB)
        bt.run(mode)
        cnt, _ = bt.results()
        got = []
        for j, (_, nb, _, npr, o, _) in enumerate(joins):
            assert int(cnt[j]) == npr - npr // 8, (mode, j)
            got.append(P.checksum_device(o, int(cnt[j]), 24))
        if sums is None:
            sums = got
        assert got == sums, mode
    bt.close()

