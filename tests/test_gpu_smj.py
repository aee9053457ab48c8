"""GPU parity of the device sort-merge join (pj_smj_*, the counterpart of
the reference's SortMergeJoinOp / MPSMJoinOp, operators/join.cpp:905-1791)
against the oracle's HashJoinOp: the same output multiset, bit for bit, at
small sizes, and count + checksum against the generator's closed form at
16M x 256M."""
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


HEADLINE = dict(build_cols=["long", "long"], probe_cols=["long", "long"], buildjattr=[0], probejattr=[0],
                projection=["B$0", "B$1", "P$1"])


def run_smj(args, build, probe, nb, npr, cap):
    j = P.SortMergeJoin(P.make_desc(**args))
    db = to_gpu(build) if build.size else torch.zeros(16, dtype=torch.uint8, device="cuda")
    dp = to_gpu(probe) if probe.size else torch.zeros(16, dtype=torch.uint8, device="cuda")
    out = torch.empty(max(cap, 1) * j.out_width, dtype=torch.uint8, device="cuda")
    n = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, out.data_ptr(), cap,
                     torch.cuda.current_stream().cuda_stream)
    rows = out[: n * j.out_width].cpu().numpy().reshape(-1, j.out_width)
    t = j.timings()
    j.close()
    return rows, t


def compare(args, build, probe, nb, npr, out_cap=None):
    d = O.join_desc(args["build_cols"], args["probe_cols"], args["buildjattr"], args["probejattr"],
                    args["projection"], buckets=1 << 12)
    exp, _ = O.hashjoin(d, build, probe, out_cap=out_cap)
    got, t = run_smj(args, build, probe, nb, npr, cap=max(exp.shape[0], 1))
    assert got.shape == exp.shape, (got.shape, exp.shape)
    assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))
    return got, t


@pytest.mark.parametrize("nb,npr", [(1, 1), (3, 1000), (1 << 16, 1 << 20), (100000, 300000), (1 << 20, 1 << 18)])
def test_pkfk_vs_oracle(nb, npr):
    _say("case: smj pkfk nb=%d np=%d" % (nb, npr))
    b, p = O.gen_pkfk(nb + npr + 7, 1 << 40, nb, npr, npr // 10)
    got, t = compare(HEADLINE, b.view(np.uint8), p.view(np.uint8), nb, npr)
    assert got.shape[0] == npr - npr // 10
    assert t.build_key_bits <= 40 and t.probe_key_bits <= 40


def test_many_to_many_signed_cross_width_keys():
    """Repeated keys on both sides, negative keys, an INTEGER build key
    against a LONG probe key (both sign-extended to 64 bits)."""
    _say("case: smj many-to-many int/long keys")
    rng = np.random.default_rng(41)
    nb, npr = 3000, 20000
    bk = rng.integers(-300, 300, nb)
    pk = rng.integers(-300, 300, npr)
    build = pack(["int", "long", "int"], [(int(k), int(v), int(w)) for k, v, w in
                                          zip(bk, rng.integers(0, 1 << 62, nb), rng.integers(0, 1 << 31, nb))])
    probe = pack(["long", "int"], [(int(k), int(v)) for k, v in zip(pk, rng.integers(-(1 << 31), 1 << 31, npr))])
    args = dict(build_cols=["int", "long", "int"], probe_cols=["long", "int"], buildjattr=[0], probejattr=[0],
                projection=["P$1", "B$1", "B$0", "P$0"])
    got, t = compare(args, build, probe, nb, npr)
    assert got.shape[0] > 50000
    assert t.build_key_bits == 64   # negative and positive keys: every bit sorted


def test_extreme_keys_and_one_heavy_key():
    """INT64_MIN / INT64_MAX / 0 keys, and one key 20000 times on the build
    side met by 10 probes (a run longer than any gallop step)."""
    _say("case: smj extreme and heavy keys")
    ext = np.array([-(1 << 63), (1 << 63) - 1, 0, -1, 1], np.int64)
    bkeys = np.concatenate([ext, np.full(20000, 77, np.int64)])
    pkeys = np.concatenate([ext[::-1], np.full(10, 77, np.int64), np.array([5, 6], np.int64)])
    build = np.stack([bkeys, np.arange(bkeys.size, dtype=np.int64)], 1).reshape(-1)
    probe = np.stack([pkeys, np.arange(pkeys.size, dtype=np.int64)], 1).reshape(-1)
    got, _ = compare(HEADLINE, build.view(np.uint8), probe.view(np.uint8), bkeys.size, pkeys.size,
                     out_cap=300000)
    assert got.shape[0] == 5 + 20000 * 10


def test_composite_int_keys():
    _say("case: smj composite (int, int) keys")
    rng = np.random.default_rng(42)
    n = 40000
    cols = ["int", "int", "long"]
    build = pack(cols, [(int(a), int(b), int(c)) for a, b, c in
                        zip(rng.integers(-25, 25, n), rng.integers(0, 50, n), rng.integers(0, 1 << 62, n))])
    probe = pack(cols, [(int(a), int(b), int(c)) for a, b, c in
                        zip(rng.integers(-30, 30, n), rng.integers(0, 60, n), rng.integers(0, 1 << 62, n))])
    args = dict(build_cols=cols, probe_cols=cols, buildjattr=[0, 1], probejattr=[0, 1],
                projection=["B$0", "B$1", "B$2", "P$2"])
    got, _ = compare(args, build, probe, n, n)
    assert got.shape[0] > 100000


def test_twelve_byte_tuples():
    _say("case: smj 12-byte tuples")
    rng = np.random.default_rng(43)
    nb, npr = 50000, 200000
    keys = rng.choice(1 << 40, nb, replace=False)
    build = pack(["long", "int"], [(int(k), int(v)) for k, v in zip(keys, rng.integers(0, 1 << 31, nb))])
    fk = keys[rng.integers(0, nb, npr)]
    probe = pack(["long", "int"], [(int(k), int(v)) for k, v in zip(fk, rng.integers(0, 1 << 31, npr))])
    args = dict(build_cols=["long", "int"], probe_cols=["long", "int"], buildjattr=[0], probejattr=[0],
                projection=["B$0", "B$1", "P$1"])
    got, _ = compare(args, build, probe, nb, npr)
    assert got.shape[0] == npr


def test_empty_sides_all_miss_and_capacity():
    _say("case: smj empty / all-miss / capacity")
    b, p = O.gen_pkfk(3, 1 << 40, 1000, 5000, 5000)
    e = np.zeros(0, np.uint8)
    assert run_smj(HEADLINE, e, p.view(np.uint8), 0, 5000, 5000)[0].shape[0] == 0
    assert run_smj(HEADLINE, b.view(np.uint8), e, 1000, 0, 5000)[0].shape[0] == 0
    assert run_smj(HEADLINE, b.view(np.uint8), p.view(np.uint8), 1000, 5000, 5000)[0].shape[0] == 0
    b, p = O.gen_pkfk(4, 1 << 40, 1000, 5000, 0)
    with pytest.raises(P.PJError) as ex:
        run_smj(HEADLINE, b.view(np.uint8), p.view(np.uint8), 1000, 5000, 100)
    assert ex.value.code == P.PJ_ERR_CAPACITY and "5000" in str(ex.value)


def test_full_16m_x_256m_equals_closed_form():
    _say("case: smj 16M x 256M")
    nb, npr = 1 << 24, 1 << 28
    nmiss = round(0.1 * npr)
    db = torch.empty(nb * 16, dtype=torch.uint8, device="cuda")
    dp = torch.empty(npr * 16, dtype=torch.uint8, device="cuda")
    P.check(P.lib().pj_gen_pkfk(202, 1 << 40, nb, npr, nmiss, C.c_void_p(db.data_ptr()), C.c_void_p(dp.data_ptr()),
                                1, None))
    j = P.SortMergeJoin(P.make_desc(**HEADLINE))
    out = torch.empty(npr * 24, dtype=torch.uint8, device="cuda")
    n = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, out.data_ptr(), npr)
    s, _ = P.checksum_device(out.data_ptr(), n, 24)
    j.close()
    es, ec = O.expected_pkfk(202, 1 << 40, nb, npr, nmiss, threads=16)
    assert (n, s) == (ec, es)
