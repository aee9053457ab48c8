"""GPU tests of the learned partitioner: the sample sort, the model's root,
least-squares fit and training gradient against the float64 oracle
(oracle/learned.py), the balance the trained model buys over the method
without it, and learned-partition joins against the CPU oracle, fallbacks
included.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import learned as OL
import oracle as O
import pjoin as P
from _util import to_gpu

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


def _dev_u64(a):
    t = torch.from_numpy(np.ascontiguousarray(a, np.uint64).view(np.uint8)).cuda()
    return t


def _clustered_keys(seed, n, **kw):
    """Build keys (int64) of the clustered workload, generated on the host."""
    t = P.cluster_table(seed, n, **kw)
    cdf = O.zipf_cdf(n, 1.0)
    build = np.zeros(n * 16, np.uint8)
    probe = np.zeros(16 * 16, np.uint8)
    P.check(P.lib().pj_gen_clustered(t.ctypes.data_as(C.c_void_p), 16, cdf.ctypes.data_as(C.c_void_p),
                                     build.ctypes.data_as(C.c_void_p), probe.ctypes.data_as(C.c_void_p), 0, None))
    return build.view(np.int64).reshape(-1, 2)[:, 0].copy()


@pytest.mark.parametrize("n", [0, 1, 2, 100, 4096, 4097, 200003, 1 << 21])
def test_sort_u64(n):
    _say("case: sort %d keys" % n)
    rng = np.random.default_rng(n + 1)
    k = rng.integers(0, np.iinfo(np.uint64).max, size=n, dtype=np.uint64, endpoint=True)
    if n > 10:
        k[: n // 3] = k[: n // 3] % np.uint64(97)   # repeats
    d = _dev_u64(k) if n else torch.zeros(8, dtype=torch.uint8, device="cuda")
    P.sort_u64(d.data_ptr(), n)
    got = d.cpu().numpy().view(np.uint64)[:n]
    assert np.array_equal(got, np.sort(k))


@pytest.mark.parametrize("kind", ["low24", "low32", "high_fixed", "low_fixed", "one_digit", "all_equal",
                                  "odd_passes"])
def test_sort_u64_skips_constant_digits(kind):
    """Digits that are equal in every key are not sorted on; the keys still
    come back sorted, also when the passes left are odd in number (their
    result then lies in the ping-pong buffer) or none."""
    _say("case: sort with constant digits, %s" % kind)
    rng = np.random.default_rng(len(kind))
    n = 300001
    k = rng.integers(0, 1 << 24, size=n, dtype=np.uint64)                        # 3 varying digits
    if kind == "low32":
        k = rng.integers(0, 1 << 32, size=n, dtype=np.uint64)                    # 4 (an even number)
    elif kind == "high_fixed":
        k = k | np.uint64(0xABCD_0000_0000_0000)
    elif kind == "low_fixed":
        k = (k << np.uint64(40)) | np.uint64(0x12_3456_789A)                  # 3 varying digits on top
    elif kind == "one_digit":
        k = (k & np.uint64(0xFF)) << np.uint64(32)
    elif kind == "all_equal":
        k = np.full(n, 0x0123_4567_89AB_CDEF, dtype=np.uint64)
    elif kind == "odd_passes":
        k = (k & np.uint64(0xFFFF)) | ((k & np.uint64(0xFF0000)) << np.uint64(32))
    d = _dev_u64(k)
    P.sort_u64(d.data_ptr(), n)
    got = d.cpu().numpy().view(np.uint64)[:n]
    assert np.array_equal(got, np.sort(k))


def _fit(keys, iters):
    u = OL.ordered(keys)
    s = u[::100].copy()
    d = _dev_u64(s)
    m = P.model_fit(d.data_ptr(), s.size, u.size, iters)
    return m, np.sort(s), u


@pytest.fixture(scope="module")
def clustered_2m():
    return _clustered_keys(45, 2_000_000, nclusters=64, log2_wmin=10, log2_wmax=18)


def test_model_root_and_fit_match_oracle(clustered_2m):
    _say("case: model root and least-squares fit")
    m, s, _ = _fit(clustered_2m, 0)
    M = OL.view(m)
    R = OL.root(s)
    assert M["S"] == R["S"] and M["L"] == R["L"]
    for k in ("seg_lo", "seg_hi", "seg_lw", "seg_leaf0"):
        assert np.array_equal(M[k], R[k]), k
    assert np.array_equal(M["leaf_seg"], R["leaf_seg"])
    R["n_keys"] = M["n_keys"]
    assert np.array_equal(M["leaf_start"], OL.leaf_starts(R, s))
    R["leaf_start"] = M["leaf_start"]
    a, b = OL.lsq(R, s)
    # FP32 moments against float64 lstsq: ranks within a window are < 2^13
    assert np.abs(M["b"] - b).max() <= 1e-3 * max(1.0, np.abs(b).max())
    assert np.abs(M["a"] - a).max() <= 0.05


def test_model_gradient_matches_oracle(clustered_2m):
    _say("case: model gradient")
    m, s, _ = _fit(clustered_2m, 0)
    rng = np.random.default_rng(7)
    L = m.nleaves
    for j in range(L):   # move off the optimum so every term contributes
        m.a[j] = m.a[j] + float(rng.normal(0, 2))
        m.b[j] = max(0.0, m.b[j] * float(1 + rng.normal(0, 0.05)))
    d = _dev_u64(s)
    loss, g = P.model_grad(m, d.data_ptr(), s.size)
    M = OL.view(m)
    ga, gb = OL.grad(M, s, M["a"], M["b"])
    want = np.concatenate([ga, gb])
    assert loss == pytest.approx(OL.loss(M, s, M["a"], M["b"]), rel=1e-4)
    assert np.abs(g - want).max() <= 2e-3 * np.abs(want).max() + 1e-7


def test_training_lowers_the_loss(clustered_2m):
    _say("case: training steps")
    m0, s, _ = _fit(clustered_2m, 0)
    m, _, _ = _fit(clustered_2m, 24)
    assert m.iters == 24 and m0.iters == 0
    assert m.loss <= m.loss_lsq * (1 + 1e-5)
    assert m.loss_lsq == pytest.approx(m0.loss, rel=1e-5)
    M0, M = OL.view(m0), OL.view(m)
    assert np.abs(OL.jumps(M, M["a"], M["b"])).sum() < np.abs(OL.jumps(M0, M0["a"], M0["b"])).sum()


def _lds_cost(bnd, u):
    """LDS cost of every build partition of ordered keys u: count / capacity
    of its table format (dense when the partition holds a whole key range),
    and whether a code table can hold it at all (key span < 2^32)."""
    p = OL.partition_of(bnd, u)
    order = np.argsort(p, kind="stable")
    ps, us = p[order], u[order]
    F = bnd.size + 1
    cnt = np.bincount(ps, minlength=F)
    lo = np.full(F, np.iinfo(np.uint64).max, np.uint64)
    hi = np.zeros(F, np.uint64)
    np.minimum.at(lo, ps, us)
    np.maximum.at(hi, ps, us)
    span = np.where(cnt > 0, hi - lo, np.uint64(0))
    dense = (cnt > 0) & ((span + np.uint64(1)).astype(np.float64) == cnt)
    return np.where(dense, cnt / OL.DENSE_CAP, cnt / OL.CODE_CAP), span >= np.uint64(1 << 32)


def test_trained_model_balances_better_than_quantiles():
    """What the trained network gains over the method without it. Boundaries
    at quantiles of the same sample (same partition count) cut through the
    gaps between clusters, so many build partitions span 2^32 key values
    or more, which no table format holds; the model's boundaries leave none,
    and their fill stays close to the plan (a partition over 1.0 is split
    in two at join time)."""
    _say("case: balance, model vs quantiles")
    keys = _clustered_keys(404, 8_000_000, nclusters=256, log2_wmin=10, log2_wmax=24)
    m, s, u = _fit(keys, 16)
    bnd = P.model_bounds(m)
    cm, wm = _lds_cost(bnd, u)
    q = OL.quantile_bounds(s, bnd.size + 1)
    cq, wq = _lds_cost(q, u)
    _say("max LDS fill: model %.3f (%d over 1), quantiles %.3f; partitions spanning 2^32 keys: model %d, "
         "quantiles %d (%d partitions)" % (cm.max(), (cm > 1).sum(), cq.max(), wm.sum(), wq.sum(), bnd.size + 1))
    assert wm.sum() == 0 and wq.sum() > 0.05 * cq.size
    assert cm.max() <= 1.15 and (cm > 1).sum() <= 0.01 * cm.size


# ---------------------------------------------------------------- joins

HEADLINE = dict(build_cols=["long", "long"], probe_cols=["long", "long"], buildjattr=[0], probejattr=[0],
                projection=["B$0", "B$1", "P$1"])


def _join(args, build, probe, nb, npr, cap, **kw):
    j = P.Join(P.make_desc(**args, **kw, timing=True))
    db = to_gpu(build)
    dp = to_gpu(probe)
    out = torch.empty(max(cap, 1) * j.out_width, dtype=torch.uint8, device="cuda")
    n = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, out.data_ptr(), cap)
    rows = out[: n * j.out_width].cpu().numpy().reshape(-1, j.out_width)
    t = j.timings()
    j.close()
    return rows, t


def _compare(args, build, probe, nb, npr, **kw):
    desc = O.join_desc(args["build_cols"], args["probe_cols"], args["buildjattr"], args["probejattr"],
                       args["projection"], buckets=1 << 12)
    exp, _ = O.hashjoin(desc, build, probe)
    got, t = _join(args, build, probe, nb, npr, max(exp.shape[0], 1), **kw)
    assert got.shape == exp.shape, (got.shape, exp.shape)
    assert np.array_equal(O.sorted_rows(got), O.sorted_rows(exp))
    return t


def _kv(keys, pay):
    return np.stack([np.asarray(keys, np.int64), np.asarray(pay, np.int64)], 1).reshape(-1).view(np.uint8)


def test_learned_join_dense_keys():
    _say("case: learned join, dense keys")
    hb, hp = O.gen_pkfk(11, 400_000, 400_000, 1_600_000, 0)
    t = _compare(HEADLINE, hb.view(np.uint8), hp.view(np.uint8), 400_000, 1_600_000, partitioner="learned")
    assert t.used_learned == 1


def test_learned_join_clusters_with_misses():
    _say("case: learned join, clustered keys and misses")
    rng = np.random.default_rng(12)
    keys = _clustered_keys(46, 500_000, nclusters=32, log2_wmin=12, log2_wmax=18)
    pay = rng.integers(-2 ** 62, 2 ** 62, keys.size)
    fk = np.concatenate([rng.choice(keys, 1_500_000), rng.integers(-2 ** 63, 2 ** 63 - 1, 300_000),
                         rng.choice(keys, 100_000) + 1, np.array([-2 ** 63, 2 ** 63 - 1])])
    probe = _kv(fk, rng.integers(-2 ** 62, 2 ** 62, fk.size))
    t = _compare(HEADLINE, _kv(keys, pay), probe, keys.size, fk.size, partitioner="learned")
    assert t.used_learned == 1


def test_learned_falls_back_on_fragmented_keys():
    _say("case: learned join, keys too spread for the root")
    rng = np.random.default_rng(13)
    keys = rng.permutation(np.unique(rng.integers(-2 ** 63, 2 ** 63 - 1, 300_000)))
    fk = rng.choice(keys, 900_000)
    t = _compare(HEADLINE, _kv(keys, rng.integers(0, 1 << 40, keys.size)), _kv(fk, np.arange(fk.size)),
                 keys.size, fk.size, partitioner="learned")
    assert t.used_learned == 0


def test_learned_join_repeated_keys():
    _say("case: learned join, many-to-many")
    rng = np.random.default_rng(14)
    keys = rng.permutation(np.repeat(np.arange(100_000, dtype=np.int64) * 3, 3))
    fk = rng.choice(np.arange(300_000, dtype=np.int64), 600_000)
    t = _compare(HEADLINE, _kv(keys, rng.integers(0, 1 << 40, keys.size)), _kv(fk, np.arange(fk.size)),
                 keys.size, fk.size, partitioner="learned")
    assert t.used_learned == 1


@pytest.mark.parametrize("n", [20_000, 200_000])
def test_learned_dense_looking_partition_with_repeats(n):
    """Every 997th key is replaced by its successor: partitions look dense
    (count == key span) but hold a repeated key. Small partitions are
    rebuilt as code tables in the kernel; large ones send the whole join
    back to radix partitions."""
    _say("case: learned join, dense-looking partitions with repeats, n=%d" % n)
    rng = np.random.default_rng(n)
    k = np.arange(n, dtype=np.int64)
    k[5::997] += 1
    keys = rng.permutation(k)
    fk = rng.choice(np.arange(n + 2, dtype=np.int64), 4 * n)
    _compare(HEADLINE, _kv(keys, rng.integers(0, 1 << 40, n)), _kv(fk, np.arange(fk.size)), n, fk.size,
             partitioner="learned")


def test_learned_join_composite_keys():
    _say("case: learned join, composite keys")
    n, m = 200_000, 800_000
    cdf = O.zipf_cdf(n, 1.0)
    b, p = O.gen_pairs_zipf(77, n, m, cdf)
    args = dict(build_cols=["int", "int", "long"], probe_cols=["int", "int", "long"], buildjattr=[0, 1],
                probejattr=[0, 1], projection=["B$0", "B$1", "B$2", "P$2"])
    _compare(args, b.view(np.uint8), p.view(np.uint8), n, m, partitioner="learned")


def test_learned_join_int_keys_and_projection():
    _say("case: learned join, INTEGER keys, reordered projection")
    rng = np.random.default_rng(15)
    n, m = 150_000, 600_000
    bk = rng.permutation(np.arange(-n // 2, n - n // 2, dtype=np.int32))
    build = np.zeros((n, 2), np.int32)
    build[:, 0] = bk
    build[:, 1] = rng.integers(-2 ** 31, 2 ** 31 - 1, n)
    probe = np.zeros((m, 4), np.int32)
    probe[:, 1] = rng.choice(bk, m)
    probe[:, 0] = rng.integers(-2 ** 31, 2 ** 31 - 1, m)
    args = dict(build_cols=["int", "int"], probe_cols=["int", "int", "long"], buildjattr=[0], probejattr=[1],
                projection=["P$0", "B$1", "B$0"])
    _compare(args, build.view(np.uint8).reshape(-1), probe.view(np.uint8).reshape(-1), n, m,
             partitioner="learned")


def test_auto_picks_radix_for_small_build():
    _say("case: auto partitioner, small build side")
    hb, hp = O.gen_pkfk(16, 1 << 40, 100_000, 400_000, 40_000)
    t = _compare(HEADLINE, hb.view(np.uint8), hp.view(np.uint8), 100_000, 400_000, partitioner="auto")
    assert t.used_learned == 0
