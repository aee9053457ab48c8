"""Pins the learned partitioner's oracle (oracle/learned.py) on the CPU, and
checks the library's host-side boundary plan (pj_model_bounds) against it.

The model has no counterpart in the reference, so every check here rests on
mathematics: finite differences for the gradient, numpy.linalg.lstsq
optimality for the fit, gap cuts found independently for the root, and
closed forms on dense key ranges for the boundaries.
"""
import ctypes as C

import numpy as np
import pytest

import learned as OL
import pjoin as P


def _clusters(rng, spec, frac=0.01):
    """Sorted sample of ordered keys drawn from dense clusters
    spec = [(lo, count, width)] (every width-th... keys uniform in range)."""
    keys = []
    for lo, cnt, width in spec:
        off = rng.choice(width, size=cnt, replace=False)
        keys.append(np.uint64(lo) + off.astype(np.uint64))
    k = np.sort(np.concatenate(keys))
    s = k[:: int(1 / frac)] if frac < 1 else k
    return k, np.sort(s)


def _model(sample, n_keys, leaves=OL.LEAVES, fit=True):
    M = OL.root(sample, leaves=leaves)
    M["n_keys"] = n_keys
    M["leaf_start"] = OL.leaf_starts(M, sample)
    if fit:
        M["a"], M["b"] = OL.lsq(M, sample)
    return M


def test_root_cuts_at_gaps():
    rng = np.random.default_rng(1)
    spec = [(1 << 40, 3000, 10000), (1 << 50, 500, 600), ((1 << 50) + (1 << 33), 2000, 2000)]
    _, s = _clusters(rng, spec, frac=1.0)
    M = OL.root(s, leaves=64)
    assert M["S"] == 3
    for q, (lo, cnt, width) in enumerate(spec):
        assert lo <= int(M["seg_lo"][q]) <= int(M["seg_hi"][q]) < lo + width
    # leaves in proportion to counts (rounded down), at least one each
    nl = np.diff(M["seg_leaf0"])
    assert 64 - 3 <= nl.sum() <= 64 and nl.min() >= 1
    assert nl[0] > nl[2] > nl[1]
    # every sample's leaf lies in its own segment; leaves are monotone
    lf = OL.leaf_of(M, s)
    assert np.all(np.diff(lf) >= 0)
    seg = np.searchsorted(M["seg_lo"], s, side="right") - 1
    assert np.array_equal(M["leaf_seg"][lf], seg)


def test_root_too_many_segments():
    s = (np.arange(OL.MAX_SEGS + 1, dtype=np.uint64) << np.uint64(40))
    assert OL.root(s) is None


def test_lsq_exact_on_linear_cdf():
    # keys 7, 10, 13, ...: rank is exactly linear in the key in every window
    s = (np.arange(5000, dtype=np.uint64) * np.uint64(3)) + np.uint64(7)
    M = _model(s, 5000, leaves=16)
    for j in range(M["L"]):
        x, y = OL.window_xy(M, s, j)
        r = M["a"][j] + M["b"][j] * x - y
        assert np.abs(r).max() < 1e-6
    # the CDF is continuous, so the continuity term vanishes too
    assert np.abs(OL.jumps(M, M["a"], M["b"])).max() < 1e-6


def test_lsq_is_stationary_point_of_data_loss():
    rng = np.random.default_rng(2)
    _, s = _clusters(rng, [(1 << 20, 20000, 100000), (1 << 45, 5000, 5000)])
    M = _model(s, 25000 * 100, leaves=40)
    ga, gb = OL.grad(M, s, M["a"], M["b"], lam=0.0)
    # zero data-term gradient wherever the slope was not clamped at 0
    free = M["b"] > 0
    assert np.abs(ga[free]).max() < 1e-9 and np.abs(gb[free]).max() < 1e-9


def test_grad_matches_finite_differences():
    rng = np.random.default_rng(3)
    _, s = _clusters(rng, [(1 << 30, 8000, 20000), (1 << 55, 3000, 1 << 20)], frac=0.1)
    M = _model(s, 110000, leaves=24)
    a = M["a"] + rng.normal(0, 3, M["L"])
    b = M["b"] * (1 + rng.normal(0, 0.05, M["L"]))
    ga, gb = OL.grad(M, s, a, b)
    h = 1e-4
    for j in rng.choice(M["L"], 8, replace=False):
        e = np.zeros(M["L"])
        e[j] = h
        fa = (OL.loss(M, s, a + e, b) - OL.loss(M, s, a - e, b)) / (2 * h)
        fb = (OL.loss(M, s, a, b + e) - OL.loss(M, s, a, b - e)) / (2 * h)
        assert fa == pytest.approx(ga[j], rel=1e-5, abs=1e-8)
        assert fb == pytest.approx(gb[j], rel=1e-5, abs=1e-8)


def test_bounds_closed_form_dense_range():
    # dense keys 0..N-1 in one segment, model exact: every build partition
    # holds N / nparts keys (the cost per key is that of a dense table)
    N = 1 << 22
    s = np.arange(0, N, 100, dtype=np.uint64)
    M = _model(s, N, leaves=64)
    bnd, div = OL.bounds(M, max_coarse=2048)
    nparts = bnd.size + 1
    C = N / OL.DENSE_CAP
    assert nparts == int(np.ceil(C / OL.FILL)) and div == 1
    sizes = np.diff(np.concatenate([[0], bnd.astype(np.int64), [N]]))
    assert sizes.max() - sizes.min() <= 2 + N // 10 ** 6


def test_bounds_put_one_boundary_in_every_gap():
    rng = np.random.default_rng(4)
    spec = [((i + 1) << 40, 20000, 20000) for i in range(5)]
    k, s = _clusters(rng, spec)
    M = _model(s, k.size, leaves=32)
    bnd, _ = OL.bounds(M)
    part = OL.partition_of(bnd, k)
    for q, (lo, cnt, width) in enumerate(spec):
        mine = np.unique(part[(k >= np.uint64(lo)) & (k < np.uint64(lo + width))])
        others = np.unique(part[(k < np.uint64(lo)) | (k >= np.uint64(lo + width))])
        assert not set(mine) & set(others), "a build partition spans a gap"


def _to_ctypes(M):
    m = P.Model()
    m.nleaves, m.nseg, m.n, m.n_keys = M["L"], M["S"], M["n"], M["n_keys"]
    m.window, m.lam = OL.WINDOW, OL.LAMBDA
    for q in range(M["S"]):
        m.seg_lo[q], m.seg_hi[q], m.seg_lw[q] = int(M["seg_lo"][q]), int(M["seg_hi"][q]), int(M["seg_lw"][q])
    for q in range(M["S"] + 1):
        m.seg_leaf0[q] = int(M["seg_leaf0"][q])
    for j in range(M["L"]):
        m.leaf_seg[j] = int(M["leaf_seg"][j])
        m.a[j], m.b[j] = float(M["a"][j]), float(M["b"][j])
    for j in range(M["L"] + 1):
        m.leaf_start[j] = int(M["leaf_start"][j])
    return m


@pytest.mark.parametrize("seed", [5, 6])
def test_library_bounds_match_oracle(seed):
    rng = np.random.default_rng(seed)
    spec = [(int(c), int(rng.integers(2000, 60000)), 0) for c in np.sort(rng.integers(0, 1 << 62, 12))]
    spec = [(lo, cnt, cnt * int(rng.choice([1, 1, 3]))) for lo, cnt, _ in spec]
    k, s = _clusters(rng, spec)
    M = _model(s, k.size, leaves=200)
    # the library computes in float32 parameters: give both the same ones
    M["a"] = M["a"].astype(np.float32).astype(np.float64)
    M["b"] = M["b"].astype(np.float32).astype(np.float64)
    want, _ = OL.bounds(M, max_coarse=64)
    got = P.model_bounds(_to_ctypes(M), max_coarse=64)
    assert got.size == want.size
    assert np.abs(got.astype(np.int64) - want.astype(np.int64)).max() <= 1


def test_library_bounds_rejects_bad_model():
    m = P.Model()
    nb = C.c_uint32()
    assert P.lib().pj_model_bounds(C.byref(m), 64, None, 0, C.byref(nb)) == P.PJ_ERR_INVALID
