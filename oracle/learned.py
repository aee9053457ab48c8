"""CPU restatement, in float64 numpy, of the learned partitioner's model:
the root, the windowed least-squares leaf fit, the training loss and its
gradient, and the boundary plan. It follows
pythia-core_amd/csrc/pj_learned.hip (the model has no counterpart in the
reference; it replaces the partition choice of TupleHasher, hash.h:109-197,
and of PartitionOp's ranges, partition.cpp:166-200).

Checker only: tests/ use it to check the GPU kernels; the product never
imports it. Pinned by tests/test_learned_oracle.py: the gradient against
central finite differences of loss(), the leaf fit against
numpy.linalg.lstsq on each window, the root against gap cuts found with
numpy.diff, and the boundaries against closed forms on dense key ranges.
"""
import numpy as np

GAP = 1 << 32
WINDOW = 16
LAMBDA = 64.0
DENSE_CAP = 20096
CODE_CAP = 12928
FILL = 0.85
FILL_MAX = 0.92
DENSE_DENSITY = 0.97
LEAVES = 4096
LEAF_SAMPLES = 128
MAX_SEGS = 1024

U64 = np.uint64


def ordered(keys):
    """Ordered keys (sign bit flipped) of int64 keys."""
    return np.asarray(keys, np.int64).view(U64) ^ U64(1 << 63)


def view(m):
    """numpy views of a pjoin.Model (ctypes)."""
    L, S = m.nleaves, m.nseg
    return {
        "L": L, "S": S, "n": int(m.n), "n_keys": int(m.n_keys),
        "seg_lo": np.array(m.seg_lo[:S], U64), "seg_hi": np.array(m.seg_hi[:S], U64),
        "seg_lw": np.array(m.seg_lw[:S], U64), "seg_leaf0": np.array(m.seg_leaf0[:S + 1], np.int64),
        "leaf_seg": np.array(m.leaf_seg[:L], np.int64), "leaf_start": np.array(m.leaf_start[:L + 1], np.int64),
        "a": np.array(m.a[:L], np.float64), "b": np.array(m.b[:L], np.float64),
    }


# ---------------------------------------------------------------- root

def root(sorted_keys, leaves=None, gap=GAP):
    """Segments of the sorted sample (cut where neighbours are more than gap
    apart) and their leaves, in proportion to sample counts (at least one
    each, the largest trimmed until the total fits the budget: one leaf per
    LEAF_SAMPLES samples, at most LEAVES, unless given). Returns a dict like
    view() without a, b and leaf_start, or None past MAX_SEGS segments."""
    k = np.asarray(sorted_keys, U64)
    n = k.size
    if leaves is None:
        leaves = min(LEAVES, max(1, n // LEAF_SAMPLES))
    cut = np.flatnonzero(np.diff(k) > U64(gap)) + 1
    starts = np.concatenate([[0], cut]).astype(np.int64)
    ends = np.concatenate([cut, [n]]).astype(np.int64)
    S = starts.size
    if S > MAX_SEGS:
        return None
    cnt = ends - starts
    nl = np.maximum(1, np.floor(leaves * cnt.astype(np.float64) / n)).astype(np.int64)
    while nl.sum() > max(leaves, S):
        nl[np.argmax(nl)] -= 1
    lo = k[starts]
    hi = k[ends - 1]
    lw = (hi - lo) // nl.astype(U64) + U64(1)
    leaf0 = np.concatenate([[0], np.cumsum(nl)]).astype(np.int64)
    return {"L": int(leaf0[-1]), "S": S, "n": n, "seg_lo": lo, "seg_hi": hi, "seg_lw": lw, "seg_leaf0": leaf0,
            "leaf_seg": np.repeat(np.arange(S), nl)}


def leaf_of(M, u):
    """Leaf of ordered keys u."""
    u = np.asarray(u, U64)
    s = np.maximum(np.searchsorted(M["seg_lo"], u, side="right") - 1, 0)
    l0 = M["seg_leaf0"][s]
    nl = M["seg_leaf0"][s + 1] - l0
    below = u < M["seg_lo"][s]
    q = np.where(below, U64(0), (u - np.where(below, u, M["seg_lo"][s])) // M["seg_lw"][s])
    return l0 + np.minimum(q.astype(np.float64), nl - 1).astype(np.int64)


def leaf_lo(M, j):
    s = M["leaf_seg"][j]
    return M["seg_lo"][s] + (np.asarray(j, np.int64) - M["seg_leaf0"][s]).astype(U64) * M["seg_lw"][s]


def leaf_starts(M, sorted_keys):
    """First sample of each leaf (leaf_start[L] = n)."""
    lf = leaf_of(M, sorted_keys)
    return np.searchsorted(lf, np.arange(M["L"] + 1), side="left").astype(np.int64)


def leaf_x(M, j, u):
    """x of keys u in leaf j's own frame."""
    s = M["leaf_seg"][j]
    return (np.asarray(u, U64) - leaf_lo(M, j)).astype(np.float64) / float(M["seg_lw"][s])


def window(M, j):
    s = M["leaf_seg"][j]
    l0, l1 = M["seg_leaf0"][s], M["seg_leaf0"][s + 1]
    return max(l0, j - WINDOW), min(l1, j + WINDOW + 1)


def window_xy(M, keys, j):
    """(x, y) of the samples of leaf j's window in leaf j's frame."""
    st = M["leaf_start"]
    w0, w1 = window(M, j)
    xs, ys = [], []
    for m in range(w0, w1):
        i = np.arange(st[m], st[m + 1])
        if i.size == 0:
            continue
        xs.append(leaf_x(M, m, keys[i]) + (m - j))
        ys.append((i - st[j]).astype(np.float64))
    if not xs:
        return np.zeros(0), np.zeros(0)
    return np.concatenate(xs), np.concatenate(ys)


# ---------------------------------------------------------------- fit

def lsq(M, keys):
    """Least-squares line (a, b) of every leaf over its window, b clamped
    at 0 (then a = mean y)."""
    L = M["L"]
    a = np.zeros(L)
    b = np.zeros(L)
    for j in range(L):
        x, y = window_xy(M, keys, j)
        if x.size == 0:
            continue
        if x.size >= 2 and np.var(x) * x.size > 1e-12 * x.size:
            A = np.stack([np.ones_like(x), x], 1)
            (aa, bb), *_ = np.linalg.lstsq(A, y, rcond=None)
            if bb < 0:
                aa, bb = y.mean(), 0.0
            a[j], b[j] = aa, bb
        else:
            a[j] = y.mean()
    return a, b


def jumps(M, a, b):
    """delta_j at inner leaf edges (j, j+1 in one segment), else 0."""
    st = M["leaf_start"].astype(np.float64)
    L = M["L"]
    d = np.zeros(L)
    seg = M["leaf_seg"]
    for j in range(L - 1):
        if seg[j] == seg[j + 1]:
            d[j] = (st[j] + a[j] + b[j]) - (st[j + 1] + a[j + 1])
    return d


def loss(M, keys, a, b, lam=LAMBDA):
    tot = 0.0
    for j in range(M["L"]):
        x, y = window_xy(M, keys, j)
        r = a[j] + b[j] * x - y
        tot += float(np.dot(r, r))
    d = jumps(M, a, b)
    return (tot + lam * float(np.dot(d, d))) / M["n"]


def grad(M, keys, a, b, lam=LAMBDA):
    """dL/da, dL/db of loss()."""
    L = M["L"]
    n = M["n"]
    ga = np.zeros(L)
    gb = np.zeros(L)
    for j in range(L):
        x, y = window_xy(M, keys, j)
        r = a[j] + b[j] * x - y
        ga[j] = 2.0 * r.sum() / n
        gb[j] = 2.0 * np.dot(r, x) / n
    d = jumps(M, a, b)
    ga += 2.0 * lam * d / n
    gb += 2.0 * lam * d / n
    ga[1:] -= 2.0 * lam * d[:-1] / n
    return ga, gb


# ---------------------------------------------------------------- boundaries

def bounds(M, max_coarse=2048):
    """Range boundaries: equal LDS cost per build partition plus the middle
    of every gap between segments; a key's cost is that of a dense table
    when its segment's sample density reaches DENSE_DENSITY. Returns
    (bounds, div)."""
    L = M["L"]
    scale = M["n_keys"] / M["n"] if M["n"] else 0.0
    keys = np.maximum(0.0, M["b"][:L]) * scale
    seg = M["leaf_seg"][:L]
    st, l0 = M["leaf_start"], M["seg_leaf0"]
    seg_keys = (st[l0[1:]] - st[l0[:-1]]).astype(np.float64) * scale
    seg_span = (M["seg_hi"] - M["seg_lo"]).astype(np.float64) + 1.0
    cost = np.where((seg_keys / seg_span)[seg] >= DENSE_DENSITY, 1.0 / DENSE_CAP, 1.0 / CODE_CAP)
    pre = np.concatenate([[0.0], np.cumsum(cost * keys)])
    C = pre[L]
    gaps = M["S"] - 1
    nmin = int(max(1.0, np.ceil(C / FILL_MAX)))
    div = max(1, (nmin + gaps + max_coarse - 1) // max_coarse)
    room = div * max_coarse - gaps if div * max_coarse > gaps else 1
    ncost = max(nmin, int(np.ceil(C / FILL)))
    if div > 1:
        ncost = max(ncost, room)
    ncost = max(nmin, min(ncost, room))
    out = []
    j = 0
    for t in range(1, ncost):
        T = C * t / ncost
        while j + 1 < L and pre[j + 1] <= T:
            j += 1
        mass = pre[j + 1] - pre[j]
        frac = (T - pre[j]) / mass if mass > 0 else 0.0
        frac = min(1.0, max(0.0, frac))
        w = int(M["seg_lw"][M["leaf_seg"][j]])
        out.append(int(leaf_lo(M, j)) + int(np.ceil(frac * w)))
    for s in range(M["S"] - 1):
        hi, lo = int(M["seg_hi"][s]), int(M["seg_lo"][s + 1])
        out.append(hi + (lo - hi) // 2)
    out = sorted(set(v for v in out if v != 0))
    ndiv = max(1, (len(out) + 1 + max_coarse - 1) // max_coarse)
    return np.array(out, U64), ndiv


def partition_of(bnd, u):
    """Range partition of ordered keys u: the number of boundaries <= u."""
    return np.searchsorted(np.asarray(bnd, U64), np.asarray(u, U64), side="right")


def quantile_bounds(sorted_sample, nparts):
    """The method without the model: boundaries at equal-count quantiles of
    the sorted sample."""
    s = np.asarray(sorted_sample, U64)
    idx = (np.arange(1, nparts) * s.size) // nparts
    return np.unique(s[idx])
