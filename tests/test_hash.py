"""The reference's TupleHasher functions on the device (pj_hash_buckets,
pj_tuple_hash_device) against the oracle's restatement (oracle/oracle.c,
itself pinned against the reference's own hashers): bucket counts on the
CPU, every tuple's bucket on the GPU."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import pjoin as P

SPECS = [
    ("modulo", 1, {}), ("modulo", 7, {}), ("modulo", 1024, {}), ("modulo", 1 << 20, {}),
    ("parammodulo", 1024, dict(offset=5, skipbits=2)), ("parammodulo", 100, dict(skipbits=7)),
    ("knuth", 64, {}), ("knuth", 4096, dict(skipbits=3)),
    ("range", 16, dict(range_min=1, range_max=20)), ("range", 1000, dict(range_min=-(1 << 30), range_max=1 << 30)),
    ("exactrange", 13, dict(range_min=0, range_max=999)), ("exactrange", 127, dict(range_min=-500, range_max=5000)),
    ("tpchorderkey", 512, {}), ("willis", 256, {}), ("willis", 1 << 16, {}),
    ("alwayszero", 1, {}), ("tpchq1magic", 4, {}),
]


def _oracle_hasher(fn, buckets, kw):
    return O.hasher(fn, buckets, kw.get("range_min", 0), kw.get("range_max", 0), kw.get("offset", 0),
                    kw.get("skipbits", 0))


@pytest.mark.parametrize("fn,buckets,kw", SPECS + [("bytes", 16, {}), ("bytes", 1000, {})])
def test_bucket_counts_equal_the_oracle(fn, buckets, kw):
    h = _oracle_hasher(fn, buckets, kw)
    assert P.hash_buckets(P.hash_spec(fn, buckets, **kw)) == O.lib().or_hasher_buckets(C.byref(h))


def test_bad_specs_are_rejected():
    lo, hi = -(1 << 63), (1 << 63) - 1
    for spec in (P.hash_spec("modulo", 0), P.hash_spec("range", 8, range_min=5, range_max=4),
                 P.hash_spec("exactrange", 8, range_min=5, range_max=4), P.HashSpec(99, 8, 0, 0, 0, 0),
                 # spans whose LONG arithmetic (max - min + 1, max - min + buckets) would overflow
                 P.hash_spec("range", 8, range_min=lo, range_max=hi),
                 P.hash_spec("range", 8, range_min=-1, range_max=hi),
                 P.hash_spec("exactrange", 8, range_min=lo, range_max=hi),
                 P.hash_spec("exactrange", 8, range_min=0, range_max=hi - 7)):
        with pytest.raises(P.PJError) as e:
            P.hash_buckets(spec)
        assert e.value.code == P.PJ_ERR_INVALID
    # the widest spans that still fit
    for fn, kw in (("range", dict(range_min=0, range_max=hi - 1)), ("exactrange", dict(range_min=0, range_max=hi - 8))):
        h = _oracle_hasher(fn, 8, kw)
        assert P.hash_buckets(P.hash_spec(fn, 8, **kw)) == O.lib().or_hasher_buckets(C.byref(h)) == 8


def test_generate_equals_the_reference_test_and_the_oracle():
    # unit_tests/testhash.cpp:53-69: parammodulo(0, 1024, skipbits 2).generate(3)
    # gives functions of 8, 8 and 16 buckets
    g = P.hash_generate(P.hash_spec("parammodulo", 1024, skipbits=2), 3)
    assert [P.hash_buckets(f) for f in g] == [8, 8, 16]
    for buckets in (2, 3, 100, 1024, 1 << 16, 1 << 31):
        for skip in (0, 2, 9):
            for passes in (1, 2, 3, 4, 7):
                g = P.hash_generate(P.hash_spec("parammodulo", buckets, offset=11, skipbits=skip), passes)
                b = (C.c_uint * passes)()
                s = (C.c_uint * passes)()
                assert O.lib().or_parammodulo_generate(11, buckets, skip, passes, b, s) == passes
                assert [P.hash_buckets(f) for f in g] == list(b)
                assert [f.skipbits for f in g] == list(s)
                assert all(f.offset == 11 and f.fn == P.HASH_FNS["parammodulo"] for f in g)
                # the passes' masks are disjoint and cover the function's, except
                # at 2 buckets, where getlogarithm(buckets() - 1) = getlogarithm(1)
                # is 0 and the reference's split keeps no bit
                masks = [(P.hash_buckets(f) - 1) << f.skipbits for f in g]
                whole = (P.hash_buckets(P.hash_spec("parammodulo", buckets, skipbits=skip)) - 1) << skip
                if buckets == 2:
                    whole = 0
                assert sum(masks) == whole and all(a & c == 0 for i, a in enumerate(masks) for c in masks[i + 1:])


def test_generate_rejects_what_has_no_split():
    for spec, passes in ((P.hash_spec("parammodulo", 1024), 0), (P.hash_spec("parammodulo", 1), 2),
                         (P.hash_spec("modulo", 1024), 2), (P.hash_spec("parammodulo", 0), 2)):
        with pytest.raises(P.PJError) as e:
            P.hash_generate(spec, passes)
        assert e.value.code == P.PJ_ERR_INVALID


@pytest.mark.gpu
@pytest.mark.parametrize("passes", [2, 3])
def test_generated_passes_concatenate_to_the_function(passes):
    """A two- or three-pass radix split on the device: each pass's bucket ids
    of the same tuples, most significant first, give the one-pass ids."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(passes)
    n = 1 << 16
    keys = rng.integers(-(1 << 40), 1 << 40, n, dtype=np.int64)
    d = torch.from_numpy(keys.view(np.uint8).copy()).cuda()
    whole = P.hash_spec("parammodulo", 1 << 14, offset=3, skipbits=4)
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    P.tuple_hash_device(whole, d.data_ptr(), n, 8, 0, 8, out.data_ptr())
    exp = out.cpu().numpy().view(np.uint32).copy()
    acc = np.zeros(n, np.uint64)
    for f in P.hash_generate(whole, passes):
        P.tuple_hash_device(f, d.data_ptr(), n, 8, 0, 8, out.data_ptr())
        part = out.cpu().numpy().view(np.uint32).astype(np.uint64)
        assert part.max() < P.hash_buckets(f)
        acc = acc * np.uint64(P.hash_buckets(f)) + part
    assert np.array_equal(acc.astype(np.uint32), exp)
    h = _oracle_hasher("parammodulo", 1 << 14, dict(offset=3, skipbits=4))
    sample = rng.integers(0, n, 500)
    assert all(O.lib().or_hasher_hash(C.byref(h), keys[i:i + 1].ctypes.data_as(C.c_void_p), 8) == exp[i] for i in sample)


def _keys(fn, kw, n, rng):
    if fn in ("range", "exactrange"):
        return rng.integers(kw["range_min"], kw["range_max"] + 1, n)
    return np.concatenate([rng.integers(-(1 << 63), (1 << 63) - 1, n - 5, dtype=np.int64),
                           np.array([0, -1, 1, -(1 << 63), (1 << 63) - 1], np.int64)])


@pytest.mark.gpu
@pytest.mark.parametrize("fn,buckets,kw", SPECS)
@pytest.mark.parametrize("size", [4, 8])
def test_value_hashers_equal_the_oracle(fn, buckets, kw, size):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(buckets + size)
    n = 20000
    keys = _keys(fn, kw, n, rng)
    # 20-byte tuples: 3 bytes of padding, the key at byte 3 (unaligned), the rest filler
    t = np.zeros((n, 20), np.uint8)
    kb = keys.astype("<i4" if size == 4 else "<i8").view(np.uint8).reshape(n, size)
    t[:, 3:3 + size] = kb
    t[:, 3 + size:] = rng.integers(0, 256, (n, 17 - size))
    d = torch.from_numpy(t.reshape(-1).copy()).cuda()
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    P.tuple_hash_device(P.hash_spec(fn, buckets, **kw), d.data_ptr(), n, 20, 3, size, out.data_ptr())
    got = out.cpu().numpy().view(np.uint32)
    h = _oracle_hasher(fn, buckets, kw)
    exp = np.array([O.lib().or_hasher_hash(C.byref(h), t[i, 3:].ctypes.data_as(C.c_void_p), size)
                    for i in range(n)], np.uint32)
    assert np.array_equal(got, exp)


@pytest.mark.gpu
@pytest.mark.parametrize("buckets", [2, 1000, 1 << 16])
@pytest.mark.parametrize("size", [1, 5, 8, 12, 16])
def test_byte_hasher_equals_the_oracle(buckets, size):
    """FNV-1a over sign-extended bytes (bytes >= 0x80 included), folded."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(size * 7 + buckets)
    n = 5000
    t = rng.integers(0, 256, (n, 24), dtype=np.uint8)
    d = torch.from_numpy(t.reshape(-1).copy()).cuda()
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    P.tuple_hash_device(P.hash_spec("bytes", buckets), d.data_ptr(), n, 24, 2, size, out.data_ptr())
    got = out.cpu().numpy().view(np.uint32)
    h = _oracle_hasher("bytes", buckets, {})
    exp = np.array([O.lib().or_hasher_hash(C.byref(h), t[i, 2:].ctypes.data_as(C.c_void_p), size)
                    for i in range(n)], np.uint32)
    assert np.array_equal(got, exp)
