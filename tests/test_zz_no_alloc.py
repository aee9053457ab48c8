"""pj_join_run_device_noalloc against pj_join_run_device.

The no-allocation entry point runs the join in the library's fixed device
scratch. Each case calls it through the C ABI twice in a row on one handle,
the second time with smaller relations (the scratch keeps the first call's
contents), and checks both outputs against pj_join_run_device on the same
inputs: the same sorted rows, byte for byte. Oversized joins and a handle
with a staged build side are refused before any device work.
"""
import sys

import numpy as np
import pytest

import oracle as O
import pjoin as P
from _util import dtype_of, to_gpu

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
LONGLONG = dict(build_cols=["long", "long", "long"], probe_cols=["long", "long", "dec"], buildjattr=[0, 1],
                probejattr=[0, 1], projection=["B$2", "P$2", "B$0"])


def _rows(out, n, w):
    return out[: n * w].cpu().numpy().reshape(-1, w)


def _both(j, build, probe, nb, npr, cap):
    """(rows of pj_join_run_device_noalloc, rows of pj_join_run_device)."""
    stream = torch.cuda.current_stream().cuda_stream
    db, dp = to_gpu(build), to_gpu(probe)
    a = torch.full((cap * j.out_width + 16,), 0xAB, dtype=torch.uint8, device="cuda")
    b = torch.zeros(cap * j.out_width + 16, dtype=torch.uint8, device="cuda")
    na = j.run_device_noalloc(db.data_ptr(), nb, dp.data_ptr(), npr, a.data_ptr(), cap, stream)
    nb2 = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, b.data_ptr(), cap, stream)
    return _rows(a, na, j.out_width), _rows(b, nb2, j.out_width)


def _ll_tables(nb, npr, seed):
    rng = np.random.default_rng(seed)
    b = np.zeros(nb, dtype=dtype_of(LONGLONG["build_cols"]))
    p = np.zeros(npr, dtype=dtype_of(LONGLONG["probe_cols"]))
    b["f0"] = rng.integers(0, nb // 2, nb) * 0x100000001
    b["f1"] = rng.integers(-2, 2, nb)
    b["f2"] = rng.integers(-(1 << 62), 1 << 62, nb)
    p["f0"] = rng.integers(0, nb // 2 + nb // 20, npr) * 0x100000001
    p["f1"] = rng.integers(-2, 2, npr)
    p["f2"] = rng.standard_normal(npr)
    return b.view(np.uint8).reshape(-1).copy(), p.view(np.uint8).reshape(-1).copy()


def test_noalloc_equals_run_device_two_calls_second_smaller():
    j = P.Join(P.make_desc(**HEADLINE))
    try:
        for (nb, npr, miss, seed) in ((1 << 20, 1 << 22, 1 << 18, 31), (5000, 40000, 3000, 32)):
            _say("case: noalloc headline %d x %d" % (nb, npr))
            b, p = O.gen_pkfk(seed, 1 << 40, nb, npr, miss)
            got, ref = _both(j, b.view(np.uint8), p.view(np.uint8), nb, npr, npr)
            assert got.shape[0] == npr - miss
            assert np.array_equal(O.sorted_rows(got), O.sorted_rows(ref))
    finally:
        j.close()


def test_noalloc_row_id_form_two_calls_second_smaller():
    j = P.Join(P.make_desc(**LONGLONG))
    try:
        for (nb, npr, seed) in ((40000, 400000, 33), (3000, 20000, 34)):
            _say("case: noalloc row-id %d x %d" % (nb, npr))
            b, p = _ll_tables(nb, npr, seed)
            got, ref = _both(j, b, p, nb, npr, npr)
            assert got.shape[0] > npr // 4
            assert np.array_equal(O.sorted_rows(got), O.sorted_rows(ref))
    finally:
        j.close()


def test_noalloc_capacity_error_counts_rows():
    _say("case: noalloc capacity error")
    j = P.Join(P.make_desc(**HEADLINE))
    try:
        nb, npr, miss = 4000, 30000, 1000
        b, p = O.gen_pkfk(35, 1 << 40, nb, npr, miss)
        db, dp = to_gpu(b.view(np.uint8)), to_gpu(p.view(np.uint8))
        out = torch.empty(1000 * 24, dtype=torch.uint8, device="cuda")
        with pytest.raises(P.PJError) as e:
            j.run_device_noalloc(db.data_ptr(), nb, dp.data_ptr(), npr, out.data_ptr(), 1000,
                                 torch.cuda.current_stream().cuda_stream)
        assert e.value.code == P.PJ_ERR_CAPACITY and ("need %d output" % (npr - miss)) in str(e.value)
    finally:
        j.close()


def test_noalloc_refuses_a_join_beyond_its_scratch():
    _say("case: noalloc oversized join refused")
    j = P.Join(P.make_desc(**HEADLINE))
    try:
        nb, npr = 1 << 22, 1 << 24   # 64 MiB + 256 MiB of partitioned copies
        out = torch.empty(16, dtype=torch.uint8, device="cuda")
        with pytest.raises(P.PJError) as e:
            # refused before any device work: the (unused) pointers are never read
            j.run_device_noalloc(out.data_ptr(), nb, out.data_ptr(), npr, out.data_ptr(), 1,
                                 torch.cuda.current_stream().cuda_stream)
        assert e.value.code == P.PJ_ERR_UNSUPPORTED and "scratch" in str(e.value)
    finally:
        j.close()
