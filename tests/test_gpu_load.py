"""GPU tests of pj_load_binary: the device buffer holds the matching files'
bytes back to back in glob's sorted order (MemMappedTable::load,
operators/loaders/table.cpp:170-347), across many staging chunks and
repeated calls, and a join over tables loaded from files equals the
generator's closed form."""
import sys

import numpy as np
import pytest

import oracle as O
import pjoin as P

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


def _write(tmp_path, name, data):
    p = tmp_path / name
    p.write_bytes(np.ascontiguousarray(data).tobytes())
    return p


@pytest.mark.parametrize("sizes,width", [([1], 16), ([5, 0, 1000, 3], 12), ([1 << 20, 777, (3 << 20) + 5], 24)])
def test_loaded_bytes_equal_the_files(tmp_path, sizes, width):
    _say("case: load sizes=%s width=%d" % (sizes, width))
    rng = np.random.default_rng(sum(sizes))
    datas = []
    for i, n in enumerate(sizes):
        d = rng.integers(0, 256, n * width, dtype=np.uint8)
        _write(tmp_path, "t%03d.bin" % (len(sizes) - i), d)   # names sort in reverse of creation
        datas.append(d)
    (tmp_path / "t999.bin").mkdir()
    expect = np.concatenate(datas[::-1])
    n = sum(sizes)
    d = torch.full((n * width + 64,), 0xEE, dtype=torch.uint8, device="cuda")
    for _ in range(2):   # the staging buffers are reused
        got = P.load_binary(str(tmp_path / "t*.bin"), width, d.data_ptr(), n)
        assert got == n
        host = d.cpu().numpy()
        assert np.array_equal(host[: n * width], expect)
        assert (host[n * width:] == 0xEE).all()   # nothing written past the tables


def test_large_table_on_a_side_stream(tmp_path):
    """A 600 MB table (many 32 MiB chunks, the last one partial) through a
    stream of its own, checked by the device checksum."""
    _say("case: load 600 MB")
    n = 600 * (1 << 20) // 16 + 3
    b, p = O.gen_pkfk(61, 1 << 40, 1 << 10, n, 0)
    _write(tmp_path, "big.bin", p)
    d = torch.empty(n * 16, dtype=torch.uint8, device="cuda")
    s = torch.cuda.Stream()
    assert P.load_binary(str(tmp_path / "big.bin"), 16, d.data_ptr(), n, s.cuda_stream) == n
    got, _ = P.checksum_device(d.data_ptr(), n, 16)
    exp, _ = O.checksum(p.view(np.uint8).reshape(n, 16), 16)
    assert got == exp


def test_join_over_loaded_tables(tmp_path):
    """Build and probe relations written to files (the probe side split over
    three files), loaded into HBM and joined: count and checksum equal the
    generator's closed form."""
    _say("case: join over loaded tables")
    nb, npr, nmiss = 1 << 16, 1 << 20, 1 << 16
    b, p = O.gen_pkfk(62, 1 << 40, nb, npr, nmiss)
    _write(tmp_path, "build.bin", b)
    pv = p.view(np.uint8).reshape(npr, 16)
    cuts = [0, 1000, 600000, npr]
    for i in range(3):
        _write(tmp_path, "probe_%d.bin" % i, pv[cuts[i]:cuts[i + 1]])
    db = torch.empty(nb * 16, dtype=torch.uint8, device="cuda")
    dp = torch.empty(npr * 16, dtype=torch.uint8, device="cuda")
    assert P.load_binary(str(tmp_path / "build.bin"), 16, db.data_ptr(), nb) == nb
    assert P.load_binary(str(tmp_path / "probe_*.bin"), 16, dp.data_ptr(), npr) == npr
    j = P.Join(P.make_desc(["long", "long"], ["long", "long"], [0], [0], ["B$0", "B$1", "P$1"]))
    out = torch.empty(npr * j.out_width, dtype=torch.uint8, device="cuda")
    n = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, out.data_ptr(), npr)
    s, _ = P.checksum_device(out.data_ptr(), n, j.out_width)
    es, ec = O.expected_pkfk(62, 1 << 40, nb, npr, nmiss)
    assert (n, s) == (ec, es)
    j.close()
