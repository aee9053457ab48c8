"""CPU checks of the raw binary table loader's file handling: which files a
pattern names and in what order (MemMappedTable::load,
operators/loaders/table.cpp:170-347: glob's sorted order, directories, empty
and non-regular files skipped), and the errors raised before any device
work."""
import os

import numpy as np
import pytest

import pjoin as P


def _files(tmp_path, sizes, width):
    rng = np.random.default_rng(len(sizes))
    datas = []
    for i, n in enumerate(sizes):
        d = rng.integers(0, 256, n * width, dtype=np.uint8)
        (tmp_path / ("part%02d.bin" % i)).write_bytes(d.tobytes())
        datas.append(d)
    return datas


def test_tuples_counted_over_every_matching_file(tmp_path):
    _files(tmp_path, [3, 0, 1000, 7], 12)
    (tmp_path / "part99.bin").mkdir()          # a directory that matches: skipped
    assert P.binary_tuples(str(tmp_path / "part*.bin"), 12) == 1010
    assert P.binary_tuples(str(tmp_path / "part0[0-1].bin"), 12) == 3


def test_no_match_is_an_empty_table(tmp_path):
    assert P.binary_tuples(str(tmp_path / "nothing*"), 16) == 0


def test_partial_tuple_is_an_error(tmp_path):
    (tmp_path / "a.bin").write_bytes(b"x" * 25)
    with pytest.raises(P.PJError) as e:
        P.binary_tuples(str(tmp_path / "*.bin"), 12)
    assert e.value.code == P.PJ_ERR_INVALID
    assert b"a.bin" in P.lib().pj_last_error()


def test_capacity_checked_before_device_work(tmp_path):
    _files(tmp_path, [100], 16)
    with pytest.raises(P.PJError) as e:
        P.load_binary(str(tmp_path / "*.bin"), 16, 0, 99)
    assert e.value.code == P.PJ_ERR_CAPACITY


def test_empty_table_loads_nothing(tmp_path):
    (tmp_path / "empty.bin").write_bytes(b"")
    assert P.load_binary(str(tmp_path / "*.bin"), 16, 0, 0) == 0
