"""The CPU legs of bench.py (the baselines it times beside the device run),
exercised on tiny workloads so that the script's host-side code is checked
without a GPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402
import oracle as O  # noqa: E402


def _check(res, n):
    assert res["unit"] == "tuples/s" and res["value"] > 0
    assert res["kind"] in ("reference", "port") and res["cores"] >= 1
    assert res["out_tuples"] == n


def test_cpu_baseline_uniform():
    nb, npr, nmiss = 1 << 10, 1 << 12, 400
    res = bench.cpu_baseline(5, 1 << 40, nb, npr, nmiss)
    _check(res, npr - nmiss)


def test_cpu_baseline_zipf():
    nb, npr = 1 << 10, 1 << 12
    res = bench.cpu_baseline_zipf(6, nb, npr, O.zipf_cdf(nb, 1.0))
    _check(res, npr)


def test_cpu_baseline_batch():
    res = bench.cpu_baseline_batch(7, [1024, 2048])
    assert res["value"] > 0 and res["cores"] >= 1


def test_bench_refuses_several_gpus(monkeypatch, capsys):
    monkeypatch.setattr(sys, "argv", ["bench.py", "--gpus", "2"])
    with pytest.raises(SystemExit) as e:
        bench.main()
    assert e.value.code == 2
    assert "one GPU" in capsys.readouterr().err
