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


def _tensors(seed, nb, npr, nmiss):
    torch = pytest.importorskip("torch")
    b, p = O.gen_pkfk(seed, 1 << 40, nb, npr, nmiss)
    return torch.from_numpy(b.view(np.uint8).copy()), torch.from_numpy(p.view(np.uint8).copy()), b, p


def test_cpu_baseline_full_join_under_consume():
    import reference_run as R
    if not R.available():
        pytest.skip("oracle/_ref not built")
    nb, npr, nmiss = 1 << 11, 1 << 14, 1000
    db, dp, b, p = _tensors(8, nb, npr, nmiss)
    d = O.join_desc(["long", "long"], ["long", "long"], [0], [0], ["B$0", "B$1", "P$1"], buckets=1 << 11)
    rows, _ = O.hashjoin(d, b.view(np.uint8), p.view(np.uint8))
    x = O.checksum(rows, 24)[1]
    res = bench.cpu_baseline_full(db, dp, nb, npr, x)
    assert res["kind"] == "reference" and res["value"] > 0 and 1 <= res["cores"] <= 16
    assert res["parity"]["equals_device_xor32"]
    assert len(res["runs_seconds"]) == 5 and set(res["seconds_by_prefetch"]) == {"off", "8", "16", "32"}
    assert res["probe_seconds"] > 0 and res["build_seconds"] > 0 and "nproc" in res["host"]


def test_cpu_baseline_sample_counts_the_sample():
    import reference_run as R
    if not R.available():
        pytest.skip("oracle/_ref not built")
    nb, npr = 1 << 11, 1 << 14
    db, dp, _, _ = _tensors(9, nb, npr, 0)
    res = bench.cpu_baseline_sample(db, dp, nb, npr, 1 << 12)
    assert res["value"] > 0 and "first 4096 of 16384 probe" in res["sample"]


def test_step_kernels_are_labelled_in_launch_order():
    rows = [("void pj::k_gen(...)", 1.0), ("void pj::k_hist<1, false>(...)", 10.0), ("pj::k_scan_small", 0.5),
            ("void pj::k_scatter<1, false>(...)", 20.0), ("void pj::k_hist<1, false>(...)", 30.0),
            ("void pj::k_scatter_wc<1, 8, false, 2>(...)", 40.0), ("pj::k_plan", 0.25),
            ("void pj::k_join<false>(...)", 50.0)]
    got = bench._label_step_kernels(rows)
    assert got == {"k_hist build": [10.0], "scans, partition starts, k_plan": [0.5, 0.25],
                   "k_scatter build": [20.0], "k_hist probe": [30.0], "k_scatter_wc probe": [40.0],
                   "k_join": [50.0]}


def test_step_kernels_of_the_appended_and_two_pass_layouts():
    app = [("void pj::k_hist<1, false>(...)", 10.0), ("void pj::k_scatter<1, false>(...)", 20.0),
           ("void pj::k_scatter_wc<1, 8, false, 2, true>(...)", 40.0), ("pj::k_scan_reduce", 0.5),
           ("pj::k_chunk_place(...)", 3.0), ("pj::k_chunk_bounds(...)", 0.125), ("pj::k_plan", 0.25),
           ("void pj::k_join<false, true>(...)", 50.0)]
    got = bench._label_step_kernels(app)
    assert got == {"k_hist build": [10.0], "k_scatter build": [20.0],
                   bench.PROBE_SCATTER_APP: [40.0, 3.0, 0.125], "scans, partition starts, k_plan": [0.5, 0.25],
                   "k_join": [50.0]}
    two = [("void pj::k_hist<1, false>(...)", 10.0), ("void pj::k_scatter<1, false>(...)", 20.0),
           ("void pj::k_hist<1, false>(...)", 30.0), ("pj::k_coarsen(...)", 0.5),
           ("void pj::k_scatter_wc<1, 8, false, 2, false>(...)", 40.0), ("pj::k_subsplit(...)", 45.0),
           ("void pj::k_join<false, false>(...)", 50.0)]
    got = bench._label_step_kernels(two)
    assert got == {"k_hist build": [10.0], "k_scatter build": [20.0], "k_hist probe": [30.0],
                   bench.PROBE_SCATTER_TWO: [0.5, 40.0, 45.0], "k_join": [50.0]}


def test_traffic_summary_is_used_only_for_its_library(tmp_path, monkeypatch):
    monkeypatch.setattr(bench, "PROFILES", str(tmp_path))
    d = {"library_sha256_16": bench.lib_sha(), "kernels": {"k_join": {"read": 1.0, "write": 2.0}}}
    (tmp_path / "traffic_w.md").write_text("# t\n\n```json\n%s\n```\n" % json_dumps(d))
    assert bench.read_traffic("w") == d["kernels"]
    d["library_sha256_16"] = "0" * 16
    (tmp_path / "traffic_w.md").write_text("```json\n%s\n```\n" % json_dumps(d))
    assert bench.read_traffic("w") is None
    assert bench.read_traffic("absent") is None


def json_dumps(d):
    import json
    return json.dumps(d)
