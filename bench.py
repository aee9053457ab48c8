"""Benchmark of the MI355X hash join: join tuples/s (build + probe) and HBM
GB/s on the 16M x 256M uniform int64 join (keys a seeded sample of
[0, 2^40), 10% of probes miss), plain radix partitioning, inputs resident in
HBM. One step is one whole join: partition R, partition S, LDS build and
probe, output written to HBM. The default run also times the clustered
64M x 512M join under radix and learned partitions (learned_vs_radix) and
the reference HashJoinOp on the host cores.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--partitioner radix]
Prints one JSON line.

python bench.py --collect-traffic [--workload W]
Runs this benchmark twice under rocprofv3 (FETCH_SIZE, then WRITE_SIZE),
scales the counters by a 1 GiB copy's, and writes the HBM bytes of every
kernel of one step to profiles/traffic_<workload>.md together with the
library's hash; the benchmark reports those bytes as `traffic` only while
that hash is the current library's.
"""
import argparse
import ctypes as C
import glob
import hashlib
import json
import os
import shutil
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pythia-core_amd"))

HBM_PEAK_GBS = 8000.0

LIB_SO = os.path.join(ROOT, "pythia-core_amd", "lib", "libpjoin.so")
PROFILES = os.path.join(ROOT, "profiles")

WORKLOADS = {
    # name: (seed, key domain, |R|, |S|, misses)
    "uniform_16Mx256M": (202, 1 << 40, 1 << 24, 1 << 28, round(0.1 * (1 << 28))),
    "uniform_1Mx16M": (101, 1 << 40, 1 << 20, 1 << 24, 0),
    "dense_1Mx16M": (101, 1 << 20, 1 << 20, 1 << 24, 0),
}
# Composite (int32 a, int32 b) keys, probe side Zipf(z) over build ranks.
ZIPF_WORKLOADS = {
    # name: (seed, |R|, |S|)
    "zipf_composite_8Mx128M": (303, 1 << 23, 1 << 27),
}
# int64 build keys from 256 dense clusters (widths 2^10..2^24, weights
# Zipf 1.2), probe keys Zipf(--zipf, default 1.0) over build ranks.
CLUSTERED_WORKLOADS = {
    # name: (seed, |R|, |S|)
    "clustered_64Mx512M": (404, 1 << 26, 1 << 29),
}
# Many independent (int64 key, int32 payload) joins: build sizes log-uniform
# in [nmin, nmax], probe = 4 x build with a uniform fk.
BATCH_WORKLOADS = {
    # name: (seed, joins, nmin, nmax)
    "batch_512": (505, 512, 1 << 10, 1 << 22),
}


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def bench_batch(args):
    """The batch workload: every join a whole radix join with the radix bits
    tuned for its size class. value: the joins spread over four streams and
    captured into one graph, run as one launch (PJ_BATCH_CONCURRENT_GRAPH);
    beside it the same four streams launched directly (PJ_BATCH_CONCURRENT,
    which also gives each join's own device time), the joins one after
    another (PJ_BATCH_SEQUENTIAL, one launch sequence per join) and with
    untuned bits. One step is the whole batch; value is aggregate build +
    probe tuples/s."""
    import numpy as np
    import torch
    import pjoin as P

    seed, nj, nmin, nmax = BATCH_WORKLOADS[args.workload]
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    sizes = [int(x) for x in P.batch_sizes(seed, nj, nmin, nmax)]

    def up(n):
        return (n + 255) // 256 * 256

    offs, tot = [], 0
    for nb in sizes:
        ob = tot
        op = ob + up(nb * 12)
        oo = op + up(4 * nb * 12)
        tot = oo + up(4 * nb * 16)
        offs.append((ob, op, oo))
    arena = torch.empty(tot, dtype=torch.uint8, device=dev)
    base = arena.data_ptr()
    joins = []
    for j, (nb, (ob, op, oo)) in enumerate(zip(sizes, offs)):
        P.check(P.lib().pj_gen_batch_join(seed, j, nb, 4 * nb, C.c_void_p(base + ob), C.c_void_p(base + op), 1,
                                          C.c_void_p(stream)))
        joins.append((base + ob, nb, base + op, 4 * nb, base + oo, 4 * nb))
    torch.cuda.synchronize()
    _say("batch: %d joins generated, %.2f GB resident" % (nj, tot / 1e9))
    desc = P.make_desc(["long", "int"], ["long", "int"], [0], [0], ["B$0", "B$1", "P$1"])
    bt = P.Batch(desc)
    bt.prepare(joins)
    tuples = sum(5 * nb for nb in sizes)

    def timed(mode, steps):
        for _ in range(args.warmup):
            bt.run(mode, stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            bt.run(mode, stream)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        r = {"ms_per_step": ms, "value": tuples / (ms * 1e-3)}
        if mode != P.Batch.CONCURRENT_GRAPH:
            own = bt.run_times()
            r.update({"latency_p50_us": float(np.percentile(own, 50)), "latency_p99_us": float(np.percentile(own, 99)),
                      "latency_max_us": float(own.max())})
        return r

    untuned = timed(P.Batch.SEQUENTIAL, max(1, args.steps // 2))
    _say("batch: sequential, untuned bits %.2f ms/step" % untuned["ms_per_step"])
    t0 = time.perf_counter()
    bt.tune(joins, trial=4, stream=stream)
    tune_s = time.perf_counter() - t0
    bt.prepare(joins)
    seq = timed(P.Batch.SEQUENTIAL, args.steps)
    _say("batch: sequential %.2f ms/step (tuning %.1f s)" % (seq["ms_per_step"], tune_s))
    conc = timed(P.Batch.CONCURRENT, args.steps)
    _say("batch: concurrent %.2f ms/step" % conc["ms_per_step"])
    graph = timed(P.Batch.CONCURRENT_GRAPH, args.steps)
    cnt, done = bt.results()
    _say("batch: one graph launch %.2f ms/step" % graph["ms_per_step"])

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    bad = 0
    for j, (b, nb, p, npr, o, cap) in enumerate(joins):
        es, ec = O.expected_batch_join(seed, j, nb, npr, threads=16)
        s_, _ = P.checksum_device(o, int(cnt[j]), 16, stream)
        bad += (int(cnt[j]), s_) != (ec, es)
        if j % 128 == 127:
            _say("batch: parity checked for %d joins" % (j + 1))
    out_rows = int(cnt.sum())
    ms = graph["ms_per_step"]
    b_alg = sum(nb * 12 + 4 * nb * 12 for nb in sizes) + out_rows * 16
    achieved = b_alg / (ms * 1e-3) / 1e9
    tuned = {}
    for c in range(32):
        if any(int(np.log2(nb)) == c for nb in sizes):
            tuned[c] = bt.config(c)[0]
    res = {
        "metric": "aggregate join tuples/s (build+probe), 512 int64-key joins, build 1K-4M log-uniform, probe 4x; "
                  "one graph launch (4 parallel streams captured), tuned radix bits",
        "value": graph["value"],
        "unit": "tuples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic (seeded generator, inputs resident in HBM)",
        "config": {"workload": args.workload, "seed": seed, "joins": nj, "build_min": nmin, "build_max": nmax,
                   "build_tuples": sum(sizes), "probe_tuples": 4 * sum(sizes), "tuple_bytes": 12,
                   "out_tuple_bytes": 16, "out_tuples": out_rows, "mode": "one graph launch of 4 concurrent streams",
                   "tuned_bits_by_size_class": tuned, "pipeline": "switched off"},
        "latency_us": {"p50": conc["latency_p50_us"], "p99": conc["latency_p99_us"], "max": conc["latency_max_us"],
                       "note": "per join, its own device time on its stream, 4 concurrent streams launched directly"},
        "concurrent_streams": conc,
        "one_launch_per_join": seq,
        "untuned_bits_one_launch_per_join": untuned,
        "tune_seconds": tune_s,
        "parity": {"joins": nj, "mismatched_joins": bad},
        "roofline": {"bound": "hbm", "kernel": "whole batch (all kernels of every join)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None},
    }
    del arena
    torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline_batch(seed, sizes)
    print(json.dumps(res))
    if bad:
        sys.exit(1)


def cpu_baseline(seed, dom, nb, npr, nmiss):
    """Times the reference HashJoinOp (oracle/_ref) when it was built, else
    the oracle's single-threaded restatement, on a bounded sample: the whole
    build relation and the first probe tuples of the same probe relation."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O

    sample = min(npr, 1 << 25)
    b, p = O.gen_pkfk(seed, dom, nb, npr, nmiss, probe_prefix=sample)
    what = "|R|=%d build + first %d of %d probe tuples" % (nb, sample, npr)
    import reference_run as R
    ref = R if R.available() else None
    if ref is not None:
        threads = ref.default_threads()
        # probeprefetchstep / buildprefetchstep swept, the fastest kept
        runs = {}
        for pf in (-1, 8, 16, 32):
            r = ref.run_pkfk(b, p, threads, prefetch=pf)
            r.pop("rows", None)
            runs[pf] = r
            _say("cpu baseline: prefetch %d, %.3f s" % (pf, r["seconds"]))
        best = min(runs, key=lambda k: runs[k]["seconds"])
        r = runs[best]
        return {"value": (nb + sample) / r["seconds"], "unit": "tuples/s", "cores": threads,
                "kind": "reference", "sample": what + "; reference HashJoinOp built -O3 from its sources, "
                "under MergeOp, modulo hash over next_pow2(|R|) buckets of 2 tuples, 1 MiB pages, %d threads, "
                "prefetch step %s (fastest of off, 8, 16, 32)" % (threads, "off" if best < 0 else best),
                "seconds": r["seconds"], "build_seconds": r["build_seconds"], "out_tuples": r["out_tuples"],
                "seconds_by_prefetch": {("off" if k < 0 else str(k)): v["seconds"] for k, v in runs.items()}}
    d = O.join_desc(["long", "long"], ["long", "long"], [0], [0], ["B$0", "B$1", "P$1"],
                    buckets=1 << max(1, (nb - 1).bit_length()), tuplesperbucket=2)
    t0 = time.perf_counter()
    out, st = O.hashjoin(d, b.view(np.uint8), p.view(np.uint8))
    dt = time.perf_counter() - t0
    return {"value": (nb + sample) / dt, "unit": "tuples/s", "cores": 1, "kind": "port",
            "sample": what + "; restated HashJoinOp, one thread", "out_tuples": int(st.out_tuples)}


def cpu_baseline_zipf(seed, nb, npr, cdf):
    """The reference HashJoinOp (bytes hash over the two key columns, as
    queryhashjoincomposite.cpp runs it) on the whole build side and the first
    probe tuples; the oracle restatement when the reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    import reference_run as R
    sample = min(npr, 1 << 24)
    b, p = O.gen_pairs_zipf(seed, nb, sample, cdf)
    cols = ["int", "int", "long"]
    proj = ["B$0", "B$1", "B$2", "P$2"]
    buckets = 1 << max(1, (nb - 1).bit_length())
    what = "|R|=%d build + %d Zipf probe tuples" % (nb, sample)
    if R.available():
        threads = R.default_threads()
        rows, tb, tt = R.run_join(cols, cols, [0, 1], [0, 1], proj, b, p, threads=threads, fn="bytes",
                                  buckets=buckets, tuplesperbucket=2)
        return {"value": (nb + sample) / tt, "unit": "tuples/s", "cores": threads, "kind": "reference",
                "sample": what + "; reference HashJoinOp, bytes hash, %d threads" % threads, "seconds": tt,
                "out_tuples": int(rows.shape[0])}
    d = O.join_desc(cols, cols, [0, 1], [0, 1], proj, fn="bytes", buckets=buckets)
    t0 = time.perf_counter()
    out, st = O.hashjoin(d, b, p)
    dt = time.perf_counter() - t0
    return {"value": (nb + sample) / dt, "unit": "tuples/s", "cores": 1, "kind": "port",
            "sample": what + "; restated HashJoinOp, one thread", "out_tuples": int(st.out_tuples)}




def cpu_baseline_batch(seed, sizes):
    """The reference HashJoinOp (oracle/_ref), else the oracle restatement,
    on every 8th join of the batch (a sample of the same size mix), one join
    after the other, each with a modulo hash over next_pow2(|R|) buckets."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    import reference_run as R
    import pjoin as P
    cols = ["long", "int"]
    proj = ["B$0", "B$1", "P$1"]
    total, secs, used, cores, kind = 0, 0.0, 0, 1, "port"
    for j, nb in enumerate(sizes):
        if j % 8:
            continue
        b = np.zeros(nb * 12, np.uint8)
        p = np.zeros(4 * nb * 12, np.uint8)
        P.check(P.lib().pj_gen_batch_join(seed, j, nb, 4 * nb, b.ctypes.data_as(C.c_void_p),
                                          p.ctypes.data_as(C.c_void_p), 0, None))
        buckets = 1 << max(1, (nb - 1).bit_length())
        if R.available():
            cores, kind = R.default_threads(), "reference"
            _, _, tt = R.run_join(cols, cols, [0], [0], proj, b, p, threads=cores, buckets=buckets)
        else:
            d = O.join_desc(cols, cols, [0], [0], proj, buckets=buckets)
            t0 = time.perf_counter()
            O.hashjoin(d, b, p)
            tt = time.perf_counter() - t0
        secs += tt
        total += 5 * nb
        used += 1
    return {"value": total / secs, "unit": "tuples/s", "cores": cores, "kind": kind,
            "sample": "every 8th join of the batch, %d joins (%d tuples), one after the other; %s" % (
                used, total, "reference HashJoinOp built -O3 from its sources, %d threads" % cores
                if kind == "reference" else "restated HashJoinOp, one thread"), "seconds": secs}


PCIE_PEAK_GBS = 64.0   # PCIe 5.0 x16, one direction


def bench_groupjoin(args, P, torch, desc, db, nb, dp, npr, nmiss, dom, seed, clustered, stream):
    """The count per build key of the join, computed inside the join
    (pj_join_count_by_key: no join rows written), against the same count
    from the join's rows in HBM (pj_join_run_device then pj_agg_run_device),
    which also checks it."""
    j = P.Join(desc)
    groups = torch.empty(nb * 16, dtype=torch.uint8, device=db.device)
    for _ in range(args.warmup):
        g = j.count_by_key(db.data_ptr(), nb, dp.data_ptr(), npr, groups.data_ptr(), nb, stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g = j.count_by_key(db.data_ptr(), nb, dp.data_ptr(), npr, groups.data_ptr(), nb, stream)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # the unfused pipeline on the same inputs: its time (one run after a warm-up) and its groups
    rows = torch.empty(npr * j.out_width, dtype=torch.uint8, device=db.device)
    agg = P.Aggregate(["long"] * 3, [0], "count", None, max_groups=nb)
    ref = torch.empty(nb * 16, dtype=torch.uint8, device=db.device)
    for k in range(2):
        torch.cuda.synchronize()
        u0 = time.perf_counter()
        n = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, rows.data_ptr(), npr, stream)
        g2 = agg.run_device(rows.data_ptr(), n, ref.data_ptr(), nb, stream)
        torch.cuda.synchronize()
        unfused_ms = (time.perf_counter() - u0) * 1e3
    agg.close()
    a = groups[: g * 16].view(torch.int64).view(-1, 2)
    b = ref[: g2 * 16].view(torch.int64).view(-1, 2)
    same = g == g2 and torch.equal(a[torch.argsort(a[:, 0])], b[torch.argsort(b[:, 0])])
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    if clustered:
        es, ec = O.expected_clustered(P.cluster_table(seed, nb), npr, P.zipf_cdf(nb, args.zipf), threads=16)
    else:
        es, ec = O.expected_pkfk(seed, dom, nb, npr, nmiss, threads=16)
    parity = same and int(a[:, 1].sum()) == ec == n
    b_alg = (nb + npr) * 16 + g * 16
    res = {
        "metric": "join + count per build key, tuples/s (build+probe), %s; counted inside the join" % args.workload,
        "value": (nb + npr) / (ms * 1e-3), "unit": "tuples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int64", "data": "synthetic (seeded generator, inputs resident in HBM)",
        "config": {"workload": args.workload, "aggregate": "count_by_key_fused", "build_tuples": nb,
                   "probe_tuples": npr, "groups": g, "join_rows_counted": int(a[:, 1].sum())},
        "join_then_aggregate_ms": unfused_ms,
        "parity": {"groups": g, "groups_from_join_rows": g2, "same_groups": bool(same), "expected_rows": ec,
                   "checksum_ok": bool(parity)},
        "roofline": {"bound": "hbm", "kernel": "whole step (partition passes, k_groupjoin, emit)",
                     "achieved": b_alg / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": b_alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None},
        "cpu_baseline": None,
    }
    j.close()
    print(json.dumps(res))
    if not parity:
        sys.exit(1)


def bench_sortmerge(args, P, torch, desc, db, nb, dp, npr, nmiss, dom, seed, clustered, zipf, stream):
    """The device sort-merge join on the same resident inputs and output rows
    as the hash join: both sides radix-sorted by key, then merged. Checked
    against the same closed form."""
    j = P.SortMergeJoin(desc)
    out = torch.empty(npr * j.out_width, dtype=torch.uint8, device=db.device)
    for _ in range(args.warmup):
        n = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, out.data_ptr(), npr, stream)
    torch.cuda.synchronize()
    ph = {"extract_ms": 0.0, "sort_ms": 0.0, "merge_ms": 0.0, "total_ms": 0.0}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        n = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, out.data_ptr(), npr, stream)
        t = j.timings()
        for k in ph:
            ph[k] += getattr(t, k) / args.steps
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    s, _ = P.checksum_device(out.data_ptr(), n, j.out_width, stream)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    if clustered or zipf:
        cdf = P.zipf_cdf(nb, args.zipf)
        es, ec = (O.expected_clustered(P.cluster_table(seed, nb), npr, cdf, threads=16) if clustered else
                  O.expected_pairs_zipf(seed, nb, npr, cdf, threads=16))
    else:
        es, ec = O.expected_pkfk(seed, dom, nb, npr, nmiss, threads=16)
    parity = (n, s) == (ec, es)
    b_alg = (nb + npr) * 16 + n * j.out_width
    res = {
        "metric": "join tuples/s (build+probe), %s; device sort-merge join" % args.workload,
        "value": (nb + npr) / (ms * 1e-3), "unit": "tuples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int64", "data": "synthetic (seeded generator, inputs resident in HBM)",
        "config": {"workload": args.workload, "algorithm": "sortmerge", "build_tuples": nb, "probe_tuples": npr,
                   "out_tuples": n, "build_key_bits": t.build_key_bits, "probe_key_bits": t.probe_key_bits},
        "phases_ms": ph,
        "parity": {"count": n, "expected_count": ec, "checksum_ok": parity},
        "roofline": {"bound": "hbm", "kernel": "whole sort-merge join", "achieved": b_alg / (ms * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": b_alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None},
        "cpu_baseline": None,
    }
    j.close()
    print(json.dumps(res))
    if not parity:
        sys.exit(1)


def load_from_files(args, P, torch, db, nb, dp, npr, stream):
    """Writes the generated relations to raw binary files (the probe side in
    four files), drops the device copies' contents, and loads them back with
    pj_load_binary: files -> pinned staging -> HBM. The files sit in the page
    cache, so this measures the host-to-device path, not a disk."""
    import shutil
    import tempfile
    d = tempfile.mkdtemp(prefix="pj_tables_")
    try:
        (db.cpu().numpy()).tofile(os.path.join(d, "build.bin"))
        hp = dp.cpu().numpy().reshape(npr, 16)
        cuts = [0, npr // 4, npr // 2, 3 * npr // 4, npr]
        for i in range(4):
            hp[cuts[i]:cuts[i + 1]].tofile(os.path.join(d, "probe_%d.bin" % i))
        del hp
        times = []
        for _ in range(3):
            db.zero_()
            dp.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            got_b = P.load_binary(os.path.join(d, "build.bin"), 16, db.data_ptr(), nb, stream)
            got_p = P.load_binary(os.path.join(d, "probe_*.bin"), 16, dp.data_ptr(), npr, stream)
            times.append(time.perf_counter() - t0)
            assert (got_b, got_p) == (nb, npr)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    t = sorted(times)[len(times) // 2]
    nbytes = (nb + npr) * 16
    _say("load: %d bytes in %.1f ms" % (nbytes, t * 1e3))
    return {"bytes": nbytes, "files": 5, "ms_median": t * 1e3, "gbs": nbytes / t / 1e9,
            "pcie_peak_gbs": PCIE_PEAK_GBS, "frac": nbytes / t / 1e9 / PCIE_PEAK_GBS,
            "note": "files in the page cache; the join that follows runs on the loaded tables and its "
                    "parity check covers them"}


def layout_flags(args, P):
    """PJ_FLAG_LAYOUT_* bits of --layout (auto: the library chooses)."""
    return {"auto": 0, "exact": P.PJ_FLAG_LAYOUT_EXACT, "append": P.PJ_FLAG_LAYOUT_APPEND,
            "two_pass": P.PJ_FLAG_LAYOUT_TWO_PASS}[args.layout]


def lib_sha():
    """First 16 hex digits of the SHA-256 of the built library."""
    h = hashlib.sha256()
    with open(LIB_SO, "rb") as f:
        for block in iter(lambda: f.read(1 << 20), b""):
            h.update(block)
    return h.hexdigest()[:16]


def host_info():
    """What the CPU baseline ran on: core counts, model, sockets, NUMA nodes."""
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                             timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("Model name", "Socket(s)", "NUMA node(s)", "Core(s) per socket", "Thread(s) per core"):
                info[k.lower().replace("(s)", "s").replace(" ", "_")] = v
    except (OSError, subprocess.SubprocessError):
        pass
    return info


# ---------------------------------------------------------------- traffic
# The HBM bytes of one step's kernels, per workload, from two rocprofv3
# counter passes; kept in profiles/traffic_<workload>.md as a JSON block.

def traffic_path(workload):
    return os.path.join(PROFILES, "traffic_%s.md" % workload)


def read_traffic(workload):
    """{kernel label: {"read": B, "write": B}} when the summary was made
    from the current library, else None."""
    p = traffic_path(workload)
    if not os.path.exists(p):
        return None
    txt = open(p).read()
    a = txt.find("```json")
    if a < 0:
        return None
    b = txt.find("```", a + 7)
    d = json.loads(txt[a + 7:b])
    if d.get("library_sha256_16") != lib_sha():
        return None
    return d["kernels"]


PROBE_SCATTER = "k_scatter_wc probe"
PROBE_SCATTER_APP = "k_scatter_wc probe (appended) + k_chunk_place"
PROBE_SCATTER_TWO = "k_scatter_wc + k_subsplit probe (two passes)"


def probe_scatter_label(appended, passes):
    return PROBE_SCATTER_APP if appended else PROBE_SCATTER_TWO if passes == 2 else PROBE_SCATTER


def _label_step_kernels(rows):
    """Dispatches of one join step in launch order -> {label: [counter values]}.
    The step launches k_hist for the build side, then (unless the probe side
    is appended) for the probe side, the build side's scatter, the probe
    side's write-combining scatter (with k_chunk_place, or k_coarsen and
    k_subsplit, by its layout), k_plan, k_join."""
    names = [n for n, _ in rows]
    scat = probe_scatter_label(any("k_chunk_place" in n for n in names), 2 if any("k_subsplit" in n for n in names)
                               else 1)
    out, hist_seen = {}, 0
    for name, val in rows:
        if "k_hist" in name:
            label = "k_hist build" if hist_seen == 0 else "k_hist probe"
            hist_seen += 1
        elif any(k in name for k in ("k_scatter_wc", "k_chunk_place", "k_chunk_bounds", "k_subsplit", "k_coarsen")):
            label = scat
        elif "k_scatter" in name:
            label = "k_scatter build"
        elif "k_join" in name:
            label = "k_join"
        elif "k_scan" in name or "k_part_bounds" in name or "k_plan" in name or "k_fill" in name:
            label = "scans, partition starts, k_plan"
        else:
            continue
        out.setdefault(label, []).append(val)
    return out


def _counter_rows(d, counter):
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    import csv
    rows = []
    for f in files:
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") == counter:
                    rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
    rows.sort()
    return [(k, v) for _, k, v in rows]


def collect_traffic(args):
    """Two rocprofv3 passes over one step of the workload; writes the summary."""
    base = os.path.join(ROOT, "gpurun_out", "traffic_%s" % args.workload)
    cmd_tail = ["python3", os.path.abspath(__file__), "--workload", args.workload, "--steps", "1", "--warmup", "0",
                "--no-cpu-baseline", "--no-compare", "--calibrate-copy"]
    per = {}
    cmds = []
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = base + "_" + counter
        cmd = ["rocprofv3", "--pmc", counter, "-d", d, "-o", "run", "--output-format", "csv", "--"] + cmd_tail
        cmds.append(" ".join(cmd).replace(ROOT, "$R"))
        _say("traffic: " + " ".join(cmd))
        subprocess.run(cmd, check=True, cwd="/tmp", timeout=200)
        rows = _counter_rows(d, counter)
        copies = [v for k, v in rows if "copyBuffer" in k or "copy_kernel" in k or "elementwise" in k]
        if not copies:
            raise RuntimeError("no calibration copy in the %s pass" % counter)
        per[counter] = (_label_step_kernels(rows), max(copies))
    gib = 1 << 30
    fetch_factor = gib / (per["FETCH_SIZE"][1] * 1024.0)
    write_factor = gib / (per["WRITE_SIZE"][1] * 1024.0)
    kernels = {}
    for label, vals in per["FETCH_SIZE"][0].items():
        wv = per["WRITE_SIZE"][0].get(label, [0.0])
        kernels[label] = {"read": sum(vals) * 1024.0 * fetch_factor, "write": sum(wv) * 1024.0 * write_factor,
                          "dispatches": len(vals)}
    d = {"library_sha256_16": lib_sha(), "workload": args.workload, "fetch_factor": fetch_factor,
         "write_factor": write_factor, "kernels": kernels}
    lines = ["# HBM traffic of one %s step, per kernel" % args.workload, "",
             "rocprofv3 counter passes (one counter a pass), each scaled by what the counter reads for the "
             "run's 1 GiB device copy (FETCH x %.2f, WRITE x %.2f). bench.py reports these bytes as `traffic` "
             "while the library's hash below is the built library's." % (fetch_factor, write_factor), "",
             "Commands (from /tmp, repository at $R), written by `python bench.py --collect-traffic "
             "--workload %s`:" % args.workload, ""]
    lines += ["    " + c for c in cmds]
    lines += ["", "| kernel | dispatches | HBM read (GB) | HBM write (GB) |", "|---|---|---|---|"]
    for label, v in sorted(kernels.items()):
        lines.append("| %s | %d | %.3f | %.3f |" % (label, v["dispatches"], v["read"] / 1e9, v["write"] / 1e9))
    lines += ["", "```json", json.dumps(d, indent=1), "```", ""]
    for path in (traffic_path(args.workload), os.path.join(ROOT, "gpurun_out", os.path.basename(
            traffic_path(args.workload)))):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            f.write("\n".join(lines))
    print(json.dumps(d))


# ---------------------------------------------------------------- CPU baseline

class _Relations:
    """Host copies of two device relations in files (in /dev/shm where it
    exists) for reference runs in processes of their own; removed on exit."""

    def __init__(self, db, dp):
        import tempfile
        import numpy as np
        d = "/dev/shm" if os.path.isdir("/dev/shm") else tempfile.gettempdir()
        self.paths = [os.path.join(d, "pj_cpu_%d_%s.npy" % (os.getpid(), s)) for s in ("build", "probe")]
        for t, path in zip((db, dp), self.paths):
            np.save(path, t.cpu().numpy().view(np.int64))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        for path in self.paths:
            if os.path.exists(path):
                os.remove(path)

    def run(self, threads, prefetch):
        """One reference run in a new process (a fresh reference allocator)."""
        out = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "reference_run.py"), "consume",
                              self.paths[0], self.paths[1], str(threads), str(prefetch)],
                             stdout=subprocess.PIPE, check=True, text=True, timeout=300).stdout
        return json.loads(out.strip().splitlines()[-1])


# ---------------------------------------------------------------- ceilings
# What the machine gives the join's access patterns, measured by the
# library's measurement kernels (pj_measure) on buffers of the bench's size.

CEILINGS = os.path.join(PROFILES, "ceilings_mi355x.md")


def measure_ceilings(args):
    import torch
    import pjoin as P
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    n = 1 << 28                       # 4 GiB of 16-byte records a buffer
    a = torch.empty(n * 16, dtype=torch.uint8, device=dev)
    b = torch.empty(n * 16, dtype=torch.uint8, device=dev)
    P.check(P.lib().pj_gen_pkfk(7, 1 << 40, n, 0, 0, C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), 1,
                                C.c_void_p(stream)))   # distinct random keys in a
    torch.cuda.synchronize()
    rows = []

    def run(kind, param=0, nn=n, idx=0, note=""):
        r = P.measure(kind, a.data_ptr(), b.data_ptr(), idx, nn, param, 5, stream)
        rows.append(dict(kind=kind, param=param, records=r["ops"], note=note, **r))
        _say("ceiling %-9s param %-6d %10.3f ms  %8.1f GB/s  %7.2f G/s  %s" % (kind, param, r["ms"], r["gbps"],
                                                                                r["gops"], note))

    for g in (2048, 8192, 16384):
        for kind in ("read", "write", "copy"):
            run(kind, g, note="grid %d" % g)
    ng = 1 << 27                      # 2 GiB of records for the random patterns
    perm = torch.randperm(ng, device=dev, dtype=torch.int32)
    torch.cuda.synchronize()
    run("gather", nn=ng, idx=perm.data_ptr(), note="random permutation")
    run("scatter", nn=ng, idx=perm.data_ptr(), note="random permutation")
    del perm
    for bits in (8, 11, 12, 14):
        run("histogram", bits, note="%d bins" % (1 << bits))
    def size(bits):
        b = 16 << bits
        return "%d MiB" % (b >> 20) if b >= 1 << 20 else "%d KiB" % (b >> 10)

    for bits in (12, 14, 16, 17, 18, 19, 20, 22, 24, 25, 26, 27, 28):   # tables of 64 KiB .. 4 GiB
        run("lookup", bits, note="table of %s" % size(bits))
    for bits in (12, 14, 16, 17, 18, 19, 20, 24, 25, 26):
        run("probe", bits, note="stream in, look up in %s, stream out" % size(bits))
    for addrs in (1, 64, 2048, 65536):
        run("atomic", addrs, note="%d counters" % addrs)
    best = {k: max(r["gbps"] for r in rows if r["kind"] == k) for k in ("read", "write", "copy")}
    props = torch.cuda.get_device_properties(0)
    out = {"library_sha": lib_sha(), "device": getattr(props, "gcnArchName", "") or props.name,
           "cus": props.multi_processor_count, "rows": rows, "best_gbps": best}
    with open(CEILINGS, "w") as f:
        f.write("# Ceilings of the join's access patterns on one MI355X\n\n")
        f.write("Command: `python bench.py --ceilings` (the library's pj_measure kernels, "
                "pythia-core_amd/csrc/pj_measure.hip; fastest of 5 launches after a warm-up, HIP events).\n\n")
        f.write("Bytes count what the pattern must move: read or write n x 16, copy 2 x n x 16, gather and "
                "scatter n x (16 + 4 + 16) (record, index, record); lookup n x 16 (random reads into a table of "
                "2^param records). Histogram: n x 16 bytes read, one LDS "
                "atomic per record. Atomic: 2048 x 256 lanes x 64 device atomicAdds, results used.\n\n")
        f.write("| pattern | parameter | records or atomics | ms | GB/s | G ops/s | note |\n"
                "|---|---|---|---|---|---|---|\n")
        for r in rows:
            f.write("| %s | %d | %d | %.3f | %.1f | %.3f | %s |\n" % (r["kind"], r["param"], r["records"], r["ms"],
                                                                     r["gbps"], r["gops"], r["note"]))
        f.write("\n```json\n%s\n```\n" % json.dumps(out, indent=1))
    os.makedirs(os.path.join(ROOT, "gpurun_out"), exist_ok=True)
    shutil.copy(CEILINGS, os.path.join(ROOT, "gpurun_out", os.path.basename(CEILINGS)))
    print(json.dumps({"ceilings": best, "file": os.path.relpath(CEILINGS, ROOT)}))


def cpu_baseline_full(db, dp, nb, npr, gpu_xor32):
    """The reference HashJoinOp (oracle/_ref, -O3 from its sources) on the
    whole join, under MergeOp(T) -> ConsumeOp (operators/consume.cpp), the
    relations copied from the device (the generators are bit-identical on
    host and device). Every run is a process of its own: the reference's
    allocator hands out pre-reserved arenas that the first runs of a process
    use up, after which each build faults in fresh pages and takes about
    three times as long. T = the cores this process may use, at most 16 (the
    host cores that come with one GPU on the GPU box). probeprefetchstep /
    buildprefetchstep: one run each of off, 8, 16, 32 (the warm-up), then 5
    runs with the fastest; medians reported. Parity: ConsumeOp's XOR of the
    output's 32-bit words equals the device output's."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import reference_run as R
    if not R.available():
        return None
    threads = R.default_threads()
    with _Relations(db, dp) as rel:
        sweep = {}
        for pf in (-1, 8, 16, 32):
            r = rel.run(threads, pf)
            sweep[pf] = r
            _say("cpu baseline: prefetch %d, %.3f s (build %.3f s)" % (pf, r["seconds"], r["build_seconds"]))
        best = min(sweep, key=lambda k: sweep[k]["seconds"])
        runs = [rel.run(threads, best) for _ in range(5)]
    med = lambda key: sorted(r[key] for r in runs)[2]
    secs = med("seconds")
    info = host_info()
    return {"value": (nb + npr) / secs, "unit": "tuples/s", "cores": threads, "kind": "reference",
            "sample": "the whole %d x %d join (no sample): the reference HashJoinOp built -O3 from its sources "
                      "under MergeOp(%d) -> ConsumeOp, modulo hash over next_pow2(|R|) buckets of 2 tuples, "
                      "1 MiB pages, prefetch step %s (fastest of off, 8, 16, 32); median of 5 runs after one "
                      "run of each prefetch step" % (nb, npr, threads, "off" if best < 0 else best),
            "seconds": secs, "build_seconds": med("build_seconds"), "probe_seconds": med("probe_seconds"),
            "seconds_by_prefetch": {("off" if k < 0 else str(k)): v["seconds"] for k, v in sweep.items()},
            "runs_seconds": [r["seconds"] for r in runs],
            "parity": {"xor32": runs[0]["xor32"], "equals_device_xor32": runs[0]["xor32"] == gpu_xor32},
            "threads_cap": "min(cores this process may use, 16): the GPU box's CPU share is 16 threads a GPU",
            "host": info}


def cpu_baseline_sample(db, dp, nb, npr, nsample, gpu_xor32=None):
    """The reference HashJoinOp under ConsumeOp on the first (at most 16M)
    build tuples and the first nsample probe tuples (1 warm-up, median of
    3). The reference's allocator fails on the 4 GiB chained table of 64M
    build keys (its numadeallocate then aborts the process), so larger build
    sides are cut to 16M, a 1 GiB table as in the headline join."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import reference_run as R
    if not R.available():
        return None
    nbs = min(nb, 1 << 24)
    threads = R.default_threads()
    with _Relations(db[: nbs * 16], dp[: nsample * 16]) as rel:
        rel.run(threads, 16)
        runs = [rel.run(threads, 16) for _ in range(3)]
    r = sorted(runs, key=lambda x: x["seconds"])[1]
    return {"value": (nbs + nsample) / r["seconds"], "unit": "tuples/s", "cores": threads, "kind": "reference",
            "sample": "first %d of %d build + first %d of %d probe tuples; reference HashJoinOp under "
                      "MergeOp(%d) -> ConsumeOp, modulo hash, prefetch step 16, median of 3 after a warm-up"
                      % (nbs, nb, nsample, npr, threads),
            "seconds": r["seconds"], "build_seconds": r["build_seconds"], "probe_seconds": r["probe_seconds"],
            "host": host_info()}


# ---------------------------------------------------------------- learned vs radix

def compare_learned(args, P, torch, dev, stream, with_cpu):
    """The clustered 64M x 512M join (256 dense clusters, Zipf-1.0 fk) under
    radix and under learned partitions, in this run, each checked against
    the generator's closed form (so both give the same output multiset)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    seed, nb, npr = CLUSTERED_WORKLOADS["clustered_64Mx512M"]
    table = P.cluster_table(seed, nb)
    cdf = P.zipf_cdf(nb, 1.0)
    db = torch.empty(nb * 16, dtype=torch.uint8, device=dev)
    dp = torch.empty(npr * 16, dtype=torch.uint8, device=dev)
    P.check(P.lib().pj_gen_clustered(table.ctypes.data_as(C.c_void_p), npr, cdf.ctypes.data_as(C.c_void_p),
                                     C.c_void_p(db.data_ptr()), C.c_void_p(dp.data_ptr()), 1, C.c_void_p(stream)))
    out = torch.empty(npr * 24, dtype=torch.uint8, device=dev)
    es, ec = O.expected_clustered(table, npr, cdf, threads=16)
    res = {"workload": "clustered_64Mx512M_z1", "build_tuples": nb, "probe_tuples": npr}
    for part in ("radix", "learned"):
        j = P.Join(P.make_desc(["long", "long"], ["long", "long"], [0], [0], ["B$0", "B$1", "P$1"],
                               partitioner=part, timing=True, flags=layout_flags(args, P)))
        for _ in range(max(1, args.warmup)):
            n = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, out.data_ptr(), npr, stream)
        torch.cuda.synchronize()
        steps = max(3, args.steps // 2)
        ph = {"model_ms": 0.0, "build_partition_ms": 0.0, "probe_partition_ms": 0.0, "join_ms": 0.0}
        t0 = time.perf_counter()
        for _ in range(steps):
            n = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, out.data_ptr(), npr, stream)
            t = j.timings()
            for k in ph:
                ph[k] += getattr(t, k) / steps
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        s, x = P.checksum_device(out.data_ptr(), n, 24, stream)
        res[part] = {"ms_per_step": ms, "value": (nb + npr) / (ms * 1e-3), "phases_ms": ph,
                     "used_learned": int(t.used_learned), "partitions": int(t.partitions),
                     "work_items": int(t.work_items),
                     "probe_layout": "appended" if t.probe_appended else "contiguous",
                     "probe_passes": int(t.probe_passes),
                     "parity": {"count": n, "checksum": s, "matches_closed_form": (n, s) == (ec, es)}}
        _say("learned_vs_radix: %s %.2f ms/step" % (part, ms))
        j.close()
    res["ratio_learned_over_radix"] = res["radix"]["ms_per_step"] / res["learned"]["ms_per_step"]
    res["identical_output"] = (res["radix"]["parity"]["matches_closed_form"] and
                               res["learned"]["parity"]["matches_closed_form"])
    res["target_ratio"] = 1.3
    del out
    if with_cpu:
        res["cpu_baseline"] = cpu_baseline_sample(db, dp, nb, npr, 1 << 26)
    del db, dp
    torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="uniform_16Mx256M", choices=sorted(WORKLOADS) + sorted(ZIPF_WORKLOADS) + sorted(CLUSTERED_WORKLOADS) +
                    sorted(BATCH_WORKLOADS))
    ap.add_argument("--zipf", type=float, default=1.0, help="Zipf exponent of the composite-key workload")
    ap.add_argument("--partitioner", default="radix", choices=["radix", "learned", "auto"])
    ap.add_argument("--layout", default="auto", choices=["auto", "exact", "append", "two_pass"],
                    help="probe-side partition layout: the library's choice, contiguous after a histogram pass, "
                         "appended chunks listed by partition, or two radix passes")
    ap.add_argument("--radix-bits", type=int, default=-1,
                    help="radix bits of the build side (-1: the library's choice)")
    ap.add_argument("--aggregate", default="none", choices=["none", "sum", "count_by_key", "count_by_key_fused"],
                    help="after each join, aggregate its output in HBM: sum of the probe payload, or a count "
                         "per build key (int64-key workloads); count_by_key_fused: the count per build key "
                         "computed inside the join (pj_join_count_by_key), no join rows written")
    ap.add_argument("--algorithm", default="hash", choices=["hash", "sortmerge"],
                    help="hash: the radix hash join; sortmerge: the device sort-merge join (pj_smj_*), "
                         "same inputs and output rows, to compare against")
    ap.add_argument("--from-files", action="store_true",
                    help="write the relations to raw binary files first and load them into HBM with "
                         "pj_load_binary (timed separately; the join's timing is unchanged)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 1 GiB with torch, a kernel of known traffic for scaling HBM counters")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the learned-vs-radix comparison on the clustered 64M x 512M join")
    ap.add_argument("--ceilings", action="store_true",
                    help="measure the read/write/copy, gather/scatter, histogram and atomic rates of the GPU "
                         "(pj_measure) and write profiles/ceilings_mi355x.md")
    ap.add_argument("--collect-traffic", action="store_true",
                    help="write profiles/traffic_<workload>.md from two rocprofv3 counter passes")
    args = ap.parse_args()
    if args.collect_traffic:
        return collect_traffic(args)
    if args.ceilings:
        return measure_ceilings(args)
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (the sharding design is "
              "written up, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload in BATCH_WORKLOADS:
        return bench_batch(args)

    import torch
    import pjoin as P

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    if args.calibrate_copy:
        x = torch.ones(1 << 30, dtype=torch.uint8, device=dev)
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        del x, y
    zipf = args.workload in ZIPF_WORKLOADS
    clustered = args.workload in CLUSTERED_WORKLOADS
    if clustered:
        seed, nb, npr = CLUSTERED_WORKLOADS[args.workload]
        nmiss, dom = 0, None
        table = P.cluster_table(seed, nb)
        cdf = P.zipf_cdf(nb, args.zipf)
        db = torch.empty(nb * 16, dtype=torch.uint8, device=dev)
        dp = torch.empty(npr * 16, dtype=torch.uint8, device=dev)
        P.check(P.lib().pj_gen_clustered(table.ctypes.data_as(C.c_void_p), npr, cdf.ctypes.data_as(C.c_void_p),
                                         C.c_void_p(db.data_ptr()), C.c_void_p(dp.data_ptr()), 1, C.c_void_p(stream)))
        desc = P.make_desc(["long", "long"], ["long", "long"], [0], [0], ["B$0", "B$1", "P$1"],
                           partitioner=args.partitioner, timing=True, flags=layout_flags(args, P),
                           radix_bits=args.radix_bits)
    elif zipf:
        seed, nb, npr = ZIPF_WORKLOADS[args.workload]
        nmiss, dom = 0, None
        cdf = P.zipf_cdf(nb, args.zipf)
        db = torch.empty(nb * 16, dtype=torch.uint8, device=dev)
        dp = torch.empty(npr * 16, dtype=torch.uint8, device=dev)
        P.check(P.lib().pj_gen_pairs_zipf(seed, nb, npr, cdf.ctypes.data_as(C.c_void_p), C.c_void_p(db.data_ptr()),
                                          C.c_void_p(dp.data_ptr()), 1, C.c_void_p(stream)))
        desc = P.make_desc(["int", "int", "long"], ["int", "int", "long"], [0, 1], [0, 1],
                           ["B$0", "B$1", "B$2", "P$2"], partitioner=args.partitioner, timing=True,
                           flags=layout_flags(args, P), radix_bits=args.radix_bits)
    else:
        seed, dom, nb, npr, nmiss = WORKLOADS[args.workload]
        db = torch.empty(nb * 16, dtype=torch.uint8, device=dev)
        dp = torch.empty(npr * 16, dtype=torch.uint8, device=dev)
        P.check(P.lib().pj_gen_pkfk(seed, dom, nb, npr, nmiss, C.c_void_p(db.data_ptr()),
                                    C.c_void_p(dp.data_ptr()), 1, C.c_void_p(stream)))
        desc = P.make_desc(["long", "long"], ["long", "long"], [0], [0], ["B$0", "B$1", "P$1"],
                           partitioner=args.partitioner, timing=True, flags=layout_flags(args, P),
                           radix_bits=args.radix_bits)
    load_res = None
    if args.from_files:
        load_res = load_from_files(args, P, torch, db, nb, dp, npr, stream)
    if args.aggregate == "count_by_key_fused":
        if zipf or args.partitioner != "radix":
            print("--aggregate count_by_key_fused runs on the int64-key workloads, radix", file=sys.stderr)
            sys.exit(2)
        return bench_groupjoin(args, P, torch, desc, db, nb, dp, npr, nmiss, dom, seed, clustered, stream)
    if zipf and args.aggregate != "none":
        print("--aggregate runs on the int64-key workloads", file=sys.stderr)
        sys.exit(2)
    if args.algorithm == "sortmerge":
        if args.aggregate != "none" or args.partitioner != "radix":
            print("--algorithm sortmerge runs the join alone", file=sys.stderr)
            sys.exit(2)
        return bench_sortmerge(args, P, torch, desc, db, nb, dp, npr, nmiss, dom, seed, clustered, zipf, stream)
    j = P.Join(desc)
    cap = npr
    out = torch.empty(cap * j.out_width, dtype=torch.uint8, device=dev)
    agg = None
    if args.aggregate != "none":
        # join output: (B$0 key, B$1 build payload, P$1 probe payload), three LONG columns
        agg = (P.Aggregate(["long"] * 3, [], "sum", 2, max_groups=1) if args.aggregate == "sum" else
               P.Aggregate(["long"] * 3, [0], "count", None, max_groups=nb))
        gcap = 1 if args.aggregate == "sum" else nb
        groups = torch.empty(gcap * agg.out_width, dtype=torch.uint8, device=dev)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        agg_ms = []

    def step():
        n = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, out.data_ptr(), cap, stream)
        if agg is not None:
            ev[0].record()
            step.groups = agg.run_device(out.data_ptr(), n, groups.data_ptr(), gcap, stream)
            ev[1].record()
            ev[1].synchronize()
            agg_ms.append(ev[0].elapsed_time(ev[1]))
        return n

    for _ in range(args.warmup):
        n = step()
    torch.cuda.synchronize()
    phases = {"model_ms": 0.0, "build_partition_ms": 0.0, "probe_partition_ms": 0.0, "join_ms": 0.0,
              "total_ms": 0.0, "build_hist_ms": 0.0, "probe_hist_ms": 0.0}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        n = step()
        t = j.timings()
        for k in phases:
            phases[k] += getattr(t, k)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = dt * 1e3 / args.steps
    for k in phases:
        phases[k] /= args.steps

    # parity of the measured output: count and checksum vs the closed form
    s, gpu_xor32 = P.checksum_device(out.data_ptr(), n, j.out_width, stream)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    if clustered:
        es, ec = O.expected_clustered(table, npr, cdf, threads=16)
    elif zipf:
        es, ec = O.expected_pairs_zipf(seed, nb, npr, cdf, threads=16)
    else:
        es, ec = O.expected_pkfk(seed, dom, nb, npr, nmiss, threads=16)
    parity = (n == ec and s == es)

    agg_res = None
    if agg is not None:
        # size-independent checks against torch's own reductions over the same join output
        rows = out[: n * j.out_width].view(torch.int64).view(-1, 3)
        g = step.groups
        gr = groups[: g * agg.out_width].view(torch.int64)
        if args.aggregate == "sum":
            ok = g == 1 and int(gr[0]) == int(rows[:, 2].sum())
        else:
            gr = gr.view(-1, 2)
            ok = (int(gr[:, 1].sum()) == n and int((gr[:, 0] * gr[:, 1]).sum()) == int(rows[:, 0].sum())
                  and g == int(torch.unique(rows[:, 0]).numel()))
        parity = parity and ok
        a_ms = sorted(agg_ms[args.warmup:])[len(agg_ms[args.warmup:]) // 2]
        a_bytes = n * j.out_width + g * agg.out_width
        agg_res = {"op": args.aggregate, "groups": g, "ms_median": a_ms, "checks_ok": ok,
                   "tuples_per_s": n / (a_ms * 1e-3),
                   "roofline": {"bound": "hbm",
                                "kernel": "k_agg" if args.aggregate == "sum" else
                                "whole aggregation (k_hist, k_scatter_wc, k_agg_parts: partitioned path)",
                                "achieved": a_bytes / (a_ms * 1e-3) / 1e9,
                                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": a_bytes / (a_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None}}
        if not args.no_cpu_baseline:
            # the reference's own AggregateSum / AggregateCount (oracle/_ref) when it was built, else the
            # oracle's numpy aggregation, over the first 16M rows of the same join output
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle as O
            import reference_run as RR
            m = min(n, 1 << 24)
            host = out[: m * j.out_width].cpu().numpy().reshape(m, j.out_width)
            fields, op, sf = ([], "sum", 2) if args.aggregate == "sum" else ([0], "count", None)
            if RR.available():
                threads = RR.default_threads()
                buckets = 1 << max(1, (min(nb, m) - 1).bit_length()) if fields else 1024
                _, cs = RR.run_aggregate(["long"] * 3, fields, op, sf, host, threads=threads, buckets=buckets)
                kind, cores = "reference", threads
                how = "reference GenericAggregate built -O3 from its sources, global mode, modulo hash over %d " \
                      "buckets, %d threads" % (buckets, threads)
            else:
                c0 = time.perf_counter()
                O.aggregate(["long"] * 3, fields, op, sf, host)
                cs = time.perf_counter() - c0
                kind, cores, how = "port", 1, "the oracle's numpy aggregation, one thread"
            agg_res["cpu_baseline"] = {"value": m / cs, "unit": "input tuples/s", "cores": cores, "kind": kind,
                                       "sample": "first %d rows of the join output; %s" % (m, how)}
        agg.close()
    w_in, w_out = 16, j.out_width
    b_alg = nb * w_in + npr * w_in + n * w_out
    job_gbs = b_alg / (ms * 1e-3) / 1e9
    t = j.timings()
    # per kernel: HIP-event times on the library's stream (the histogram
    # times include their scans), algorithmic bytes, and the HBM bytes of
    # the current build's counter summary when there is one
    traffic = read_traffic(args.workload)
    # the probe side by its layout: a histogram and one scatter; no
    # histogram (appended chunks, then their list); or a histogram and two
    # passes, each reading and writing every record
    scat = probe_scatter_label(t.probe_appended, t.probe_passes)
    kern = [("k_hist build", phases["build_hist_ms"], nb * w_in),
            ("k_scatter build", phases["build_partition_ms"] - phases["build_hist_ms"], 2 * nb * w_in)]
    if not t.probe_appended:
        kern.append(("k_hist probe", phases["probe_hist_ms"], npr * w_in))
    kern += [(scat, phases["probe_partition_ms"] - (0.0 if t.probe_appended else phases["probe_hist_ms"]),
              (4 if t.probe_passes == 2 else 2) * npr * w_in),
             ("k_join", phases["join_ms"], b_alg)]
    kernels = []
    for name, kms, kbytes in kern:
        ach = kbytes / (kms * 1e-3) / 1e9 if kms > 0 else None
        tr = traffic.get(name) if traffic else None
        kernels.append({"kernel": name, "ms": kms, "alg_bytes": kbytes, "achieved": ach, "unit": "GB/s",
                        "frac": ach / HBM_PEAK_GBS if ach else None,
                        "traffic": (tr["read"] + tr["write"]) if tr else None})
    dom = max(kernels, key=lambda k: k["ms"])
    dom["dominant"] = True
    step_traffic = sum(k["traffic"] for k in kernels) if traffic and all(k["traffic"] for k in kernels) else None
    if agg_res is not None:
        phases["aggregate_ms"] = agg_res["ms_median"]
    res = {
        "metric": ("join tuples/s (build+probe), 64M x 512M clustered int64 keys, Zipf z=%g fk; %s"
                   % (args.zipf, args.partitioner)) if clustered else
        ("join tuples/s (build+probe) + HBM GB/s, 16M x 256M int64; " + args.partitioner) if not zipf
        else "join tuples/s (build+probe), 8M x 128M (int32, int32) keys, Zipf z=%g; %s" % (args.zipf,
                                                                                             args.partitioner),
        "value": (nb + npr) / (ms * 1e-3),
        "unit": "tuples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32" if zipf else "int64",
        "data": "synthetic (seeded generator, inputs resident in HBM)",
        "config": {"workload": args.workload + ("_z%g" % args.zipf if zipf or clustered else ""), "seed": seed,
                   "build_tuples": nb, "probe_tuples": npr, "probe_misses": nmiss,
                   "tuple_bytes": 16, "out_tuple_bytes": w_out, "out_tuples": n,
                   "partitioner": args.partitioner, "partitions": t.partitions,
                   "subpartitions": t.subpartitions, "work_items": t.work_items,
                   "used_learned": t.used_learned, "spilled_items": t.spilled_items,
                   "probe_layout": "appended" if t.probe_appended else "contiguous",
                   "probe_passes": t.probe_passes},
        "hbm_gbs_algorithmic": job_gbs,
        "hbm_frac_of_peak": job_gbs / HBM_PEAK_GBS,
        "hbm_gbs_inputs_only": (nb + npr) * w_in / (ms * 1e-3) / 1e9,
        "phases_ms": phases,
        "parity": {"count": n, "expected_count": ec, "checksum_ok": parity},
        "roofline": {"bound": "hbm", "kernel": "whole step (partition R, partition S, join)",
                     "achieved": job_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": job_gbs / HBM_PEAK_GBS,
                     "traffic": step_traffic,
                     "traffic_source": ("profiles/traffic_%s.md (rocprofv3 FETCH_SIZE / WRITE_SIZE of this "
                                        "library build)" % args.workload) if step_traffic else
                                       "none for this library build (bench.py --collect-traffic)",
                     "kernels": kernels},
    }
    if load_res is not None:
        res["load_from_files"] = load_res
    if agg_res is not None:
        res["metric"] = res["metric"].replace("join tuples/s", "join + %s tuples/s" % args.aggregate)
        res["config"]["aggregate"] = args.aggregate
        res["aggregate"] = agg_res
    j.close()
    del out
    torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        if clustered:
            res["cpu_baseline"] = cpu_baseline_sample(db, dp, nb, npr, 1 << 26)
        elif zipf:
            res["cpu_baseline"] = cpu_baseline_zipf(seed, nb, npr, cdf)
        else:
            res["cpu_baseline"] = cpu_baseline_full(db, dp, nb, npr, gpu_xor32) or cpu_baseline(seed, dom, nb, npr,
                                                                                               nmiss)
    del db, dp
    torch.cuda.empty_cache()
    if not args.no_compare and args.workload == "uniform_16Mx256M" and args.aggregate == "none" \
            and not args.from_files:
        res["learned_vs_radix"] = compare_learned(args, P, torch, dev, stream, not args.no_cpu_baseline)
        parity = parity and res["learned_vs_radix"]["identical_output"]
    print(json.dumps(res))
    if not parity:
        sys.exit(1)


if __name__ == "__main__":
    main()
