"""Failure handling: what a plan does when an input returns Error or the
device cannot hold a buffer, and what the C ABI reports.

Reference behaviour: a build-side child's Error makes HashJoinOp::scanStart
return Error (join.cpp:586-588) before the group barrier, so the thread's
peers wait there for good (join.cpp:545-547); a probe-side Error throws
QueryExecutionError (join.cpp:785-789). Here a failed input thread still
arrives at its group's barrier, the group skips its device work, and every
thread of the group returns Error; the plan call reports how many threads
failed and the first library error one of them saw.

A thread's input is made to fail by the memscan source's failure injection
(errorthread / errorafter). Failures before the group barrier need no GPU:
the device work they cancel is never started."""
import numpy as np
import pytest

import oracle as O
import pjoin as P

HJ = """
buffsize = 64;
b: { type = "memscan"; schema = ( "long", "long" ); %(berr)s };
p: { type = "memscan"; schema = ( "long", "long" ); %(perr)s };
j: { type = "hashjoin"; hash: { fn = "modulo"; buckets = 1024; }; threadgroups = ( %(groups)s );
     tuplesperbucket = 2; allocpolicy = "local"; buildjattr = 0; probejattr = 0;
     projection = ( "B$0", "B$1", "P$1" ); probebatch = 16; };
m: { type = "merge"; threads = 4; };
treeroot: { name = "m"; input: { name = "j"; build: { name = "b"; }; probe: { name = "p"; }; }; };
"""


def err(thread, after):
    return "errorthread = %d; errorafter = %d;" % (thread, after)


def hj_plan(berr="", perr="", groups="[ 0, 1, 2, 3 ]"):
    return HJ % dict(berr=berr, perr=perr, groups=groups)


def hj_inputs(nb=2000, npr=8000, seed=61):
    b, p = O.gen_pkfk(seed, 1 << 40, nb, npr, npr // 10)
    return {"b": (b.view(np.uint8), nb), "p": (p.view(np.uint8), npr)}, b, p


def run_failing(plan, sources):
    with pytest.raises(P.PJError) as e:
        P.run_plan(plan, sources, 1 << 22)
    return e.value


@pytest.mark.timeout(60, method="thread")
@pytest.mark.parametrize("thread,after", [(0, 0), (0, 3), (2, 1), (3, 40)])
def test_build_child_error_fails_every_thread_of_the_group(thread, after):
    """One build thread's child fails mid-build (the leader or another, at
    its first page or later): all four threads return Error, none waits."""
    src, _, _ = hj_inputs()
    e = run_failing(hj_plan(berr=err(thread, after)), src)
    assert e.code == P.PJ_ERR_STATE
    assert e.msg.startswith("4 of 4 threads returned Error"), str(e)


@pytest.mark.timeout(60, method="thread")
def test_build_child_error_fails_only_its_own_group():
    """Two thread groups with a table each: a build failure in group [2, 3]
    fails its two threads; group [0, 1] is not waited on by it (its build
    and probe need the device, so only its threads' count is checked where a
    GPU runs them)."""
    src, _, _ = hj_inputs()
    e = run_failing(hj_plan(berr=err(3, 2), groups="[ 0, 1 ], [ 2, 3 ]"), src)
    assert e.code == P.PJ_ERR_STATE
    n = int(e.msg.split(" of ")[0])
    assert n in (2, 4), str(e)   # 4 without a GPU: group [0, 1] then fails at its device build


PART = """
buffsize = 256;
s: { type = "memscan"; schema = ( "long", "long" ); %(err)s };
part: { type = "partition"; attr = 0; range = [ 1, 1000 ]; buckets = 4; maxtuples = 100000; sort = "no"; };
m: { type = "merge"; threads = 4; };
treeroot: { name = "m"; input: { name = "part"; input: { name = "s"; }; }; };
"""


@pytest.mark.timeout(60, method="thread")
@pytest.mark.parametrize("thread,after", [(0, 1), (3, 0)])
def test_partition_child_error_fails_every_thread(thread, after):
    rng = np.random.default_rng(5)
    v = rng.integers(1, 1001, 20000).astype(np.int64)
    data = np.stack([v, v], 1).reshape(-1)
    e = run_failing(PART % dict(err=err(thread, after)), {"s": (data.view(np.uint8), 20000)})
    assert e.code == P.PJ_ERR_STATE
    assert e.msg.startswith("4 of 4 threads returned Error"), str(e)


SMJ = """
buffsize = 256;
build: { type = "memscan"; schema = ( "long", "long" ); %(berr)s };
probe: { type = "memscan"; schema = ( "long", "long" ); %(perr)s };
join: { type = "sortmergejoin"; threadgroups = ( [0, 1, 2, 3] ); maxbuildtuples = 2000; maxprobetuples = 8000;
        buildjattr = 0; probejattr = 0; projection = ( "B$0", "P$1" ); };
merge: { type = "merge"; threads = 4; };
treeroot: { name = "merge"; input: { name = "join"; build: { name = "build"; }; probe: { name = "probe"; }; }; };
"""


@pytest.mark.timeout(60, method="thread")
@pytest.mark.parametrize("side", ["berr", "perr"])
def test_sortmerge_input_error_fails_every_thread(side):
    src, b, p = hj_inputs()
    plan = SMJ % dict(dict(berr="", perr=""), **{side: err(1, 2)})
    e = run_failing(plan, {"build": src["b"], "probe": src["p"]})
    assert e.code == P.PJ_ERR_STATE
    assert e.msg.startswith("4 of 4 threads returned Error"), str(e)


AGG = """
buffsize = 256;
s: { type = "memscan"; schema = ( "long", "long" ); %(err)s };
agg: { type = "aggregate_count"; field = 0; hash: { fn = "modulo"; buckets = 1024; field = 0; };
       global = "yes"; threads = 4; };
merge: { type = "merge"; threads = 4; };
treeroot: { name = "merge"; input: { name = "agg"; input: { name = "s"; }; }; };
"""


@pytest.mark.timeout(60, method="thread")
def test_global_aggregate_input_error_fails_every_thread():
    rng = np.random.default_rng(6)
    data = rng.integers(0, 500, (20000, 2)).astype(np.int64).reshape(-1)
    e = run_failing(AGG % dict(err=err(2, 5)), {"s": (data.view(np.uint8), 20000)})
    assert e.code == P.PJ_ERR_STATE
    assert e.msg.startswith("4 of 4 threads returned Error"), str(e)


def test_device_limit_calls_need_no_gpu():
    before = P.device_bytes()
    P.set_device_limit(1 << 20)
    P.set_device_limit(0)
    assert P.device_bytes() == before


# ----------------------------------------------------------------- GPU


@pytest.mark.gpu
@pytest.mark.timeout(120, method="thread")
def test_probe_side_error_mid_probe_is_reported():
    """A probe child's Error after some pages throws QueryExecutionError in
    the reference (join.cpp:785-789); here the plan call fails with that
    message."""
    src, _, _ = hj_inputs()
    e = run_failing(hj_plan(perr=err(1, 3)), src)
    assert e.code == P.PJ_ERR_STATE
    assert "probe side returned Error" in str(e), str(e)


@pytest.mark.gpu
@pytest.mark.timeout(120, method="thread")
def test_probe_side_error_at_first_page_fails_that_thread():
    """An Error on a thread's first probe page, read inside scanStart: that
    thread returns Error; the table is whole, so the others finish."""
    src, _, _ = hj_inputs()
    e = run_failing(hj_plan(perr=err(2, 0)), src)
    assert e.code == P.PJ_ERR_STATE
    assert e.msg.startswith("1 of 4 threads returned Error"), str(e)


@pytest.mark.gpu
@pytest.mark.timeout(120, method="thread")
def test_device_allocation_failure_is_an_error_of_every_thread():
    """A device buffer the library cannot have (a limit far below what the
    build needs) fails the group's build with PJ_ERR_DEVICE's message, every
    thread returns Error, and the same plan runs once the limit is gone."""
    src, b, p = hj_inputs(20000, 80000)
    plan = hj_plan()
    P.set_device_limit(64 << 10)
    try:
        e = run_failing(plan, src)
    finally:
        P.set_device_limit(0)
    assert e.code == P.PJ_ERR_STATE
    assert e.msg.startswith("4 of 4 threads returned Error: "), str(e)
    assert "would pass the limit" in str(e), str(e)
    rows = P.run_plan(plan, src, 80000 * 24)
    d = O.join_desc(["long", "long"], ["long", "long"], [0], [0], ["B$0", "B$1", "P$1"], buckets=1024)
    exp, _ = O.hashjoin(d, b.view(np.uint8), p.view(np.uint8))
    assert np.array_equal(O.sorted_rows(rows), O.sorted_rows(exp))


@pytest.mark.gpu
def test_device_allocation_failure_through_the_abi():
    """pj_join_run_device under the limit: PJ_ERR_DEVICE; under no limit the
    same handle joins correctly; its buffers go back on close."""
    torch = pytest.importorskip("torch")
    from _util import to_gpu
    b, p = O.gen_pkfk(62, 1 << 40, 50000, 200000, 0)
    db, dp = to_gpu(b.view(np.uint8)), to_gpu(p.view(np.uint8))
    torch.cuda.synchronize()
    before = P.device_bytes()
    j = P.Join(P.make_desc(["long", "long"], ["long", "long"], [0], [0], ["B$0", "B$1", "P$1"]))
    out = torch.empty(200000 * j.out_width, dtype=torch.uint8, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    P.set_device_limit(before + (256 << 10))
    try:
        with pytest.raises(P.PJError) as e:
            j.run_device(db.data_ptr(), 50000, dp.data_ptr(), 200000, out.data_ptr(), 200000, st)
    finally:
        P.set_device_limit(0)
    assert e.value.code == P.PJ_ERR_DEVICE and "would pass the limit" in e.value.msg
    n = j.run_device(db.data_ptr(), 50000, dp.data_ptr(), 200000, out.data_ptr(), 200000, st)
    assert n == 200000
    rows = out[: n * j.out_width].cpu().numpy().reshape(-1, j.out_width)
    d = O.join_desc(["long", "long"], ["long", "long"], [0], [0], ["B$0", "B$1", "P$1"], buckets=1 << 12)
    exp, _ = O.hashjoin(d, b.view(np.uint8), p.view(np.uint8))
    assert np.array_equal(O.sorted_rows(rows), O.sorted_rows(exp))
    assert P.device_bytes() > before
    j.close()
    assert P.device_bytes() == before
