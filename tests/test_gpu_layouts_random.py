"""GPU parity of the three probe-side layouts on seeded random joins: the
contiguous layout after a histogram (PJ_FLAG_LAYOUT_EXACT), the appended
layout (PJ_FLAG_LAYOUT_APPEND) and two radix passes
(PJ_FLAG_LAYOUT_TWO_PASS), each against the CPU oracle's HashJoinOp (sorted
rows, bit for bit), over random sizes (empty sides included), radix bits 1
to 14 and key domains from a few hundred values (many matches a key) to
2^40 (at most one)."""
import sys

import numpy as np
import pytest

import oracle as O
import pjoin as P
from _util import to_gpu

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

ARGS = dict(build_cols=["long", "long"], probe_cols=["long", "long"], buildjattr=[0], probejattr=[0],
            projection=["B$0", "B$1", "P$1"])
FLAGS = {"exact": P.PJ_FLAG_LAYOUT_EXACT, "append": P.PJ_FLAG_LAYOUT_APPEND, "two_pass": P.PJ_FLAG_LAYOUT_TWO_PASS}


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    P.lib()
    yield


def _case(seed):
    rng = np.random.default_rng(seed)
    nb = int(rng.choice([0, 1, 7, int(rng.integers(1, 120000))]))
    npr = int(rng.choice([0, 5, int(rng.integers(1, 400000))]))
    dom = int(rng.choice([300, 1 << 16, 1 << 40]))
    bits = int(rng.integers(1, 15))
    if dom == 300:   # many matches a key: keep the output small
        nb, npr = min(nb, 2000), min(npr, 30000)
    bk = rng.integers(-dom, dom, nb)
    pkeys = rng.integers(-dom, dom, npr)
    if nb and npr:   # half the probes from the build keys
        pkeys[: npr // 2] = rng.choice(bk, npr // 2)
    b = np.stack([bk, np.arange(nb)], 1).astype(np.int64).reshape(-1).view(np.uint8)
    p = np.stack([pkeys, np.arange(npr) + (1 << 40)], 1).astype(np.int64).reshape(-1).view(np.uint8)
    return nb, npr, bits, b, p


@pytest.mark.parametrize("seed", range(40))
def test_layouts_on_random_joins(seed):
    nb, npr, bits, b, p = _case(seed)
    print("case: random join seed=%d nb=%d np=%d bits=%d" % (seed, nb, npr, bits), flush=True)
    sys.stdout.flush()
    d = O.join_desc(["long", "long"], ["long", "long"], [0], [0], ["B$0", "B$1", "P$1"], buckets=1 << 12)
    exp, _ = O.hashjoin(d, b, p)
    exp = O.sorted_rows(exp)
    cap = max(exp.shape[0], 1)
    db = to_gpu(b) if b.size else torch.zeros(16, dtype=torch.uint8, device="cuda")
    dp = to_gpu(p) if p.size else torch.zeros(16, dtype=torch.uint8, device="cuda")
    for name, flags in FLAGS.items():
        j = P.Join(P.make_desc(**ARGS, radix_bits=bits, flags=flags))
        out = torch.empty(cap * j.out_width, dtype=torch.uint8, device="cuda")
        n = j.run_device(db.data_ptr(), nb, dp.data_ptr(), npr, out.data_ptr(), cap,
                         torch.cuda.current_stream().cuda_stream)
        got = out[: n * j.out_width].cpu().numpy().reshape(-1, j.out_width)
        j.close()
        assert got.shape == exp.shape, (name, got.shape, exp.shape)
        assert np.array_equal(O.sorted_rows(got), exp), name
