"""GPU: the continuous-batching serving loop (CkFftServer*).

Every request's output must equal, bit for bit, the same transform run alone through the
single-signal drop-in call, whatever it was batched with; admission must respect the HBM
budget (back-pressure, never failure), and invalid or oversized requests are refused.
"""
import threading

import numpy as np
import pytest

import oracle as O
from test_gpu_batch import KINDS, build, plain

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available(), "GPU tests need a HIP device"
    return _t


@pytest.fixture(scope="module")
def C(torch):
    import ckfft_amd
    return ckfft_amd


@pytest.fixture(scope="module")
def ctx(torch, C):
    c = C.Context(1 << 17, C.kCkFftDirection_Both)
    yield c
    c.close()


def workload_mix(count, seed, kmin=6, kmax=16):
    """Sizes and kinds drawn as in the uneven batch workload: k uniform in [kmin, kmax],
    complex/real and forward/inverse 50/50."""
    rng = np.random.default_rng(seed)
    mix = []
    for _ in range(count):
        k = int(rng.integers(kmin, kmax + 1))
        real = rng.random() < 0.5
        fwd = rng.random() < 0.5
        mix.append((("rfwd" if fwd else "rinv") if real else ("cfwd" if fwd else "cinv"), 1 << k))
    return mix


def poisson_arrivals(count, rate_per_s, seed):
    gaps = np.random.default_rng(seed).exponential(1e6 / rate_per_s, count)
    return np.cumsum(gaps)


def assert_alone_equals_mix(torch, ctx, mix, ins, outs):
    for i, (kind, n) in enumerate(mix):
        want = plain(torch, ctx, kind, n, ins[i]).cpu().numpy()
        got = outs[i].cpu().numpy()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (i, kind, n)


def test_serving_alone_equals_mix(torch, C, ctx):
    """Poisson arrivals at 200k requests/s: each result equals the same transform alone."""
    mix = workload_mix(600, 51)
    mix += [("cfwd", 1 << 17), ("rinv", 1 << 17), ("cfwd", 2), ("rfwd", 4)]   # ordinary-launch sizes
    ins, outs, descs = build(torch, C, ctx, mix, 7000)
    torch.cuda.synchronize()
    with C.Server(ctx, 4 << 30, max_in_flight=4) as srv:
        tickets = srv.replay(descs, poisson_arrivals(len(descs), 200e3, 52))
        assert srv.drain() == 1
        assert all(srv.query(t) == 1 for t in tickets)
        st = srv.stats()
        assert st["completed"] == len(mix) and st["failed"] == 0
        assert 0 < st["p50_us"] <= st["p99_us"] <= st["max_us"]
        assert st["batches"] >= 1 and st["mean_batch"] >= 1
    assert_alone_equals_mix(torch, ctx, mix, ins, outs)


def test_serving_tuned_layouts_alone_equals_mix(torch, C, tmp_path):
    """With the 32-points-per-thread layout chosen for every tunable class (a tuning file),
    every class from 2 to 2^16 points served in a Poisson stream equals the single call."""
    path = tmp_path / "all_var1.json"
    keys = [f'"{k}_{(c + 1) if k[0] == "r" else c}": 1' for k in KINDS for c in range(5, 14)]
    path.write_text('{\n "arch": "gfx950",\n "choice": {' + ", ".join(keys) + "}\n}\n")
    mix = [(k, 1 << e) for k in KINDS for e in range(1 if k[0] == "c" else 3, 17)] * 3
    with C.Context(1 << 16) as c:
        assert c.load_tuning(str(path)) == 1
        ins, outs, descs = build(torch, C, c, mix, 9000)
        torch.cuda.synchronize()
        with C.Server(c, 1 << 30, max_in_flight=4) as srv:
            srv.replay(descs, poisson_arrivals(len(descs), 200e3, 54))
            assert srv.drain() == 1
        assert_alone_equals_mix(torch, c, mix, ins, outs)


def test_serving_against_oracle(torch, C, ctx):
    mix = workload_mix(40, 53, kmin=3, kmax=16)
    ins, outs, descs = build(torch, C, ctx, mix, 8000)
    torch.cuda.synchronize()
    with C.Server(ctx, 1 << 30) as srv:
        tickets = [srv.submit(*d) for d in descs]
        assert all(t >= 0 for t in tickets)
        assert all(srv.wait(t) == 1 for t in tickets)
    for i, (kind, n) in enumerate(mix):
        x = ins[i].cpu().numpy()[None]
        o = O.Restatement(n, 3)
        ref = {"cfwd": lambda: o.complex(x, False), "cinv": lambda: o.complex(x, True),
               "rfwd": lambda: o.real_forward(x), "rinv": lambda: o.real_inverse(x, n)}[kind]()
        assert O.split_norm_err(outs[i].cpu().numpy()[None], ref) <= 1e-5, (i, kind, n)


def test_budget_back_pressure(torch, C, ctx):
    """A budget that holds about two of the largest requests: everything still completes,
    in-flight bytes never exceed the budget, and results are unchanged."""
    with C.Server(ctx, 1 << 30, max_in_flight=2) as probe:
        resident = probe.stats()["resident_bytes"]
    big = 16 << 16                                   # complex 2^16: read + write (the intermediate is the slot's)
    mix = workload_mix(200, 54)
    ins, outs, descs = build(torch, C, ctx, mix, 9000)
    torch.cuda.synchronize()
    with C.Server(ctx, resident + 2 * big + 4096, max_in_flight=2) as srv:
        srv.replay(descs, np.zeros(len(descs)))     # all arrive at once
        assert srv.drain() == 1
        st = srv.stats()
        assert st["completed"] == len(mix)
        assert 0 < st["peak_inflight_bytes"] <= st["budget_bytes"] < 3 * big
    assert_alone_equals_mix(torch, ctx, mix, ins, outs)


def test_refusals(torch, C, ctx):
    x = torch.zeros(1 << 18, device="cuda")
    y = torch.zeros(1 << 18, device="cuda")
    h = np.zeros(64, np.float32)
    with C.Server(ctx, 1 << 30) as probe:
        resident = probe.stats()["resident_bytes"]
    with pytest.raises(ValueError):
        C.Server(ctx, resident // 2)                 # below the server's own buffers
    with pytest.raises(ValueError):
        C.Server(ctx, 1 << 30, max_in_flight=0)
    with C.Server(ctx, resident + (16 << 10)) as srv:
        assert srv.submit(0, 1 << 12, x, y) == -1    # 64 KiB alone exceeds a 16 KiB budget
        assert srv.submit(0, 1 << 8, x, y) >= 0
        assert srv.submit(0, 1 << 8, x, x) == -1     # aliased
        assert srv.submit(7, 1 << 8, x, y) == -1     # bad kind
        assert srv.submit(0, 100, x, y) == -1        # not a power of two
        assert srv.submit(0, 1 << 18, x, y) == -1    # above the context's nMax
        assert srv.submit(0, 16, h, y) == -1         # host input
        assert srv.query(12345) == -1
        assert srv.drain() == 1
    with C.Context(1 << 10, C.kCkFftDirection_Forward) as f:
        with C.Server(f, 1 << 28) as srv:
            assert srv.submit(1, 64, x, y) == -1     # inverse on a forward-only context
            assert srv.submit(0, 64, x, y) >= 0
            assert srv.drain() == 1


def test_concurrent_submitters(torch, C, ctx):
    """Four client threads submit at once; every request completes with the alone result."""
    mix = workload_mix(400, 55, kmin=4, kmax=15)
    ins, outs, descs = build(torch, C, ctx, mix, 10000)
    torch.cuda.synchronize()
    with C.Server(ctx, 2 << 30, max_in_flight=3) as srv:
        tickets = [None] * len(descs)

        def client(w):
            for i in range(w, len(descs), 4):
                tickets[i] = srv.submit(*descs[i])

        threads = [threading.Thread(target=client, args=(w,)) for w in range(4)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        assert all(t is not None and t >= 0 for t in tickets)
        assert len(set(tickets)) == len(tickets)
        assert srv.drain() == 1
        assert srv.stats()["completed"] == len(mix)
    assert_alone_equals_mix(torch, ctx, mix, ins, outs)


def test_ticket_ring_wraps(torch, C, ctx):
    """More submissions than the server's ticket ring (2^18): old entries are reused once
    finished, every request completes, old tickets report done."""
    n = 16
    x = torch.from_numpy(np.ascontiguousarray(O.uniform_pm1(56, 2 * n))).cuda()
    y = torch.full((2 * n,), float("nan"), device="cuda")
    want = plain(torch, ctx, "cfwd", n, x).cpu().numpy()
    count = (1 << 18) + 5000
    torch.cuda.synchronize()
    with C.Server(ctx, 1 << 30, max_in_flight=4) as srv:
        tickets = srv.replay([(0, n, x, y)] * count, np.zeros(count))   # same buffers: identical writes
        assert srv.drain() == 1
        st = srv.stats()
        assert st["completed"] == count and st["failed"] == 0
        assert tickets[-1] == count - 1
        assert srv.query(0) == 1 and srv.query(count - 1) == 1 and srv.wait(1) == 1
        assert srv.query(count) == -1                   # not issued yet
    assert np.array_equal(y.cpu().numpy().view(np.uint32), want.view(np.uint32))
