"""GPU: the opt-in split-BF16x3 precision mode (radix-16 stages on the matrix cores).

Gates: against the FP64 DFT, relative L2 <= 1e-6 (FP32-level error); against the CPU
oracle (the reference's scalar algorithm) the same max-abs gate as the FP32 path; the
mode is deterministic, leaves real transforms and other sizes on the FP32 kernels, and
batch plans / servers built on a context in this mode give the single-call results bit
for bit.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-6
TOL_REF = 1e-5


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
    c = C.Context(1 << 16, C.kCkFftDirection_Both)
    assert c.set_precision(C.kCkFftPrecision_SplitBf16x3) == 1
    yield c
    c.close()


@pytest.fixture(scope="module")
def ctx32(torch, C):
    c = C.Context(1 << 16, C.kCkFftDirection_Both)
    yield c
    c.close()


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to("cuda")


def run(torch, c, inv, n, x):
    xt = dev(torch, x)
    yt = torch.full(xt.shape, float("nan"), dtype=torch.float32, device="cuda")
    fn = c.complex_inverse_batch if inv else c.complex_forward_batch
    assert fn(n, x.shape[0], xt, yt) == 1
    torch.cuda.synchronize()
    return yt.cpu().numpy()


@pytest.mark.parametrize("inv", [False, True])
@pytest.mark.parametrize("k", range(4, 17))
def test_split_bf16_sizes(torch, ctx, inv, k):
    n = 1 << k
    count = max(3, min(37, (1 << 18) // n))          # 37: not a multiple of a group
    x = O.uniform_pm1(7000 + 2 * k + inv, count * 2 * n).reshape(count, 2 * n)
    got = run(torch, ctx, inv, n, x)
    assert np.all(np.isfinite(got))
    e64 = O.rel_l2(got, O.interleave(O.fp64_complex(x, inv)))
    assert e64 <= TOL_FP64, (n, inv, e64)
    sample = x[:3]
    e_ref = O.split_norm_err(got[:3], O.Restatement(n, 3).complex(sample, inv, 8))
    assert e_ref <= TOL_REF, (n, inv, e_ref)


def test_split_bf16_single_call_and_round_trip(torch, ctx):
    """The drop-in single-signal call uses the mode too; forward then inverse returns
    n * x within FP32-level error."""
    n = 4096
    x = O.uniform_pm1(77, 2 * n)
    xt = dev(torch, x)
    y = torch.empty_like(xt)
    z = torch.empty_like(xt)
    assert ctx.complex_forward(n, xt, y) == 1
    assert ctx.complex_inverse(n, y, z) == 1
    torch.cuda.synchronize()
    batch = run(torch, ctx, False, n, x[None])[0]
    assert np.array_equal(y.cpu().numpy().view(np.uint32), batch.view(np.uint32))
    assert O.rel_l2(z.cpu().numpy() / n, x) <= 2e-6


def test_split_bf16_deterministic_and_distinct_from_fp32(torch, ctx, ctx32):
    n = 1 << 12
    x = O.uniform_pm1(78, 16 * 2 * n).reshape(16, 2 * n)
    a = run(torch, ctx, False, n, x)
    b = run(torch, ctx, False, n, x)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    f = run(torch, ctx32, False, n, x)
    assert not np.array_equal(a.view(np.uint32), f.view(np.uint32))   # the mode is really on
    assert O.rel_l2(a, f) <= 2 * TOL_FP64


def test_mode_leaves_real_and_other_sizes_on_fp32(torch, C, ctx, ctx32):
    for n in (1 << 10, 1 << 14):
        x = O.normal01(79, 4 * n).reshape(4, n)
        xt = dev(torch, x)
        ya = torch.empty((4, n + 2), device="cuda")
        yb = torch.empty((4, n + 2), device="cuda")
        assert ctx.real_forward_batch(n, 4, xt, ya) == 1
        assert ctx32.real_forward_batch(n, 4, xt, yb) == 1
        torch.cuda.synchronize()
        assert torch.equal(ya.view(torch.int32), yb.view(torch.int32))
    for n in (2, 8):                                   # below the mode's range
        x = O.uniform_pm1(80, 4 * 2 * n).reshape(4, 2 * n)
        assert np.array_equal(run(torch, ctx, False, n, x).view(np.uint32),
                              run(torch, ctx32, False, n, x).view(np.uint32))


def test_precision_api(C, ctx32):
    with C.Context(1 << 10) as c:
        assert c.precision() == C.kCkFftPrecision_Fp32
        assert c.set_precision(5) == 0
        assert c.set_precision(C.kCkFftPrecision_SplitBf16x3) == 1
        assert c.precision() == C.kCkFftPrecision_SplitBf16x3
        assert c.set_precision(C.kCkFftPrecision_Fp32) == 1
    assert C.CkFftSetPrecision(None, 1) == 0
    assert C.CkFftGetPrecision(None) == -1


def test_plan_and_server_in_split_mode_equal_single_calls(torch, C, ctx):
    from test_gpu_batch import build, plain
    rng = np.random.default_rng(81)
    mix = [(("cfwd", "cinv", "rfwd", "rinv")[int(rng.integers(0, 4))], 1 << int(rng.integers(3, 17)))
           for _ in range(60)]
    ins, outs, descs = build(torch, C, ctx, mix, 12000)
    plan = C.BatchPlan(ctx, descs, graph=True)
    n_split = sum(1 for kind, n in mix if kind in ("cfwd", "cinv") and n >= 16)
    assert plan.info()["extra_launches"] >= n_split
    assert plan.execute() == 1
    torch.cuda.synchronize()
    for i, (kind, n) in enumerate(mix):
        want = plain(torch, ctx, kind, n, ins[i]).cpu().numpy()
        assert np.array_equal(outs[i].cpu().numpy().view(np.uint32), want.view(np.uint32)), (i, kind, n)
    plan.close()
    for o in outs:
        o.fill_(float("nan"))
    torch.cuda.synchronize()
    with C.Server(ctx, 1 << 30) as srv:
        srv.replay(descs, np.arange(len(descs)) * 3.0)
        assert srv.drain() == 1
    for i, (kind, n) in enumerate(mix):
        want = plain(torch, ctx, kind, n, ins[i]).cpu().numpy()
        assert np.array_equal(outs[i].cpu().numpy().view(np.uint32), want.view(np.uint32)), (i, kind, n)
