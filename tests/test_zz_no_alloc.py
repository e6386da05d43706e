"""GPU: CkFftTransformBatchNoAlloc, the batched call that allocates nothing.

Its multi-pass sizes take their intermediate from the library's fixed per-GPU scratch
array instead of a stream-ordered allocation.  It runs the same kernels as the batched
calls (CkFft*Batch), so every output must equal theirs bit for bit.  Each case makes two
calls in a row on the shared scratch, the second with a smaller input, so that a stale
intermediate left by the first call would show in the second's output.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

KIND = {"cfwd": 0, "cinv": 1, "rfwd": 2, "rinv": 3}


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
    c = C.Context(1 << 22, C.kCkFftDirection_Both)
    yield c
    c.close()


def in_len(kind, n):
    return {"cfwd": 2 * n, "cinv": 2 * n, "rfwd": n, "rinv": 2 * (n // 2 + 1)}[kind]


def out_len(kind, n):
    return {"cfwd": 2 * n, "cinv": 2 * n, "rfwd": 2 * (n // 2 + 1), "rinv": n}[kind]


def scratch_bytes(kind, n, count):
    """The intermediate the multi-pass route needs: 8n bytes per complex signal above
    2^15 points, 4n per real signal above 2^16 samples, 0 for the one-pass sizes."""
    if kind in ("cfwd", "cinv"):
        return 8 * n * count if n > (1 << 15) else 0
    return 4 * n * count if n > (1 << 16) else 0


def make_input(torch, kind, n, count, seed):
    if kind == "rfwd":
        x = O.normal01(seed, count * n)
    else:
        # complex inputs, and real-inverse half spectra (any values: Im Y[0] and
        # Im Y[n/2] enter the transform as in the reference)
        x = O.uniform_pm1(seed, count * in_len(kind, n))
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32).reshape(count, in_len(kind, n))).cuda()


def batched(torch, ctx, kind, n, count, xt, stream=None):
    y = torch.full((count, out_len(kind, n)), float("nan"), dtype=torch.float32, device="cuda")
    fn = {"cfwd": ctx.complex_forward_batch, "cinv": ctx.complex_inverse_batch,
          "rfwd": ctx.real_forward_batch, "rinv": ctx.real_inverse_batch}[kind]
    assert fn(n, count, xt, y, stream) == 1
    return y


def noalloc(torch, ctx, kind, n, count, xt, stream=None):
    y = torch.full((count, out_len(kind, n)), float("nan"), dtype=torch.float32, device="cuda")
    ok = ctx.transform_batch_noalloc(KIND[kind], n, count, xt, y, stream)
    assert ok == 1, f"no-alloc {kind} n={n} x{count} refused"
    return y


# (kind, first (n, count), second (n, count)); every first call fills or nearly fills the
# 64 MiB scratch, every second one is smaller
CASES = [
    ("cfwd", (1 << 20, 8), (1 << 17, 3)),     # 64 MiB exactly, then 3 MiB
    ("cinv", (1 << 21, 2), (1 << 16, 5)),     # 32 MiB, then 2.5 MiB
    ("rfwd", (1 << 22, 4), (1 << 18, 2)),     # 64 MiB, then 2 MiB
    ("rinv", (1 << 21, 3), (1 << 17, 1)),     # 24 MiB, then 0.5 MiB
    ("cfwd", (1 << 15, 64), (1 << 12, 16)),   # one-pass sizes: no scratch used
]


@pytest.mark.parametrize("kind,first,second", CASES, ids=[f"{c[0]}-{c[1][0]}x{c[1][1]}" for c in CASES])
def test_noalloc_equals_batched(torch, C, ctx, kind, first, second):
    for seed, (n, count) in enumerate((first, second)):
        assert scratch_bytes(kind, n, count) <= C.NOALLOC_SCRATCH_BYTES
        print(f"no-alloc {kind} n={n} x{count} (scratch {scratch_bytes(kind, n, count)} B)", flush=True)
        xt = make_input(torch, kind, n, count, 100 + seed)
        ref = batched(torch, ctx, kind, n, count, xt)
        got = noalloc(torch, ctx, kind, n, count, xt)
        torch.cuda.synchronize()
        r, g = ref.cpu().numpy(), got.cpu().numpy()
        assert np.isfinite(r).all()
        assert np.array_equal(r.view(np.uint32), g.view(np.uint32)), \
            f"{kind} n={n} x{count}: no-alloc output differs from the batched call"


def test_noalloc_on_a_stream_follows_new_inputs(torch, C, ctx):
    """Two calls in a row on one non-default stream, the second smaller, each compared
    with the batched call on the same stream."""
    s = torch.cuda.Stream()
    outs = []
    for seed, (n, count) in enumerate(((1 << 19, 16), (1 << 18, 4))):
        xt = make_input(torch, "cfwd", n, count, 200 + seed)
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            ref = batched(torch, ctx, "cfwd", n, count, xt, s)
            got = noalloc(torch, ctx, "cfwd", n, count, xt, s)
        outs.append((ref, got))
    s.synchronize()
    for ref, got in outs:
        assert np.array_equal(ref.cpu().numpy().view(np.uint32), got.cpu().numpy().view(np.uint32))


def test_noalloc_matches_the_oracle(torch, C, ctx):
    """A multi-pass size through the scratch, against the CPU restatement's gate."""
    n, count = 1 << 17, 2
    xt = make_input(torch, "cfwd", n, count, 300)
    got = noalloc(torch, ctx, "cfwd", n, count, xt)
    torch.cuda.synchronize()
    ref = O.Restatement(n, 3).complex(xt.cpu().numpy(), False, 8)
    assert O.split_norm_err(got.cpu().numpy(), ref) <= 1e-5


def test_noalloc_rejects_what_does_not_fit(torch, C, ctx):
    n = 1 << 20
    count = C.NOALLOC_SCRATCH_BYTES // (8 * n) + 1            # one signal over the scratch
    x = torch.zeros((count, 2 * n), dtype=torch.float32, device="cuda")
    y = torch.zeros_like(x)
    assert ctx.transform_batch_noalloc(KIND["cfwd"], n, count, x, y) == 0
    assert b"exceeds the fixed scratch" in C.CkFftGpuLastError()
    # invalid arguments, as in the batched calls
    assert ctx.transform_batch_noalloc(7, n, 1, x, y) == 0            # unknown kind
    assert ctx.transform_batch_noalloc(KIND["cfwd"], 3 * n, 1, x, y) == 0   # not a power of two
    assert ctx.transform_batch_noalloc(KIND["cfwd"], n, 1, x, x) == 0       # in == out
    assert ctx.transform_batch_noalloc(KIND["cfwd"], n, 0, x, y) == 1       # nothing to do
    with C.Context(1 << 10, C.kCkFftDirection_Forward) as fwd_only:
        a = torch.zeros(2 * 1024, dtype=torch.float32, device="cuda")
        b = torch.zeros_like(a)
        assert fwd_only.transform_batch_noalloc(KIND["cinv"], 1024, 1, a, b) == 0
        assert fwd_only.transform_batch_noalloc(KIND["cfwd"], 1024, 1, a, b) == 1
    torch.cuda.synchronize()
