"""GPU parity: the HIP path through the C ABI against the CPU oracles.

Gates (per transform, worst over the batch):
  * vs the bit-exact CPU restatement of the reference (oracle.Restatement):
        max|gpu - ref| <= 1e-5 * ||ref||_2 / sqrt(len(ref))
    (the output-norm reading of "max abs diff <= 1e-5 ||x||_2 / sqrt(N)";
    ckfft does not normalise, so the outputs carry the scale);
  * vs the FP64 oracle: relative L2 error <= 1e-6 over the whole batch.
Closed-form sizes (complex n <= 2, real n <= 4) must match the reference bit for bit.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TOL_REF = 1e-5
TOL_FP64 = 1e-6
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


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
    yield c
    c.close()


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to("cuda")


def out_len(kind, n):
    return {"cfwd": 2 * n, "cinv": 2 * n, "rfwd": 2 * (n // 2 + 1), "rinv": n}[kind]


def run_batch(torch, ctx, kind, n, x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    count = x.shape[0]
    xt = dev(torch, x)
    yt = torch.full((count, out_len(kind, n)), float("nan"), dtype=torch.float32, device="cuda")
    fn = {"cfwd": ctx.complex_forward_batch, "cinv": ctx.complex_inverse_batch,
          "rfwd": ctx.real_forward_batch, "rinv": ctx.real_inverse_batch}[kind]
    assert fn(n, count, xt, yt) == 1
    torch.cuda.synchronize()
    return yt.cpu().numpy()


def oracle_out(kind, n, x, threads=8):
    o = O.Restatement(max(n, 1), 3)
    if kind in ("cfwd", "cinv"):
        return o.complex(x, kind == "cinv", threads)
    if kind == "rfwd":
        return o.real_forward(x, threads)
    return o.real_inverse(x, n, threads)


def fp64_out(kind, n, x):
    if kind in ("cfwd", "cinv"):
        return O.interleave(O.fp64_complex(x, kind == "cinv"))
    if kind == "rfwd":
        return O.interleave(O.fp64_real_forward(x))
    return O.fp64_real_inverse(x, n)


def make_input(kind, n, count, seed):
    if kind in ("cfwd", "cinv"):
        return O.uniform_pm1(seed, count * 2 * n).reshape(count, 2 * n)
    if kind == "rfwd":
        return O.normal01(seed, count * n).reshape(count, n)
    # real inverse: half spectra of real signals (what callers pass), as the oracle makes them
    x = O.normal01(seed, count * n).reshape(count, n)
    return O.Restatement(n, 3).real_forward(x)


def check(kind, n, got, x, exact_small=True):
    ref = oracle_out(kind, n, x)
    assert np.all(np.isfinite(got)), (kind, n)
    if exact_small and ((kind in ("cfwd", "cinv") and n <= 2) or (kind in ("rfwd", "rinv") and n <= 4)):
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (kind, n)
        return
    e_ref = O.split_norm_err(got, ref)
    e64 = O.rel_l2(got, fp64_out(kind, n, x))
    assert e_ref <= TOL_REF, (kind, n, e_ref)
    assert e64 <= TOL_FP64, (kind, n, e64)


# ------------------------------------------------------------------------- twiddles

def test_twiddle_table_is_fp64_rounded(ctx, C):
    for lg in range(0, 17):
        L = 1 << lg
        buf = np.zeros(2 * L, np.float32)
        assert ctx.twiddles(L, buf) == 1
        k = np.arange(L)
        exact = O.exact_twiddle(k, L, -1)
        want = O.interleave(exact).astype(np.float32)
        err = np.abs(buf.astype(np.float64) - O.interleave(exact))
        # correctly rounded FP64 -> FP32: within half an FP32 ulp of the exact value
        assert float(err.max()) <= 2.0 ** -25, (L, float(err.max()))
        assert np.mean(buf != want) < 1e-4, L
        if L >= 4:
            assert buf[2 * (L // 4)] == 0.0 and buf[2 * (L // 4) + 1] == -1.0


def test_fine_and_transposed_twiddle_tables(torch, C):
    """The tables above the pyramid, read directly from a caller's arena (nMax = 2^28): the
    fine tables (levels 2^17 .. 2^28, entries m < 2^12) within half an FP32 ulp of the
    exact exp(-2 pi i m / L), and the transposed inter-pass table, table[r * 256 + j] =
    pyramid level 2^16 entry r j mod 2^16, bit for bit.  The arena holds the pyramid (2 *
    2^16 entries), then the fine tables, then the transposed table, each starting on 256
    bytes (table_layout in ckfft_amd/csrc/ckfft_api.cpp; the sizes are checked against
    CkFftDeviceTableBytes here)."""
    n = 1 << 28
    log_pyr, log_fine, nfine = 16, 12, 28 - 16
    pyr_bytes = 8 * (2 << log_pyr)
    fine_bytes = 8 * (nfine << log_fine)
    xtw_bytes = 8 * (1 << log_pyr)
    need = C.CkFftDeviceTableBytes(n, C.kCkFftDirection_Both)
    assert need == pyr_bytes + fine_bytes + xtw_bytes, need
    arena = torch.zeros(need, dtype=torch.uint8, device="cuda")
    with C.Context(n, C.kCkFftDirection_Both, arena=arena) as ac:
        torch.cuda.synchronize()
        tab = arena.cpu().numpy().view(np.float32)
        # the pyramid sits at offset 0: level 2^16 as CkFftGetTwiddles reports it
        L = 1 << log_pyr
        lvl = np.zeros(2 * L, np.float32)
        assert ac.twiddles(L, lvl) == 1
        assert np.array_equal(tab[2 * L:4 * L].view(np.uint32), lvl.view(np.uint32))
        fine = tab[pyr_bytes // 4:(pyr_bytes + fine_bytes) // 4].reshape(nfine, 2 << log_fine)
        m = np.arange(1 << log_fine)
        for i in range(nfine):
            Lf = 1 << (log_pyr + 1 + i)
            exact = O.interleave(O.exact_twiddle(m, Lf, -1))
            err = np.abs(fine[i].astype(np.float64) - exact)
            assert float(err.max()) <= 2.0 ** -25, (Lf, float(err.max()))
            assert np.mean(fine[i] != exact.astype(np.float32)) < 1e-3, Lf
        xtw = tab[(pyr_bytes + fine_bytes) // 4:].view(np.uint64).reshape(256, 256)   # one complex per uint64
        r = np.arange(256)[:, None]
        j = np.arange(256)[None, :]
        want = lvl.view(np.uint64)[(r * j) % L]
        assert np.array_equal(xtw, want)


# ------------------------------------------------------------------------- all sizes

@pytest.mark.parametrize("kind", ["cfwd", "cinv"])
@pytest.mark.parametrize("k", range(0, 16))
def test_complex_sizes(torch, ctx, kind, k):
    n = 1 << k
    count = max(3, min(257, (1 << 16) // n))
    x = make_input(kind, n, count, 1000 + 10 * k)
    check(kind, n, run_batch(torch, ctx, kind, n, x), x)


@pytest.mark.parametrize("k", range(0, 17))
def test_real_forward_sizes(torch, ctx, k):
    n = 1 << k
    count = max(3, min(131, (1 << 17) // n))
    x = make_input("rfwd", n, count, 2000 + k)
    check("rfwd", n, run_batch(torch, ctx, "rfwd", n, x), x)


@pytest.mark.parametrize("k", range(0, 17))
def test_real_inverse_sizes(torch, ctx, k):
    n = 1 << k
    count = max(3, min(131, (1 << 17) // n))
    x = make_input("rinv", n, count, 3000 + k)
    check("rinv", n, run_batch(torch, ctx, "rinv", n, x), x)


@pytest.mark.parametrize("k", [3, 4, 5, 9, 12, 15, 16])
def test_real_inverse_arbitrary_spectra(torch, ctx, k):
    """Inputs that are not half spectra of real data: Im Y[0] and Im Y[N/2] enter the
    result exactly as in the reference's pre-pass."""
    n = 1 << k
    x = O.uniform_pm1(4000 + k, 3 * 2 * (n // 2 + 1)).reshape(3, -1)
    check("rinv", n, run_batch(torch, ctx, "rinv", n, x), x)


# ------------------------------------------------------------------------- API behaviour

@pytest.mark.parametrize("kind,k", [("cfwd", 0), ("cinv", 1), ("cfwd", 2), ("cinv", 6), ("cfwd", 12),
                                    ("rfwd", 2), ("rfwd", 3), ("rinv", 8), ("rfwd", 13), ("rinv", 16),
                                    ("cfwd", 15)])
def test_plain_call_equals_batch_bitwise(torch, ctx, kind, k):
    """A signal's result does not depend on the batch it is in (same kernel, same bits)."""
    n = 1 << k
    x = make_input(kind, n, 5, 5000 + k)
    batch = run_batch(torch, ctx, kind, n, x)
    xt = dev(torch, x)
    for b in range(x.shape[0]):
        y = torch.full((out_len(kind, n),), float("nan"), dtype=torch.float32, device="cuda")
        if kind == "cfwd":
            ok = ctx.complex_forward(n, xt[b], y)
        elif kind == "cinv":
            ok = ctx.complex_inverse(n, xt[b], y)
        elif kind == "rfwd":
            ok = ctx.real_forward(n, xt[b], y)
        else:
            tmp = torch.empty(2 * (n // 2 + 1), dtype=torch.float32, device="cuda")
            ok = ctx.real_inverse(n, xt[b], y, tmp)
        assert ok == 1
        assert np.array_equal(y.cpu().numpy().view(np.uint32), batch[b].view(np.uint32)), b


@pytest.mark.parametrize("kind,k", [("cfwd", 10), ("cinv", 4), ("rfwd", 11), ("rinv", 12), ("rinv", 2)])
def test_host_pointers_are_staged(torch, ctx, kind, k):
    """Host buffers (the reference's only kind) work and give the device-path bits."""
    n = 1 << k
    x = make_input(kind, n, 1, 6000 + k)
    want = run_batch(torch, ctx, kind, n, x)[0]
    xh = np.ascontiguousarray(x[0])
    y = np.full(out_len(kind, n), np.nan, np.float32)
    if kind == "cfwd":
        ok = ctx.complex_forward(n, xh, y)
    elif kind == "cinv":
        ok = ctx.complex_inverse(n, xh, y)
    elif kind == "rfwd":
        ok = ctx.real_forward(n, xh, y)
    else:
        ok = ctx.real_inverse(n, xh, y, np.zeros(2 * (n // 2 + 1), np.float32))
    assert ok == 1
    assert np.array_equal(y.view(np.uint32), want.view(np.uint32))


def host_call(ctx, kind, n, x, y):
    if kind == "cfwd":
        return ctx.complex_forward(n, x, y)
    if kind == "cinv":
        return ctx.complex_inverse(n, x, y)
    if kind == "rfwd":
        return ctx.real_forward(n, x, y)
    return ctx.real_inverse(n, x, y, np.zeros(2 * (n // 2 + 1), np.float32))


@pytest.mark.parametrize("kind,k", [("cfwd", 17), ("rinv", 20), ("cinv", 21)])
def test_host_pointers_large_chunked(torch, ctx_big, kind, k):
    """Host buffers of several MiB go through the chunked pinned pipeline both ways."""
    n = 1 << k
    x = make_input(kind, n, 1, 6100 + k)
    want = run_batch(torch, ctx_big, kind, n, x)[0]
    y = np.full(out_len(kind, n), np.nan, np.float32)
    assert host_call(ctx_big, kind, n, np.ascontiguousarray(x[0]), y) == 1
    assert np.array_equal(y.view(np.uint32), want.view(np.uint32))


def test_pinned_and_mixed_buffers(torch, ctx):
    """Pinned host input (DMA'd directly), device output, and the other way round."""
    n = 1 << 12
    x = make_input("cfwd", n, 1, 6200)
    want = run_batch(torch, ctx, "cfwd", n, x)[0]
    xp = torch.from_numpy(np.ascontiguousarray(x[0])).pin_memory()
    yd = torch.full((2 * n,), float("nan"), device="cuda")
    assert ctx.complex_forward(n, xp, yd) == 1
    assert np.array_equal(yd.cpu().numpy().view(np.uint32), want.view(np.uint32))
    xd = torch.from_numpy(np.ascontiguousarray(x[0])).cuda()
    yp = torch.full((2 * n,), float("nan")).pin_memory()
    assert ctx.complex_forward(n, xd, yp) == 1
    assert np.array_equal(yp.numpy().view(np.uint32), want.view(np.uint32))


def test_concurrent_host_pointer_calls(torch, ctx):
    """Eight threads sharing one context (allowed by the reference's header) each run
    host-pointer calls; every result equals the device-path bits."""
    import threading
    cases = [("cfwd", 1 << 14), ("rfwd", 1 << 16), ("rinv", 1 << 13), ("cinv", 1 << 10)]
    inputs = {c: make_input(c[0], c[1], 1, 6300 + i) for i, c in enumerate(cases)}
    wants = {c: run_batch(torch, ctx, c[0], c[1], inputs[c])[0] for c in cases}
    errors = []

    def worker(w):
        for rep in range(6):
            kind, n = cases[(w + rep) % len(cases)]
            y = np.full(out_len(kind, n), np.nan, np.float32)
            ok = host_call(ctx, kind, n, np.ascontiguousarray(inputs[(kind, n)][0]), y)
            if ok != 1 or not np.array_equal(y.view(np.uint32), wants[(kind, n)].view(np.uint32)):
                errors.append((w, rep, kind, n, ok))

    threads = [threading.Thread(target=worker, args=(w,)) for w in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors[:5]


def test_concurrent_device_pointer_calls_overlap(torch, ctx):
    """Eight threads sharing one context call the drop-in entry points on device buffers.
    Each call runs on its calling thread's own default stream, so calls from different
    threads overlap on the device (on the null stream they would queue behind each other):
    eight threads finish 8 x 40 complex 2^15 transforms (one workgroup each) in well under
    the time one thread needs for the same 320 calls.  Every result equals the batched
    call's bits."""
    import threading
    import time
    n, per, nthr = 1 << 15, 40, 8
    x = make_input("cfwd", n, nthr, 6400)
    xt = dev(torch, x)
    want = run_batch(torch, ctx, "cfwd", n, x)
    outs = [torch.full((2 * n,), float("nan"), device="cuda") for _ in range(nthr)]
    torch.cuda.synchronize()

    def worker(w, calls, res):
        ok = True
        for i in range(calls):
            k = (w + i) % nthr if calls > per else w
            ok &= ctx.complex_forward(n, xt[k], outs[k]) == 1
        res[w] = ok

    res = [None] * nthr
    worker(0, nthr, res)                       # warm-up: every input once
    t0 = time.perf_counter()
    worker(0, nthr * per, res)
    serial = time.perf_counter() - t0
    assert res[0]
    t0 = time.perf_counter()
    threads = [threading.Thread(target=worker, args=(w, per, res)) for w in range(nthr)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    concurrent = time.perf_counter() - t0
    assert all(res), res
    for w in range(nthr):
        assert np.array_equal(outs[w].cpu().numpy().view(np.uint32), want[w].view(np.uint32))
    print(f"320 calls: one thread {serial * 1e3:.2f} ms, eight threads {concurrent * 1e3:.2f} ms")
    assert concurrent < 0.7 * serial, (serial, concurrent)


def test_device_arena_context(torch, C):
    """Tables in a caller's device arena: same bits as a library-allocated context; the
    arena is left to the caller at shutdown and can hold a new context; bad arenas refused."""
    assert C.CkFftDeviceTableBytes(1024, C.kCkFftDirection_Both) == 2 * 1024 * 8
    assert C.CkFftDeviceTableBytes(1 << 16, C.kCkFftDirection_Both) == (2 << 16) * 8 + (1 << 16) * 8
    assert C.CkFftDeviceTableBytes(1 << 20, C.kCkFftDirection_Both) == (2 << 16) * 8 + 4 * 4096 * 8 + (1 << 16) * 8
    assert C.CkFftDeviceTableBytes(1000, C.kCkFftDirection_Both) == 0
    assert C.CkFftDeviceTableBytes(1024, 7) == 0
    n = 1 << 16
    need = C.CkFftDeviceTableBytes(n, C.kCkFftDirection_Both)
    arena = torch.zeros(need + 512, dtype=torch.uint8, device="cuda")
    x = make_input("cfwd", n, 2, 6400)
    r = make_input("rfwd", n, 2, 6401)
    with C.Context(n) as plain_ctx:
        want_c = run_batch(torch, plain_ctx, "cfwd", n, x)
        want_r = run_batch(torch, plain_ctx, "rfwd", n, r)
    for rep in range(2):                                # the arena is reused after shutdown
        with C.Context(n, C.kCkFftDirection_Both, arena=arena[:need]) as ac:
            assert np.array_equal(run_batch(torch, ac, "cfwd", n, x).view(np.uint32), want_c.view(np.uint32))
            assert np.array_equal(run_batch(torch, ac, "rfwd", n, r).view(np.uint32), want_r.view(np.uint32))
    for bad in (arena[: need - 256], arena[1:need + 1], torch.zeros(need, dtype=torch.uint8)):
        with pytest.raises(RuntimeError):
            C.Context(n, C.kCkFftDirection_Both, arena=bad)
    size = ctypes.c_size_t(0)
    assert not C.CkFftInitArena(n, C.kCkFftDirection_Both, None, ctypes.byref(size), None, 0)
    assert size.value > 0                               # query mode reports the host part


def test_device_arena_multipass_plan_and_server(torch, C):
    """An arena context at nMax = 2^20 (pyramid, fine tables and the transposed pass table
    all in the arena): multi-pass transforms of every kind, a batch plan and a server give
    the bits of a library-allocated context."""
    n = 1 << 20
    need = C.CkFftDeviceTableBytes(n, C.kCkFftDirection_Both)
    arena = torch.zeros(need, dtype=torch.uint8, device="cuda")
    cases = [(k, n >> s) for k in ("cfwd", "cinv", "rfwd", "rinv") for s in (0, 3)]
    ins = {c: make_input(c[0], c[1], 1, 6500 + i) for i, c in enumerate(cases)}
    with C.Context(n) as lib_ctx:
        want = {c: run_batch(torch, lib_ctx, c[0], c[1], ins[c]) for c in cases}
    with C.Context(n, C.kCkFftDirection_Both, arena=arena) as ac:
        for c in cases:
            assert np.array_equal(run_batch(torch, ac, c[0], c[1], ins[c]).view(np.uint32),
                                  want[c].view(np.uint32)), c
        kinds = ("cfwd", "cinv", "rfwd", "rinv")
        xs = [dev(torch, ins[c][0]) for c in cases]
        for route in ("plan", "server"):
            ys = [torch.full((out_len(k, m),), float("nan"), dtype=torch.float32, device="cuda") for k, m in cases]
            descs = [(kinds.index(k), m, xs[i], ys[i]) for i, (k, m) in enumerate(cases)]
            if route == "plan":
                plan = C.BatchPlan(ac, descs)
                assert plan.execute() == 1
                plan.close()
            else:
                with C.Server(ac, 1 << 30) as srv:
                    assert all(srv.wait(srv.submit(*d)) == 1 for d in descs)
            torch.cuda.synchronize()
            for i, c in enumerate(cases):
                assert np.array_equal(ys[i].cpu().numpy().view(np.uint32), want[c][0].view(np.uint32)), (route, c)


def test_max_count_does_not_change_results(torch, C):
    """Like the reference (float twiddle angle scaling by 2^j is exact), results are
    identical for any nMax >= n."""
    for kind, n in (("cfwd", 1024), ("rfwd", 4096), ("rinv", 256), ("cinv", 8)):
        x = make_input(kind, n, 2, 7000 + n)
        outs = []
        for m in (n, 2 * n, 1 << 16):
            with C.Context(m) as c:
                outs.append(run_batch(torch, c, kind, n, x))
        assert all(np.array_equal(o.view(np.uint32), outs[0].view(np.uint32)) for o in outs), (kind, n)


def test_argument_validation_on_device(torch, C):
    """src/ckfft/ckfft.cpp:36-114: 0 for every invalid call, 1 otherwise."""
    x = torch.zeros(64, dtype=torch.float32, device="cuda")
    y = torch.zeros(64, dtype=torch.float32, device="cuda")
    t = torch.zeros(64, dtype=torch.float32, device="cuda")
    with C.Context(16, C.kCkFftDirection_Forward) as f, C.Context(16, C.kCkFftDirection_Inverse) as i:
        assert f.complex_forward(16, x, y) == 1
        assert f.complex_inverse(16, x, y) == 0            # no inverse direction
        assert i.complex_inverse(16, x, y) == 1
        assert i.complex_forward(16, x, y) == 0
        assert f.real_forward(16, x, y) == 1
        assert f.real_inverse(16, x, y, t) == 0
        assert i.real_inverse(16, x, y, t) == 1
        assert i.real_inverse(16, x, y, None) == 0          # tmpBuf NULL
        assert f.complex_forward(32, x, y) == 0             # n > nMax
        assert f.complex_forward(12, x, y) == 0             # not a power of two
        assert f.complex_forward(0, x, y) == 0
        assert f.complex_forward(-2147483648, x, y) == 0    # INT_MIN is rejected here
        assert f.complex_forward(16, x, x) == 0             # in == out
        assert f.complex_forward(16, None, y) == 0
        assert f.real_forward(16, x, None) == 0
        assert f.complex_forward_batch(16, -1, x, y) == 0
        assert f.complex_forward_batch(16, 0, x, y) == 1
        assert f.complex_forward_batch(16, 2, x, y) == 1


def test_caller_buffer_context(torch, C):
    """Query mode, then a context placed in caller memory (src/ckfft/context.cpp:47-67)."""
    size = ctypes.c_size_t(0)
    assert not C.CkFftInit(1024, 3, None, ctypes.byref(size))
    buf = ctypes.create_string_buffer(size.value)
    h = C.CkFftInit(1024, 3, buf, ctypes.byref(size))
    assert h == ctypes.addressof(buf)
    x = make_input("cfwd", 1024, 1, 11)
    xt = dev(torch, x)
    y = torch.empty_like(xt)
    assert C.CkFftComplexForwardBatch(h, 1024, 1, xt.data_ptr(), y.data_ptr(), None) == 1
    torch.cuda.synchronize()
    check("cfwd", 1024, y.cpu().numpy(), x)
    C.CkFftShutdown(h)                 # does not free the caller's buffer
    assert ctypes.sizeof(buf) == size.value


# ------------------------------------------------------------------------- reference fixtures

def test_golden_vectors_on_gpu(torch, ctx):
    """Against the genuine reference's outputs on its own input (src/test/input.txt)."""
    g = dict(np.load(os.path.join(GOLDEN, "golden_small.npz")))
    inp = O.reference_input_txt()
    for k in range(0, 13):
        n = 1 << k
        x = inp[: 2 * n][None]
        for kind in ("cfwd", "cinv"):
            got = run_batch(torch, ctx, kind, n, x)[0]
            assert O.split_norm_err(got[None], g[f"{kind}_{n}"][None]) <= TOL_REF, (kind, n)
        xr = np.ascontiguousarray(x[:, 0::2])
        got = run_batch(torch, ctx, "rfwd", n, xr)[0]
        assert O.split_norm_err(got[None], g[f"rfwd_{n}"][None]) <= TOL_REF, n
        got = run_batch(torch, ctx, "rinv", n, g[f"rfwd_{n}"][None])[0]
        assert O.split_norm_err(got[None], g[f"rinv_{n}"][None]) <= TOL_REF, n


def test_reference_regression_checks_on_gpu(torch, ctx):
    """src/test/test.cpp:754-864 on the GPU path: RMS vs an independent FFT <= 1e-3, and
    real forward * 0.5 == complex forward of the same data with imag = 0."""
    inp = O.reference_input_txt()
    for k in range(0, 13):
        n = 1 << k
        x = inp[: 2 * n][None]
        for kind in ("cfwd", "cinv"):
            got = run_batch(torch, ctx, kind, n, x)[0]
            want = fp64_out(kind, n, x)[0]
            assert float(np.sqrt(np.mean((got - want) ** 2))) <= 1e-3
        xr = np.ascontiguousarray(x[:, 0::2])
        xc = np.zeros((1, 2 * n), np.float32)
        xc[:, 0::2] = xr
        c = run_batch(torch, ctx, "cfwd", n, xc)[0][: 2 * (n // 2 + 1)]
        r = run_batch(torch, ctx, "rfwd", n, xr)[0] * 0.5
        assert float(np.sqrt(np.mean((r - c) ** 2))) <= 1e-3


def test_example_program_on_gpu(torch, C):
    """src/example/main.cpp:44-85 through the drop-in calls with host buffers."""
    ex = np.fromfile(os.path.join(GOLDEN, "example_in.f32"), dtype="<f4")
    with C.Context(1024) as c:
        y = np.zeros(2 * 513, np.float32)
        assert c.real_forward(1024, ex, y) == 1
        x = np.zeros(1024, np.float32)
        assert c.real_inverse(1024, y, x, np.zeros(2 * 513, np.float32)) == 1
    assert float(np.sum((x / 2048.0 - ex) ** 2)) < 1e-9


# ------------------------------------------------------------------------- workload configs

def test_config_complex_4096_x1024_round_trip(torch, ctx):
    n, count = 4096, 1024
    x = O.uniform_pm1(0xC0FFEE, count * 2 * n).reshape(count, 2 * n)
    y = run_batch(torch, ctx, "cfwd", n, x)
    check("cfwd", n, y, x)
    z = run_batch(torch, ctx, "cinv", n, y)
    check("cinv", n, z, y)
    assert np.max(np.abs(z / n - x)) < 1e-5


def test_config_real_65536_x256_round_trip(torch, ctx):
    n, count = 65536, 256
    x = O.normal01(0xBEEF, count * n).reshape(count, n)
    y = run_batch(torch, ctx, "rfwd", n, x)
    check("rfwd", n, y, x)
    z = run_batch(torch, ctx, "rinv", n, y)
    check("rinv", n, z, y)
    assert np.max(np.abs(z / (2 * n) - x)) < 1e-4


# ------------------------------------------------------------------------- multi-pass sizes

@pytest.fixture(scope="module")
def ctx_big(torch, C):
    c = C.Context(1 << 25, C.kCkFftDirection_Both)
    yield c
    c.close()


@pytest.mark.parametrize("kind", ["cfwd", "cinv"])
@pytest.mark.parametrize("k", range(16, 26))
def test_complex_multipass_sizes(torch, ctx_big, kind, k):
    """2 passes for 2^16..2^22 (radix splits 2^8 x 2^8 .. 2^11 x 2^11), 3 passes from 2^23."""
    n = 1 << k
    count = 2 if k <= 20 else 1
    x = make_input(kind, n, count, 8000 + 10 * k)
    check(kind, n, run_batch(torch, ctx_big, kind, n, x), x)


@pytest.mark.parametrize("kind", ["rfwd", "rinv"])
@pytest.mark.parametrize("k", [17, 18, 19, 20, 21, 22, 23, 24, 25])
def test_real_multipass_sizes(torch, ctx_big, kind, k):
    """Half-size complex passes with the pack fused into the last pass (forward) or the
    unpack into the first (inverse): 2 passes up to 2^23 real samples, 3 from 2^24 (where
    an intermediate goes through the output buffer)."""
    n = 1 << k
    count = 3 if k <= 18 else 1
    x = make_input(kind, n, count, 9000 + k)
    check(kind, n, run_batch(torch, ctx_big, kind, n, x), x)


def test_multipass_plain_equals_batch(torch, ctx_big):
    n = 1 << 17
    x = make_input("cfwd", n, 3, 9100)
    batch = run_batch(torch, ctx_big, "cfwd", n, x)
    xt = dev(torch, x)
    y = torch.empty(2 * n, dtype=torch.float32, device="cuda")
    assert ctx_big.complex_forward(n, xt[1], y) == 1
    assert np.array_equal(y.cpu().numpy().view(np.uint32), batch[1].view(np.uint32))


def test_multipass_max_count_independent(torch, C):
    n = 1 << 18
    x = make_input("rfwd", n, 1, 9200)
    outs = []
    for m in (n, 1 << 22):
        with C.Context(m) as c:
            outs.append(run_batch(torch, c, "rfwd", n, x))
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


def test_config_single_complex_2_24(torch, ctx_big):
    """One complex transform of 2^24 points (128 MiB in, 128 MiB out), forward and inverse."""
    n = 1 << 24
    x = O.uniform_pm1(0x2424, 2 * n)[None]
    y = run_batch(torch, ctx_big, "cfwd", n, x)
    check("cfwd", n, y, x)
    z = run_batch(torch, ctx_big, "cinv", n, y)
    assert np.max(np.abs(z / n - x)) < 1e-5
