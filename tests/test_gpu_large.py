"""GPU: the largest transforms, 2^26 .. 2^28 points (complex) or samples (real).

These sizes take pass plans and twiddle levels that no smaller size uses: a radix-2^10
first pass at complex 2^28, radix-2^9 later passes from complex 2^26 up, the radix-2^9
pack pass (real 2^28) and unpack pass (real 2^26), and the fine twiddle tables of levels
2^26 .. 2^28.  The CPU oracles would need minutes per transform here, so the oracle is a
closed form checked on every output value:

  * complex: a sum of tones x[n] = sum_t b_t exp(-+2 pi i f_t n / N) (sign of the
    direction's kernel) has the unnormalised spectrum N b_t at k = f_t and 0 at every
    other bin;
  * real forward: x[n] = sum_t a_t cos(2 pi f_t n / N + p_t) gives 2 * DFT(x)[k] (the
    reference's scale, src/ckfft/fft_real_default.cpp:13-63) = N a_t exp(i p_t) at
    k = f_t and 0 at the other bins 0..N/2;
  * real inverse: a half spectrum with Y[f_t] = c_t and zeros elsewhere gives
    x[n] = sum_t 2 Re(c_t exp(2 pi i f_t n / N)) (the Hermitian extension, unnormalised,
    src/ckfft/fft_real_default.cpp:65-114).

A wrong twiddle anywhere in the transform leaks a tone into other bins (or distorts the
wave), so comparing every output value catches it.  Gates: relative L2 <= 1e-6; the dense
real-inverse output within the parity tests' per-value gate 1e-5 ||ref||_2 / sqrt(len);
for the sparse spectra each tone bin within 1e-6 of its value and every other bin below
1e-6 ||ref||_2 (spectrum_gate says why).  Inputs and expected
values are formed in FP64 on the GPU (exact integer phase reduction), with torch as
plumbing only.  tests/test_oracle.py checks on the CPU that this closed-form check rejects
transforms with a planted twiddle mistake.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_REF = 1e-5
TOL_REL = 1e-6


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
    c = C.Context(1 << 28, C.kCkFftDirection_Both)
    yield c
    c.close()


def tones(seed, n, count, lo, hi):
    """`count` distinct tone bins in [lo, hi) and unit-scale complex amplitudes."""
    rng = np.random.default_rng(seed)
    f = rng.choice(np.arange(lo, hi, dtype=np.int64), size=count, replace=False) if hi - lo < 1 << 20 else \
        np.unique(rng.integers(lo, hi, size=count, dtype=np.int64))
    amp = rng.uniform(0.5, 1.0, f.size) * np.exp(1j * rng.uniform(0, 2 * np.pi, f.size))
    return f, amp


def phase(torch, f, n, sign):
    """sign * 2 pi ((f * i) mod n) / n for i < n, in FP64 on the GPU."""
    i = torch.arange(n, dtype=torch.int64, device="cuda")
    return (sign * 2.0 * math.pi / n) * torch.remainder(i * int(f), n).to(torch.float64)


def gate(d_max, d_sq, ref_sq, length):
    ref_norm = math.sqrt(ref_sq)
    return d_max <= TOL_REF * ref_norm / math.sqrt(length) and math.sqrt(d_sq) <= TOL_REL * ref_norm


def spectrum_gate(torch, y, f, expect):
    """A sparse expected spectrum (values `expect` at bins f, 0 elsewhere): relative L2 error
    <= 1e-6 overall, each tone within 1e-6 of its value, and no other bin above 1e-6
    ||ref||_2.  The parity tests' per-value gate 1e-5 ||ref||_2 / sqrt(len) reads the
    output's RMS as its scale, which fits dense spectra; here the energy sits in a few bins
    and FP32 rounding scales with them (intermediate values grow to N |b| along a tone's
    path, and a 1e-7 rounding there spreads into other bins: ~4 at 2^26 against a tone of
    5e7), so the scale is ||ref||_2 itself.  A wrong twiddle moves a fraction of the energy
    to other bins and fails the L2 gate by orders of magnitude."""
    ref_norm = float(np.sqrt(np.sum(np.abs(expect) ** 2)))
    idx = torch.from_numpy(f).cuda()
    tone_err = (y[idx] - torch.from_numpy(expect).cuda()).abs()
    y[idx] = 0
    rest = y.abs()
    d_sq = float((rest * rest).sum() + (tone_err * tone_err).sum())
    ok_tones = bool((tone_err <= TOL_REL * torch.from_numpy(np.abs(expect)).cuda()).all())
    ok_rest = float(rest.max()) <= TOL_REL * ref_norm
    ok_l2 = math.sqrt(d_sq) <= TOL_REL * ref_norm
    return ok_tones and ok_rest and ok_l2, (float(tone_err.max()), float(rest.max()), math.sqrt(d_sq) / ref_norm)


@pytest.mark.parametrize("kind", ["cfwd", "cinv"])
@pytest.mark.parametrize("k", [26, 27, 28])
def test_complex_tones(torch, ctx, kind, k):
    n = 1 << k
    print(f"large: {kind} n=2^{k}", flush=True)
    f, amp = tones(100 * k + (kind == "cinv"), n, 4, 1, n)
    sign = -1.0 if kind == "cinv" else 1.0      # the tone this direction maps to a single bin
    x = torch.zeros(n, dtype=torch.complex128, device="cuda")
    for ft, a in zip(f, amp):
        x += complex(a) * torch.polar(torch.ones(n, dtype=torch.float64, device="cuda"), phase(torch, ft, n, sign))
    xf = torch.view_as_real(x.to(torch.complex64)).reshape(-1).contiguous()
    del x
    y = torch.empty_like(xf)
    fn = ctx.complex_forward_batch if kind == "cfwd" else ctx.complex_inverse_batch
    assert fn(n, 1, xf, y) == 1
    torch.cuda.synchronize()
    del xf
    yc = torch.view_as_complex(y.view(n, 2)).to(torch.complex128)
    ok, errs = spectrum_gate(torch, yc, f, n * amp)
    assert ok, f"{kind} 2^{k}: tone err {errs[0]:.3e}, other bins max {errs[1]:.3e}, relL2 {errs[2]:.3e}"


@pytest.mark.parametrize("k", [26, 27, 28])
def test_real_forward_tones(torch, ctx, k):
    n = 1 << k
    h = n // 2
    print(f"large: rfwd n=2^{k}", flush=True)
    f, amp = tones(300 + k, n, 4, 1, h)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    for ft, a in zip(f, amp):
        x += abs(a) * torch.cos(phase(torch, ft, n, 1.0) + float(np.angle(a)))
    xf = x.to(torch.float32).contiguous()
    del x
    y = torch.empty(2 * (h + 1), dtype=torch.float32, device="cuda")
    assert ctx.real_forward_batch(n, 1, xf, y) == 1
    torch.cuda.synchronize()
    del xf
    yc = torch.view_as_complex(y.view(h + 1, 2)).to(torch.complex128)
    ok, errs = spectrum_gate(torch, yc, f, n * amp)
    assert ok, f"rfwd 2^{k}: tone err {errs[0]:.3e}, other bins max {errs[1]:.3e}, relL2 {errs[2]:.3e}"


@pytest.mark.parametrize("k", [26, 27, 28])
def test_real_inverse_tones(torch, ctx, k):
    n = 1 << k
    h = n // 2
    print(f"large: rinv n=2^{k}", flush=True)
    f, amp = tones(500 + k, n, 4, 1, h)
    spec = torch.zeros(h + 1, dtype=torch.complex64, device="cuda")
    spec[torch.from_numpy(f).cuda()] = torch.from_numpy(amp.astype(np.complex64)).cuda()
    yf = torch.view_as_real(spec).reshape(-1).contiguous()
    out = torch.empty(n, dtype=torch.float32, device="cuda")
    assert ctx.real_inverse_batch(n, 1, yf, out) == 1
    torch.cuda.synchronize()
    del yf, spec
    ref = torch.zeros(n, dtype=torch.float64, device="cuda")
    for ft, a in zip(f, amp.astype(np.complex64).astype(np.complex128)):
        ref += 2.0 * abs(a) * torch.cos(phase(torch, ft, n, 1.0) + float(np.angle(a)))
    d = (out.to(torch.float64) - ref).abs()
    ok = gate(float(d.max()), float((d * d).sum()), float((ref * ref).sum()), n)
    assert ok, f"rinv 2^{k}: max {float(d.max()):.3e} L2 {float(d.norm()):.3e}"
