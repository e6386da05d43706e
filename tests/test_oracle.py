"""Pin the CPU oracles (no GPU needed).

What pins what:
  * Restatement (oracle/ckfft_oracle.c) == the genuine reference bit for bit:
    golden vectors written from the reference build (tests/golden/), and oracle/_ref
    itself when it is built in this checkout.
  * FP64 oracle == an O(N^2) float64 DFT for N <= 512 (all four kinds), and == the
    restatement within FP32 rounding for arbitrary inputs (which pins the real-inverse
    map, including Im Y[0] and Im Y[N/2], against the reference's own arithmetic).
  * Closed forms for N <= 4; round-trip identities (complex N*x, real 2N*x); the
    reference's own regression checks (kiss-style RMS <= 1e-3 on src/test/input.txt,
    real == 2 * complex of the same data); the example program's 2N round trip.
  * Planted mistakes (a conjugated twiddle, a transposed real pair, a dropped factor 2)
    in copies of the FP64 oracle are caught by these checks.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def golden():
    return dict(np.load(os.path.join(GOLDEN, "golden_small.npz")))


@pytest.fixture(scope="module")
def digests():
    with open(os.path.join(GOLDEN, "golden_digests.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def inp():
    x = O.reference_input_txt()
    assert x is not None and x.size == 8192
    return x


# ---------------------------------------------------------------- restatement vs reference

@pytest.mark.parametrize("max_mult", [1, 2])
def test_restatement_matches_golden_small(golden, inp, max_mult):
    for k in range(0, 13):
        n = 1 << k
        o = O.Restatement(n * max_mult, 3)
        x = inp[: 2 * n]
        assert np.array_equal(bits(o.complex(x, False)), bits(golden[f"cfwd_{n}"])), n
        assert np.array_equal(bits(o.complex(x, True)), bits(golden[f"cinv_{n}"])), n
        xr = np.ascontiguousarray(x[0::2])
        yr = o.real_forward(xr)
        assert np.array_equal(bits(yr), bits(golden[f"rfwd_{n}"])), n
        assert np.array_equal(bits(o.real_inverse(yr, n)), bits(golden[f"rinv_{n}"])), n


def test_restatement_matches_golden_digests(digests):
    for key, d in digests["outputs"].items():
        kind, n = key.split("_")
        n = int(n)
        o = O.Restatement(n, 3)
        if kind in ("cfwd", "cinv"):
            x = O.uniform_pm1(d["seed"], 2 * n)
            y = o.complex(x, kind == "cinv")
        else:
            xr = O.normal01(d["seed"], n)
            y = o.real_forward(xr)
            if kind == "rinv":
                y = o.real_inverse(y, n)
        flat = bits(y).ravel()
        assert flat[d["idx"]].tolist() == d["bits"], key
        assert hashlib.sha256(np.ascontiguousarray(y).tobytes()).hexdigest() == d["sha256"], key


def test_restatement_twiddles_match_reference_tables(digests):
    for m_s, d in digests["twiddles"].items():
        m = int(m_s)
        if m > (1 << 20):
            continue  # the 2^21..2^24 digests are checked by test_twiddle_digests_large
        f, i = O.Restatement(m, 3).twiddles()
        assert hashlib.sha256(f.tobytes()).hexdigest() == d["fwd_sha256"], m
        assert hashlib.sha256(i.tobytes()).hexdigest() == d["inv_sha256"], m


def test_twiddle_digests_large(digests):
    m = 1 << 24
    f, i = O.Restatement(m, 3).twiddles()
    assert hashlib.sha256(f.tobytes()).hexdigest() == digests["twiddles"][str(m)]["fwd_sha256"]
    assert hashlib.sha256(i.tobytes()).hexdigest() == digests["twiddles"][str(m)]["inv_sha256"]


@pytest.mark.skipif(not O.Reference.available(), reason="oracle/_ref not built in this checkout")
def test_restatement_matches_reference_build_random():
    R = O.Reference()
    for m_log in (3, 10, 16):
        M = 1 << m_log
        o = O.Restatement(M, 3)
        ctx = R.init(M, 3)
        try:
            for k in range(0, m_log + 1):
                n = 1 << k
                x = O.uniform_pm1(77 + k, 2 * n * 2).reshape(2, 2 * n)
                for inv in (False, True):
                    assert np.array_equal(bits(o.complex(x, inv)), bits(R.complex(ctx, x, inv)))
                xr = O.normal01(88 + k, 2 * n).reshape(2, n)
                assert np.array_equal(bits(o.real_forward(xr)), bits(R.real_forward(ctx, xr)))
                y = O.uniform_pm1(99 + k, 2 * 2 * (n // 2 + 1)).reshape(2, -1)
                assert np.array_equal(bits(o.real_inverse(y, n)), bits(R.real_inverse(ctx, y, n)))
        finally:
            R.shutdown(ctx)


def test_restatement_argument_validation():
    """src/ckfft/ckfft.cpp:14-114: 0 on invalid arguments, NULL context on bad init."""
    lib = O.Restatement(16, 3).lib
    assert not lib.ocf_create(0, 3)
    assert not lib.ocf_create(12, 3)
    assert not lib.ocf_create(16, 4)
    fwd_only = O.Restatement(16, 1)
    x = np.zeros(32, np.float32)
    y = np.zeros(32, np.float32)
    fp = O._fptr
    assert fwd_only.lib.ocf_complex(fwd_only.ctx, 16, fp(x), fp(y), 0) == 1
    assert fwd_only.lib.ocf_complex(fwd_only.ctx, 16, fp(x), fp(y), 1) == 0   # no inverse table
    assert fwd_only.lib.ocf_complex(fwd_only.ctx, 32, fp(x), fp(y), 0) == 0   # n > nMax
    assert fwd_only.lib.ocf_complex(fwd_only.ctx, 12, fp(x), fp(y), 0) == 0   # not a power of two
    assert fwd_only.lib.ocf_complex(fwd_only.ctx, 16, fp(x), fp(x), 0) == 0   # in == out
    both = O.Restatement(16, 3)
    t = np.zeros(32, np.float32)
    assert both.lib.ocf_real_inverse(both.ctx, 16, fp(x), fp(y), None) == 0   # tmp NULL


# ---------------------------------------------------------------- FP64 oracle vs brute force

@pytest.mark.parametrize("k", range(0, 10))
def test_fp64_complex_vs_brute_force(k):
    n = 1 << k
    z = O.uniform_pm1(5 + k, 2 * n).astype(np.float64)
    zc = z[0::2] + 1j * z[1::2]
    np.testing.assert_allclose(O.fp64_complex(z, False), O.dft_brute(zc, -1), atol=1e-12 * n)
    np.testing.assert_allclose(O.fp64_complex(z, True), O.dft_brute(zc, +1), atol=1e-12 * n)


@pytest.mark.parametrize("k", range(0, 10))
def test_fp64_real_vs_brute_force(k):
    n = 1 << k
    x = O.normal01(7 + k, n).astype(np.float64)
    full = O.dft_brute(x.astype(np.complex128), -1)
    y = O.fp64_real_forward(x)
    np.testing.assert_allclose(y, 2 * full[: n // 2 + 1], atol=1e-11 * n)
    # inverse of a Hermitian half spectrum = brute-force inverse DFT of the full spectrum
    if n >= 2:
        yh = full[: n // 2 + 1].copy()
        yh[0] = yh[0].real
        yh[-1] = yh[-1].real
        ext = np.concatenate([yh, np.conj(yh[1:-1][::-1])])
        want = O.dft_brute(ext, +1).real
        np.testing.assert_allclose(O.fp64_real_inverse(yh, n), want, atol=1e-10 * n)


@pytest.mark.parametrize("k", [3, 4, 5, 8, 11, 14])
def test_fp64_real_inverse_arbitrary_input_matches_reference_arithmetic(k):
    """For inputs that are not a valid half spectrum (Im Y[0], Im Y[N/2] != 0) the map
    is fixed only by the reference's pre-pass; the restatement follows it bit for bit."""
    n = 1 << k
    y = O.uniform_pm1(123 + k, 4 * (n // 2 + 1)).reshape(2, -1)
    got = O.Restatement(n, 3).real_inverse(y, n)
    want = O.fp64_real_inverse(y, n)
    assert O.split_norm_err(got, want) < 2e-6


@pytest.mark.parametrize("k", [2, 5, 8, 12, 16])
def test_fp64_vs_restatement_within_fp32(k):
    n = 1 << k
    o = O.Restatement(n, 3)
    x = O.uniform_pm1(31 + k, 2 * n)
    for inv in (False, True):
        got = o.complex(x, inv)
        want = O.interleave(O.fp64_complex(x, inv))
        assert O.rel_l2(got, want) < 2e-6
    xr = O.normal01(41 + k, n)
    got = o.real_forward(xr)
    assert O.rel_l2(got, O.interleave(O.fp64_real_forward(xr))) < 2e-6


def test_closed_forms_small_n():
    """n <= 4 (src/ckfft/fft.cpp:20-31, src/ckfft/fft_real.cpp:20-93) are the textbook
    DFT_1/2/4, with the real forward scaled by 2."""
    o = O.Restatement(4, 3)
    x = np.array([0.5, -1.25, 2.0, 0.75], np.float32)
    for n in (1, 2, 4):
        got = o.real_forward(x[:n])
        want = O.interleave(2 * np.fft.rfft(x[:n].astype(np.float64)))
        assert np.array_equal(got, want.astype(np.float32)), n
    z = np.array([1.0, 2.0, -3.0, 0.5], np.float32)
    assert np.array_equal(o.complex(z, False), np.array([-2.0, 2.5, 4.0, 1.5], np.float32))


def test_round_trip_identities():
    o = O.Restatement(1 << 12, 3)
    for k in (3, 7, 12):
        n = 1 << k
        x = O.uniform_pm1(900 + k, 2 * n)
        back = o.complex(o.complex(x, False), True) / n
        assert np.max(np.abs(back - x)) < 1e-5
        xr = O.normal01(901 + k, n)
        back = o.real_inverse(o.real_forward(xr), n) / (2 * n)
        assert np.max(np.abs(back - xr)) < 1e-5


def test_reference_regression_checks(inp):
    """src/test/test.cpp:754-864: complex RMS vs an independent FFT <= 1e-3, and the
    real forward equal to twice the complex transform of the same (imag = 0) data."""
    def rms(a, b):
        return float(np.sqrt(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2)))

    for k in range(0, 13):
        n = 1 << k
        o = O.Restatement(2 * n, 3)
        x = inp[: 2 * n]
        for inv in (False, True):
            assert rms(o.complex(x, inv), O.interleave(O.fp64_complex(x, inv))) <= 1e-3
        xr = np.ascontiguousarray(x[0::2])
        xc = np.zeros(2 * n, np.float32)
        xc[0::2] = xr
        ref = o.complex(xc, False)[: 2 * (n // 2 + 1)]
        assert rms(0.5 * o.real_forward(xr), ref) <= 1e-3


def test_example_program_round_trip(golden):
    """src/example/main.cpp:44-85: 1024 reals forward then inverse, / 2048 == input."""
    ex = np.fromfile(os.path.join(GOLDEN, "example_in.f32"), dtype="<f4")
    o = O.Restatement(1024, 3)
    y = o.real_forward(ex)
    assert np.array_equal(bits(y), bits(golden["ex_rfwd"]))
    x = o.real_inverse(y, 1024)
    assert np.array_equal(bits(x), bits(golden["ex_rinv"]))
    assert float(np.sum((x / 2048.0 - ex) ** 2)) < 1e-9


# ---------------------------------------------------------------- planted mistakes

def test_planted_mistakes_are_caught(monkeypatch):
    """Each plausible slip in the FP64 oracle must fail one of the checks above."""
    n = 64
    x = O.normal01(5, n).astype(np.float64)
    full = O.dft_brute(x.astype(np.complex128), -1)
    good = O.fp64_real_forward(x)
    assert np.allclose(good, 2 * full[: n // 2 + 1])
    assert not np.allclose(np.conj(good), 2 * full[: n // 2 + 1])           # conjugated result
    assert not np.allclose(good / 2, 2 * full[: n // 2 + 1])                # dropped factor 2

    y = O.uniform_pm1(6, 2 * (n // 2 + 1))
    rest = O.Restatement(n, 3).real_inverse(y, n)
    real_twiddle = O.exact_twiddle
    monkeypatch.setattr(O, "exact_twiddle", lambda k, m, sign=-1: real_twiddle(k, m, -sign))
    assert O.split_norm_err(rest, O.fp64_real_inverse(y, n)) > 1e-2         # wrong twiddle sign
    monkeypatch.setattr(O, "exact_twiddle", real_twiddle)
    Yc = O._to_c128(y)
    swapped = O.interleave(Yc[::-1]).ravel()
    assert O.split_norm_err(rest, O.fp64_real_inverse(swapped, n)) > 1e-2  # transposed k / N-k


def _radix2_fft(x, plant=None):
    """Iterative radix-2 decimation-in-frequency FFT in complex64 (forward), used only to
    show that the closed-form tone check of tests/test_gpu_large.py rejects a transform
    with a twiddle mistake.  plant: None, or (stage, kind) with kind "sign" (conjugated
    twiddles in that stage) or "index" (twiddle index off by one in that stage)."""
    n = x.size
    a = x.astype(np.complex64).copy()
    half, stage = n // 2, 0
    while half >= 1:
        j = np.arange(half)
        idx = j * (n // (2 * half))
        if plant == (stage, "index"):
            idx = (idx + 1) % n
        w = np.exp(-2j * np.pi * idx / n).astype(np.complex64)
        if plant == (stage, "sign"):
            w = np.conj(w)
        a = a.reshape(-1, 2, half)
        top, bot = a[:, 0, :].copy(), a[:, 1, :].copy()
        a[:, 0, :] = top + bot
        a[:, 1, :] = (top - bot) * w
        a = a.reshape(n)
        half //= 2
        stage += 1
    # bit reversal
    bits = int(np.log2(n))
    rev = np.array([int(f"{i:0{bits}b}"[::-1], 2) for i in range(n)])
    return a[rev]


def _tone_check(y, n, f, amp):
    """The closed-form gate of tests/test_gpu_large.py (spectrum_gate) for a complex
    forward transform of a sum of tones."""
    d = y.astype(np.complex128).copy()
    expect = n * amp
    ref_norm = np.sqrt(np.sum(np.abs(expect) ** 2))
    tone_err = np.abs(d[f] - expect)
    d[f] = 0
    return (np.all(tone_err <= 1e-6 * np.abs(expect))
            and np.max(np.abs(d)) <= 1e-6 * ref_norm
            and np.sqrt(np.sum(np.abs(d) ** 2) + np.sum(tone_err ** 2)) <= 1e-6 * ref_norm)


def test_tone_check_rejects_planted_twiddle_mistakes():
    """The large-size GPU tests check transforms of a sum of tones against its closed-form
    spectrum on every bin; a correct FP32 FFT passes, and one wrong twiddle sign or index
    in any single stage fails."""
    n = 1 << 12
    rng = np.random.default_rng(11)
    f = rng.choice(np.arange(1, n), size=4, replace=False)
    amp = rng.uniform(0.5, 1.0, 4) * np.exp(1j * rng.uniform(0, 2 * np.pi, 4))
    t = np.arange(n)
    x = sum(a * np.exp(2j * np.pi * ((fi * t) % n) / n) for fi, a in zip(f, amp))
    assert np.allclose(_radix2_fft(x), np.fft.fft(x.astype(np.complex64)), atol=1e-2)
    assert _tone_check(_radix2_fft(x), n, f, amp)
    for stage in (0, 5, 10):
        for kind in ("sign", "index"):
            assert not _tone_check(_radix2_fft(x, (stage, kind)), n, f, amp), (stage, kind)
