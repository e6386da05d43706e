"""CPU oracles for ckfft's transform path -- TEST INFRASTRUCTURE ONLY.

Nothing in the product (the ``ckfft_amd`` package and its HIP library) imports this
package.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker or the timed CPU baseline.

Three oracles live here:

* ``Restatement`` -- ctypes binding of ``libckfft_oracle.so`` (``ckfft_oracle.c``), a
  bit-exact FP32 restatement of the reference's scalar path
  (src/ckfft/fft_default.cpp:12-266, src/ckfft/fft_real_default.cpp:13-114,
  src/ckfft/context.cpp:90-105).
* ``Reference`` -- ctypes binding of ``_ref/libckfft_ref.so``: the genuine reference built
  from its own sources by ``oracle/Makefile`` (target ``ref``).  Present only where the
  reference checkout was available at build time.
* FP64 functions ``fp64_*`` -- the same linear maps evaluated in float64 with exact
  twiddles (numpy's pocketfft in complex128), pinned against ``dft_brute`` (an O(N^2)
  DFT in float64 with exactly reduced angles).

Conventions (inc/ckfft/ckfft.h:72-81,123-125; src/test/test.cpp:813-818):
complex forward X[k] = sum x[n] e^{-2 pi i k n / N}; complex inverse uses e^{+...};
neither is normalised.  Real forward returns 2*DFT(x)[0..N/2].  Real inverse of Y is
the linear map derived in ``fp64_real_inverse``; for Y = RealForward(x) it returns 2N*x.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libckfft_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libckfft_ref.so")

_c_float_p = ctypes.POINTER(ctypes.c_float)


def _fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_c_float_p)


def as_interleaved(z: np.ndarray) -> np.ndarray:
    """complex64 array -> float32 interleaved (re, im) view copy."""
    return np.ascontiguousarray(z, dtype=np.complex64).view(np.float32)


class Restatement:
    """Bit-exact FP32 CPU restatement of ckfft (see ckfft_oracle.h)."""

    def __init__(self, max_n: int, dirs: int = 3, so_path: str = ORACLE_SO):
        if not os.path.exists(so_path):
            raise FileNotFoundError(f"{so_path} not built (run `make -C oracle`)")
        lib = ctypes.CDLL(so_path)
        lib.ocf_create.restype = ctypes.c_void_p
        lib.ocf_create.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.ocf_destroy.argtypes = [ctypes.c_void_p]
        lib.ocf_twiddles.argtypes = [ctypes.c_int, _c_float_p, _c_float_p]
        for name in ("ocf_complex",):
            getattr(lib, name).argtypes = [ctypes.c_void_p, ctypes.c_int, _c_float_p, _c_float_p, ctypes.c_int]
        lib.ocf_real_forward.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_float_p, _c_float_p]
        lib.ocf_real_inverse.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_float_p, _c_float_p, _c_float_p]
        lib.ocf_complex_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _c_float_p,
                                          _c_float_p, ctypes.c_int, ctypes.c_int]
        lib.ocf_real_forward_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _c_float_p,
                                               _c_float_p, ctypes.c_int]
        lib.ocf_real_inverse_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _c_float_p,
                                               _c_float_p, ctypes.c_int]
        for name in ("ocf_complex", "ocf_real_forward", "ocf_real_inverse", "ocf_complex_batch",
                     "ocf_real_forward_batch", "ocf_real_inverse_batch"):
            getattr(lib, name).restype = ctypes.c_int
        self.lib = lib
        self.max_n = max_n
        self.ctx = lib.ocf_create(max_n, dirs)
        if not self.ctx:
            raise ValueError(f"ocf_create({max_n}, {dirs}) failed")

    def close(self):
        if self.ctx:
            self.lib.ocf_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def twiddles(self):
        f = np.zeros(2 * self.max_n, np.float32)
        i = np.zeros(2 * self.max_n, np.float32)
        self.lib.ocf_twiddles(self.max_n, _fptr(f), _fptr(i))
        return f, i

    # Arrays: complex data as float32 interleaved, shape (..., 2n); batch on axis 0.
    def complex(self, x: np.ndarray, inverse: bool, threads: int = 1) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        n = x.shape[-1] // 2
        count = x.size // (2 * n)
        y = np.empty_like(x)
        ok = self.lib.ocf_complex_batch(self.ctx, n, count, _fptr(x), _fptr(y), int(inverse), threads)
        if not ok:
            raise ValueError("oracle rejected arguments")
        return y

    def real_forward(self, x: np.ndarray, threads: int = 1) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        n = x.shape[-1]
        count = x.size // n
        y = np.empty(x.shape[:-1] + (2 * (n // 2 + 1),), np.float32)
        if not self.lib.ocf_real_forward_batch(self.ctx, n, count, _fptr(x), _fptr(y), threads):
            raise ValueError("oracle rejected arguments")
        return y

    def real_inverse(self, y: np.ndarray, n: int, threads: int = 1) -> np.ndarray:
        y = np.ascontiguousarray(y, dtype=np.float32)
        assert y.shape[-1] == 2 * (n // 2 + 1)
        count = y.size // y.shape[-1]
        x = np.empty(y.shape[:-1] + (n,), np.float32)
        if not self.lib.ocf_real_inverse_batch(self.ctx, n, count, _fptr(y), _fptr(x), threads):
            raise ValueError("oracle rejected arguments")
        return x


class Reference:
    """The genuine reference library (oracle/_ref), driven through its own C API."""

    def __init__(self, so_path: str = REF_SO):
        if not os.path.exists(so_path):
            raise FileNotFoundError(so_path)
        lib = ctypes.CDLL(so_path)
        lib.CkFftInit.restype = ctypes.c_void_p
        lib.CkFftInit.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]
        lib.CkFftShutdown.argtypes = [ctypes.c_void_p]
        lib.CkFftComplexForward.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        lib.CkFftComplexInverse.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        lib.CkFftRealForward.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        lib.CkFftRealInverse.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p]
        for f in ("CkFftComplexForward", "CkFftComplexInverse", "CkFftRealForward", "CkFftRealInverse"):
            getattr(lib, f).restype = ctypes.c_int
        self.lib = lib

    @staticmethod
    def available(so_path: str = REF_SO) -> bool:
        return os.path.exists(so_path)

    def run_threads(self, ctx, kinds, ns, ins, outs, threads: int, passes: int = 1) -> None:
        """`passes` passes over the transforms (kinds[i]: 0 complex forward, 1 complex
        inverse, 2 real forward, 3 real inverse; ns[i]; numpy buffers ins[i] -> outs[i]),
        transform i on POSIX thread i % threads, every thread sharing `ctx`
        (oracle/ref_threads.c, built beside the reference library)."""
        if not hasattr(self, "_thr"):
            thr = ctypes.CDLL(os.path.join(os.path.dirname(REF_SO), "libckfft_ref_threads.so"))
            thr.refthr_run.restype = ctypes.c_int
            thr.refthr_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
            self._thr = thr
        count = len(ns)
        k = np.ascontiguousarray(kinds, dtype=np.int32)
        nn = np.ascontiguousarray(ns, dtype=np.int32)
        pi = np.array([a.ctypes.data for a in ins], dtype=np.uint64)
        po = np.array([a.ctypes.data for a in outs], dtype=np.uint64)
        if not self._thr.refthr_run(ctx, count, k.ctypes.data, nn.ctypes.data, pi.ctypes.data, po.ctypes.data,
                                    int(threads), int(passes)):
            raise ValueError("reference rejected a transform, or a thread could not start")

    def init(self, max_n: int, direction: int = 3):
        return self.lib.CkFftInit(max_n, direction, None, None)

    def shutdown(self, ctx):
        self.lib.CkFftShutdown(ctx)

    def complex(self, ctx, x: np.ndarray, inverse: bool) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        n = x.shape[-1] // 2
        y = np.empty_like(x)
        fn = self.lib.CkFftComplexInverse if inverse else self.lib.CkFftComplexForward
        xs, ys = x.reshape(-1, 2 * n), y.reshape(-1, 2 * n)
        for b in range(xs.shape[0]):
            if not fn(ctx, n, xs[b].ctypes.data, ys[b].ctypes.data):
                raise ValueError("reference rejected arguments")
        return y

    def real_forward(self, ctx, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        n = x.shape[-1]
        # n == 4 writes one element past n/2+1 (src/ckfft/fft_real.cpp:46-47): give it room
        width = 2 * max(n // 2 + 1, 4 if n == 4 else 0)
        xs = x.reshape(-1, n)
        y = np.zeros((xs.shape[0], width), np.float32)
        for b in range(xs.shape[0]):
            if not self.lib.CkFftRealForward(ctx, n, xs[b].ctypes.data, y[b].ctypes.data):
                raise ValueError("reference rejected arguments")
        return y[:, : 2 * (n // 2 + 1)].reshape(x.shape[:-1] + (2 * (n // 2 + 1),)).copy()

    def real_inverse(self, ctx, y: np.ndarray, n: int) -> np.ndarray:
        y = np.ascontiguousarray(y, dtype=np.float32)
        ys = y.reshape(-1, 2 * (n // 2 + 1))
        x = np.empty((ys.shape[0], n), np.float32)
        tmp = np.empty(2 * (n // 2 + 1), np.float32)
        for b in range(ys.shape[0]):
            if not self.lib.CkFftRealInverse(ctx, n, ys[b].ctypes.data, x[b].ctypes.data, tmp.ctypes.data):
                raise ValueError("reference rejected arguments")
        return x.reshape(y.shape[:-1] + (n,))


# ---------------------------------------------------------------------------
# FP64 oracle
# ---------------------------------------------------------------------------

def _to_c128(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x)
    if np.iscomplexobj(x):
        return x.astype(np.complex128)
    x = np.asarray(x, dtype=np.float64)
    return x[..., 0::2] + 1j * x[..., 1::2]


def exact_twiddle(k: np.ndarray, n: int, sign: int = -1) -> np.ndarray:
    """e^{sign * 2 pi i k / n} in float64 with the angle reduced exactly.

    x = 2k/n is split into q quarter turns and a remainder d in [-1/4, 1/4] (both exact
    for a power-of-two n), sin/cos are taken of pi*d only, and the quarter turns are
    applied by swapping/negating -- so multiples of a quarter turn come out exact."""
    k = np.asarray(k, dtype=np.int64) % n
    x = 2.0 * k.astype(np.float64) / n       # exact: n is a power of two
    q = np.rint(2.0 * x)
    d = x - 0.5 * q
    s, c = np.sin(np.pi * d), np.cos(np.pi * d)
    qi = q.astype(np.int64) & 3
    sn = np.select([qi == 0, qi == 1, qi == 2], [s, c, -s], -c)
    cs = np.select([qi == 0, qi == 1, qi == 2], [c, -s, -c], s)
    return cs + sign * 1j * sn


def fp64_complex(x, inverse: bool) -> np.ndarray:
    """Unnormalised complex DFT in complex128 (src/ckfft/fft.cpp:13-46 semantics)."""
    z = _to_c128(x)
    n = z.shape[-1]
    return np.fft.ifft(z, axis=-1) * n if inverse else np.fft.fft(z, axis=-1)


def fp64_real_forward(x) -> np.ndarray:
    """2 * DFT(x)[0..n/2] (src/ckfft/fft_real_default.cpp:13-63 yields the factor 2)."""
    return 2.0 * np.fft.rfft(np.asarray(x, np.float64), axis=-1)


def fp64_real_inverse(y, n: int) -> np.ndarray:
    """The exact linear map of src/ckfft/fft_real_default.cpp:65-114 in float64.

    With h = n/2 and Y the n/2+1 inputs, the pre-pass builds for every k in [0, h]
        t[k] = (Y[k] + conj Y[h-k]) + i e^{+2 pi i k/n} (Y[k] - conj Y[h-k])
    (the k = n/4 "middle" special case 2*conj(Y[n/4]) is this formula at k = n/4), then an
    unnormalised inverse h-point DFT of t[0..h-1] gives z; the output is x[2m] = Re z[m],
    x[2m+1] = Im z[m].  Im Y[0] and Im Y[h] are not ignored.  For n <= 4 the reference
    uses closed forms (src/ckfft/fft_real.cpp:69-93) that ignore those imaginary parts.
    """
    Y = _to_c128(y)
    if n == 1:
        return Y[..., 0].real[..., None].copy()
    if n == 2:
        return np.stack([Y[..., 0].real + Y[..., 1].real, Y[..., 0].real - Y[..., 1].real], axis=-1)
    if n == 4:
        s02 = Y[..., 0].real + Y[..., 2].real
        s13 = 2 * Y[..., 1].real
        d02 = Y[..., 0].real - Y[..., 2].real
        d13 = 2 * Y[..., 1].imag
        return np.stack([s02 + s13, d02 - d13, s02 - s13, d02 + d13], axis=-1)
    h = n // 2
    k = np.arange(h + 1)
    Yr = Y[..., ::-1]                       # Yr[k] = Y[h-k]
    w = exact_twiddle(k, n, sign=+1)
    t = (Y + np.conj(Yr)) + 1j * w * (Y - np.conj(Yr))
    z = np.fft.ifft(t[..., :h], axis=-1) * h
    out = np.empty(Y.shape[:-1] + (n,), np.float64)
    out[..., 0::2] = z.real
    out[..., 1::2] = z.imag
    return out


def dft_brute(z: np.ndarray, sign: int = -1) -> np.ndarray:
    """O(N^2) DFT in float64; angles reduced exactly via (k*n mod N)."""
    z = np.asarray(z, np.complex128)
    n = z.shape[-1]
    kn = np.outer(np.arange(n), np.arange(n)) % n
    W = exact_twiddle(kn, n, sign=sign)
    return z @ W.T


def split_norm_err(got: np.ndarray, ref: np.ndarray) -> float:
    """max|got-ref| / (||ref||_2 / sqrt(len)): the output-scale reading of the
    'max abs diff <= 1e-5 * ||x||_2 / sqrt(N)' gate, per transform, worst over batch."""
    g = np.asarray(got, np.float64).reshape(-1, got.shape[-1])
    r = np.asarray(ref, np.float64).reshape(-1, ref.shape[-1])
    scale = np.sqrt(np.sum(r * r, axis=-1) / r.shape[-1])
    scale = np.where(scale == 0, 1.0, scale)
    return float(np.max(np.max(np.abs(g - r), axis=-1) / scale))


def rel_l2(got: np.ndarray, ref: np.ndarray) -> float:
    """||got - ref||_2 / ||ref||_2 over the whole array (float64)."""
    g = np.asarray(got, np.float64).ravel()
    r = np.asarray(ref, np.float64).ravel()
    d = np.linalg.norm(r)
    return float(np.linalg.norm(g - r) / (d if d else 1.0))


def interleave(z: np.ndarray) -> np.ndarray:
    """complex -> float64 interleaved (re, im) on the last axis."""
    z = np.asarray(z)
    out = np.empty(z.shape[:-1] + (2 * z.shape[-1],), np.float64)
    out[..., 0::2] = z.real
    out[..., 1::2] = z.imag
    return out


# ---------------------------------------------------------------------------
# Seeded synthetic inputs shared by tests, smoke and bench (SplitMix64 -> float32)
# ---------------------------------------------------------------------------

def splitmix64(seed: int, count: int) -> np.ndarray:
    """``count`` successive SplitMix64 outputs as uint64 (vectorised)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform_pm1(seed: int, count: int) -> np.ndarray:
    """iid Uniform[-1, 1) float32 from the top 24 bits of SplitMix64."""
    u = (splitmix64(seed, count) >> np.uint64(40)).astype(np.float64) / float(1 << 24)
    return (2.0 * u - 1.0).astype(np.float32)


def normal01(seed: int, count: int) -> np.ndarray:
    """iid N(0,1) float32 by Box-Muller on SplitMix64 uniforms."""
    m = (count + 1) // 2
    r = splitmix64(seed, 2 * m)
    u1 = ((r[0::2] >> np.uint64(11)).astype(np.float64) + 0.5) / float(1 << 53)
    u2 = ((r[1::2] >> np.uint64(11)).astype(np.float64)) / float(1 << 53)
    rad = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * m, np.float64)
    out[0::2] = rad * np.cos(2 * np.pi * u2)
    out[1::2] = rad * np.sin(2 * np.pi * u2)
    return out[:count].astype(np.float32)


def reference_input_txt() -> Optional[np.ndarray]:
    """The reference's own test input (src/test/input.txt, 4096 complex values),
    committed as data under tests/golden/input_txt.f32."""
    p = os.path.join(os.path.dirname(HERE), "tests", "golden", "input_txt.f32")
    if not os.path.exists(p):
        return None
    return np.fromfile(p, dtype="<f4")
