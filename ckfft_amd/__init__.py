"""ckfft_amd -- MI355X-native drop-in for ckfft's transform path.

The product is the C-ABI library ``libckfft_amd.so`` (declared in ``include/ckfft/``):
the reference's six entry points (``CkFftInit``, ``CkFftShutdown``,
``CkFftComplexForward/Inverse``, ``CkFftRealForward/Inverse``) plus batched,
stream-ordered extensions.  This module is a thin ctypes binding of that library,
with the same names, argument meanings and 0/1/NULL error behaviour, so that Python
callers and the tests drive exactly the code a C caller would.

There is no CPU fallback: importing this module raises if the HIP library has not
been built (``python ckfft_amd/build.py``).
"""
from __future__ import annotations

import ctypes
import os

__all__ = [
    "LIB_PATH", "lib", "CkFftComplex", "kCkFftDirection_Forward", "kCkFftDirection_Inverse",
    "kCkFftDirection_Both", "CkFftInit", "CkFftShutdown", "CkFftComplexForward", "CkFftComplexInverse",
    "CkFftRealForward", "CkFftRealInverse", "CkFftComplexForwardBatch", "CkFftComplexInverseBatch",
    "CkFftRealForwardBatch", "CkFftRealInverseBatch", "CkFftGetTwiddles", "CkFftGpuVersion", "CkFftGpuLastError", "Context", "BatchPlan", "CkFftBatchDesc",
    "CkFftBatchInfo", "CkFftBatchPlanCreate", "CkFftBatchPlanExecute", "CkFftBatchPlanGetInfo",
    "CkFftBatchPlanDestroy", "CkFftBatchLastError", "CkFftAutotune", "CkFftLoadTuning",
    "CkFftServerStats", "CkFftServerCreate", "CkFftServerSubmit", "CkFftServerReplay", "CkFftServerQuery",
    "CkFftServerWait", "CkFftServerDrain", "CkFftServerGetStats", "CkFftServerResetStats",
    "CkFftServerDestroy", "CkFftServerLastError", "Server", "CkFftSetPrecision", "CkFftGetPrecision",
    "kCkFftPrecision_Fp32", "kCkFftPrecision_SplitBf16x3", "CkFftDeviceTableBytes", "CkFftInitArena",
    "CkFftTransformBatchNoAlloc", "NOALLOC_SCRATCH_BYTES", "CkFftAutotuneMix", "EXPORTED_SYMBOLS",
]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libckfft_amd.so")

# One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64.so.7.  If this
# library were loaded first it would bring in /opt/rocm's copy and torch would then
# load a second runtime beside it, and the two do not share devices, streams or
# allocations.  Loading torch first makes the dynamic linker bind this library to
# torch's already-loaded runtime.
try:
    import torch as _torch  # noqa: F401
except ImportError:  # a process without torch uses /opt/rocm's runtime only
    _torch = None

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build it with `python ckfft_amd/build.py` "
                      "(there is no CPU fallback)")
lib = ctypes.CDLL(LIB_PATH)


class CkFftComplex(ctypes.Structure):
    """``{float real; float imag;}`` -- inc/ckfft/ckfft.h:10-15 of the reference."""
    _fields_ = [("real", ctypes.c_float), ("imag", ctypes.c_float)]


kCkFftDirection_Forward = 1
kCkFftDirection_Inverse = 2
kCkFftDirection_Both = 3

kCkFftKind_ComplexForward = 0
kCkFftKind_ComplexInverse = 1
kCkFftKind_RealForward = 2
kCkFftKind_RealInverse = 3
kCkFftBatch_Graph = 1
kCkFftPrecision_Fp32 = 0
kCkFftPrecision_SplitBf16x3 = 1
# CKFFT_NOALLOC_SCRATCH_BYTES: the fixed per-GPU scratch of CkFftTransformBatchNoAlloc
NOALLOC_SCRATCH_BYTES = 64 << 20


class CkFftBatchDesc(ctypes.Structure):
    """One transform of an uneven batch (include/ckfft/ckfft_gpu.h)."""
    _fields_ = [("kind", ctypes.c_int), ("n", ctypes.c_int), ("input", ctypes.c_void_p),
                ("output", ctypes.c_void_p)]


class CkFftBatchInfo(ctypes.Structure):
    _fields_ = [("transforms", ctypes.c_int), ("persistent_transforms", ctypes.c_int),
                ("work_items", ctypes.c_int), ("workgroups", ctypes.c_int), ("extra_launches", ctypes.c_int),
                ("algorithmic_bytes", ctypes.c_longlong), ("pass_transforms", ctypes.c_int)]


class CkFftServerStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_longlong), ("failed", ctypes.c_longlong), ("p50_us", ctypes.c_double),
                ("p99_us", ctypes.c_double), ("max_us", ctypes.c_double), ("mean_us", ctypes.c_double),
                ("requests_per_s", ctypes.c_double), ("GB_per_s", ctypes.c_double),
                ("peak_inflight_bytes", ctypes.c_longlong), ("budget_bytes", ctypes.c_longlong),
                ("resident_bytes", ctypes.c_longlong), ("batches", ctypes.c_longlong),
                ("mean_batch", ctypes.c_double), ("max_client_us", ctypes.c_double),
                ("max_wait_us", ctypes.c_double), ("max_device_us", ctypes.c_double)]


_vp = ctypes.c_void_p
_i = ctypes.c_int
_ll = ctypes.c_longlong

_SIGS = {
    "CkFftInit": (_vp, [_i, _i, _vp, ctypes.POINTER(ctypes.c_size_t)]),
    "CkFftShutdown": (None, [_vp]),
    "CkFftComplexForward": (_i, [_vp, _i, _vp, _vp]),
    "CkFftComplexInverse": (_i, [_vp, _i, _vp, _vp]),
    "CkFftRealForward": (_i, [_vp, _i, _vp, _vp]),
    "CkFftRealInverse": (_i, [_vp, _i, _vp, _vp, _vp]),
    "CkFftComplexForwardBatch": (_i, [_vp, _i, _i, _vp, _vp, _vp]),
    "CkFftComplexInverseBatch": (_i, [_vp, _i, _i, _vp, _vp, _vp]),
    "CkFftRealForwardBatch": (_i, [_vp, _i, _i, _vp, _vp, _vp]),
    "CkFftRealInverseBatch": (_i, [_vp, _i, _i, _vp, _vp, _vp]),
    "CkFftGetTwiddles": (_i, [_vp, _i, _vp]),
    "CkFftGpuVersion": (ctypes.c_char_p, []),
    "CkFftGpuLastError": (ctypes.c_char_p, []),
    "CkFftBatchPlanCreate": (_vp, [_vp, ctypes.POINTER(CkFftBatchDesc), _i, _i]),
    "CkFftBatchPlanExecute": (_i, [_vp, _vp]),
    "CkFftBatchPlanGetInfo": (_i, [_vp, ctypes.POINTER(CkFftBatchInfo)]),
    "CkFftBatchPlanDestroy": (None, [_vp]),
    "CkFftBatchLastError": (ctypes.c_char_p, []),
    "CkFftAutotune": (_i, [_vp, ctypes.c_char_p]),
    "CkFftLoadTuning": (_i, [_vp, ctypes.c_char_p]),
    "CkFftServerCreate": (_vp, [_vp, ctypes.c_size_t, _i]),
    "CkFftServerSubmit": (_ll, [_vp, _i, _i, _vp, _vp]),
    "CkFftServerReplay": (_i, [_vp, ctypes.POINTER(CkFftBatchDesc), ctypes.POINTER(ctypes.c_double), _i,
                               ctypes.POINTER(_ll)]),
    "CkFftServerQuery": (_i, [_vp, _ll]),
    "CkFftServerWait": (_i, [_vp, _ll]),
    "CkFftServerDrain": (_i, [_vp]),
    "CkFftServerGetStats": (_i, [_vp, ctypes.POINTER(CkFftServerStats)]),
    "CkFftServerResetStats": (None, [_vp]),
    "CkFftServerDestroy": (None, [_vp]),
    "CkFftServerLastError": (ctypes.c_char_p, []),
    "CkFftSetPrecision": (_i, [_vp, _i]),
    "CkFftDeviceTableBytes": (ctypes.c_size_t, [_i, _i]),
    "CkFftInitArena": (_vp, [_i, _i, _vp, ctypes.POINTER(ctypes.c_size_t), _vp, ctypes.c_size_t]),
    "CkFftGetPrecision": (_i, [_vp]),
    "CkFftTransformBatchNoAlloc": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp]),
    "CkFftAutotuneMix": (_i, [_vp, ctypes.POINTER(CkFftBatchDesc), _i, ctypes.c_char_p,
                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
}
EXPORTED_SYMBOLS = tuple(_SIGS)

for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args
    globals()[_name] = _fn


def _ptr(x):
    """Accept a raw address, None, a ctypes object/array, or a torch tensor."""
    if x is None or isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    if hasattr(x, "ctypes"):
        return x.ctypes.data
    return ctypes.cast(x, ctypes.c_void_p).value


def _stream(s):
    if s is None:
        return None
    if isinstance(s, int):
        return s
    return getattr(s, "cuda_stream", s)


class Context:
    """Owns a ``CkFftContext*``; methods mirror the C calls and return their 0/1 result.

    Buffers may be torch tensors (device or host), numpy arrays or raw addresses.
    """

    def __init__(self, n_max: int, direction: int = kCkFftDirection_Both, arena=None):
        """``arena``: optional device buffer (at least CkFftDeviceTableBytes bytes) that
        holds the context's tables instead of memory the library allocates."""
        self.n_max = n_max
        self.direction = direction
        self.arena = arena                       # keeps a caller's arena alive
        if arena is None:
            self.handle = CkFftInit(n_max, direction, None, None)
        else:
            nbytes = arena.numel() * arena.element_size() if hasattr(arena, "numel") else None
            self.handle = CkFftInitArena(n_max, direction, None, None, _ptr(arena), nbytes or 0)
        if not self.handle:
            raise RuntimeError(f"CkFftInit({n_max}, {direction}) failed: "
                               f"{CkFftGpuLastError().decode()} (or invalid size/direction)")

    def close(self):
        if self.handle:
            CkFftShutdown(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # single-signal, synchronous (the reference's calls)
    def complex_forward(self, n, x, y) -> int:
        return CkFftComplexForward(self.handle, n, _ptr(x), _ptr(y))

    def complex_inverse(self, n, x, y) -> int:
        return CkFftComplexInverse(self.handle, n, _ptr(x), _ptr(y))

    def real_forward(self, n, x, y) -> int:
        return CkFftRealForward(self.handle, n, _ptr(x), _ptr(y))

    def real_inverse(self, n, y, x, tmp) -> int:
        return CkFftRealInverse(self.handle, n, _ptr(y), _ptr(x), _ptr(tmp))

    # batched, stream-ordered (device memory)
    def complex_forward_batch(self, n, count, x, y, stream=None) -> int:
        return CkFftComplexForwardBatch(self.handle, n, count, _ptr(x), _ptr(y), _stream(stream))

    def complex_inverse_batch(self, n, count, x, y, stream=None) -> int:
        return CkFftComplexInverseBatch(self.handle, n, count, _ptr(x), _ptr(y), _stream(stream))

    def real_forward_batch(self, n, count, x, y, stream=None) -> int:
        return CkFftRealForwardBatch(self.handle, n, count, _ptr(x), _ptr(y), _stream(stream))

    def real_inverse_batch(self, n, count, y, x, stream=None) -> int:
        return CkFftRealInverseBatch(self.handle, n, count, _ptr(y), _ptr(x), _stream(stream))

    def transform_batch_noalloc(self, kind, n, count, x, y, stream=None) -> int:
        """CkFftTransformBatchNoAlloc: the batched call for any kind (kCkFftKind_*) with the
        multi-pass intermediate in the library's fixed per-GPU scratch (no allocation)."""
        return CkFftTransformBatchNoAlloc(self.handle, kind, n, count, _ptr(x), _ptr(y), _stream(stream))

    def twiddles(self, L: int, out) -> int:
        return CkFftGetTwiddles(self.handle, L, _ptr(out))

    def autotune(self, cache_path=None) -> int:
        return CkFftAutotune(self.handle, cache_path.encode() if cache_path else None)

    def autotune_mix(self, transforms, cache_path=None):
        """CkFftAutotuneMix over (kind, n, input, output) transforms: returns (ok, us before,
        us after) for one execution of their plan."""
        arr = (CkFftBatchDesc * max(1, len(transforms)))()
        for i, (kind, n, x, y) in enumerate(transforms):
            arr[i] = CkFftBatchDesc(kind, n, _ptr(x), _ptr(y))
        before, after = ctypes.c_double(), ctypes.c_double()
        ok = CkFftAutotuneMix(self.handle, arr, len(transforms), cache_path.encode() if cache_path else None,
                              ctypes.byref(before), ctypes.byref(after))
        return ok, before.value, after.value

    def set_precision(self, precision: int) -> int:
        """kCkFftPrecision_Fp32 (default) or kCkFftPrecision_SplitBf16x3 (complex
        16 <= n <= 65536 on the matrix cores)."""
        return CkFftSetPrecision(self.handle, precision)

    def precision(self) -> int:
        return CkFftGetPrecision(self.handle)

    def load_tuning(self, cache_path) -> int:
        return CkFftLoadTuning(self.handle, cache_path.encode())


class BatchPlan:
    """An uneven batch run in persistent launches (CkFftBatchPlan*).

    ``transforms`` is a sequence of (kind, n, input, output) with device buffers.
    """

    def __init__(self, ctx: Context, transforms, graph: bool = False):
        arr = (CkFftBatchDesc * max(1, len(transforms)))()
        for i, (kind, n, x, y) in enumerate(transforms):
            arr[i] = CkFftBatchDesc(kind, n, _ptr(x), _ptr(y))
        self.ctx = ctx                         # keeps the context alive
        self.handle = CkFftBatchPlanCreate(ctx.handle, arr, len(transforms), kCkFftBatch_Graph if graph else 0)
        if not self.handle:
            raise ValueError(f"CkFftBatchPlanCreate failed: {CkFftBatchLastError().decode()}")

    def execute(self, stream=None) -> int:
        return CkFftBatchPlanExecute(self.handle, _stream(stream))

    def info(self) -> dict:
        inf = CkFftBatchInfo()
        assert CkFftBatchPlanGetInfo(self.handle, ctypes.byref(inf)) == 1
        return {f: getattr(inf, f) for f, _ in CkFftBatchInfo._fields_}

    def close(self):
        if self.handle:
            CkFftBatchPlanDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching serving loop (CkFftServer*): requests admitted against an HBM
    budget as they arrive, batched into persistent launches on up to ``max_in_flight``
    streams at once.  Buffers are device memory the caller keeps alive until done."""

    def __init__(self, ctx: Context, hbm_budget: int, max_in_flight: int = 4):
        self.ctx = ctx                         # keeps the context alive
        self.handle = CkFftServerCreate(ctx.handle, int(hbm_budget), int(max_in_flight))
        if not self.handle:
            raise ValueError(f"CkFftServerCreate failed: {CkFftServerLastError().decode()}")

    def submit(self, kind, n, x, y) -> int:
        return CkFftServerSubmit(self.handle, kind, n, _ptr(x), _ptr(y))

    def replay(self, transforms, arrival_us):
        """Submit (kind, n, input, output) transforms at the given arrival times
        (microseconds from the call); returns the list of tickets."""
        count = len(transforms)
        arr = (CkFftBatchDesc * max(1, count))()
        for i, (kind, n, x, y) in enumerate(transforms):
            arr[i] = CkFftBatchDesc(kind, n, _ptr(x), _ptr(y))
        times = (ctypes.c_double * max(1, count))(*[float(t) for t in arrival_us])
        tickets = (_ll * max(1, count))()
        ok = CkFftServerReplay(self.handle, arr, times, count, tickets)
        if not ok:
            raise ValueError(f"CkFftServerReplay rejected a request: {CkFftServerLastError().decode()}")
        return list(tickets)[:count]

    def query(self, ticket) -> int:
        return CkFftServerQuery(self.handle, ticket)

    def wait(self, ticket) -> int:
        return CkFftServerWait(self.handle, ticket)

    def drain(self) -> int:
        return CkFftServerDrain(self.handle)

    def stats(self) -> dict:
        st = CkFftServerStats()
        assert CkFftServerGetStats(self.handle, ctypes.byref(st)) == 1
        return {f: getattr(st, f) for f, _ in CkFftServerStats._fields_}

    def reset_stats(self):
        CkFftServerResetStats(self.handle)

    def close(self):
        if self.handle:
            CkFftServerDestroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
