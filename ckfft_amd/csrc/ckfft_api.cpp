// ckfft_api.cpp -- the C-ABI boundary: ckfft's six entry points plus the batched
// extensions, backed by the HIP kernels.
//
// Argument validation follows the reference's C API (src/ckfft/ckfft.cpp:14-114) rule
// for rule and returns the same 0 / 1 / NULL values; the context's size query and
// caller-buffer mode follow src/ckfft/context.cpp:26-67.  Differences, all deliberate:
//  * n <= 0 is rejected (the reference's unsigned power-of-two test lets INT_MIN in);
//  * a context pointer is also checked for a live magic value;
//  * 0 is also returned if the GPU reports an error for the call;
//  * the size written in query mode covers only the host part of the context, since
//    the twiddle tables live in device memory the library allocates itself.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>

#include "ckfft/ckfft.h"
#include "ckfft/ckfft_gpu.h"
#include "ckfft_internal.h"

using namespace ckfft_gpu;

namespace {

bool pow2_positive(int n) { return n > 0 && (n & (n - 1)) == 0; }

int ilog2(int n)
{
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

size_t host_context_bytes() { return (sizeof(_CkFftContext) + 7) & ~size_t(7); }

bool context_ok(const CkFftContext* c, int need_direction)
{
    return c && c->magic == kContextMagic && (c->direction & need_direction);
}

// src/ckfft/ckfft.cpp:42-49 / 84-91: size, pointer and aliasing rules.
bool size_ok(const CkFftContext* c, int n) { return pow2_positive(n) && n <= c->max_count; }

size_t in_bytes(Kind k, int n)
{
    switch (k) {
    case kRealForward: return sizeof(float) * (size_t)n;
    case kRealInverse: return sizeof(CkFftComplex) * ((size_t)n / 2 + 1);
    default: return sizeof(CkFftComplex) * (size_t)n;
    }
}

size_t out_bytes(Kind k, int n)
{
    switch (k) {
    case kRealForward: return sizeof(CkFftComplex) * ((size_t)n / 2 + 1);
    case kRealInverse: return sizeof(float) * (size_t)n;
    default: return sizeof(CkFftComplex) * (size_t)n;
    }
}

// Device memory (hipMalloc, torch) runs in place; anything else is staged.
bool is_device_pointer(const void* p)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();   // plain pageable host memory is unknown to HIP
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// Last failure on this thread, for CkFftGpuLastError().
thread_local char g_last_error[256] = "no error";

bool record(hipError_t e, const char* where)
{
    if (e == hipSuccess) return true;
    std::snprintf(g_last_error, sizeof(g_last_error), "%s: %s (%d)", where, hipGetErrorString(e), (int)e);
    return false;
}

int fail(hipError_t e, const char* where)
{
    record(e, where);
    return 0;
}

// One synchronous transform, as every drop-in call is: device buffers in place on the
// calling thread's default stream, host buffers through a staging slot.
int run_plain(CkFftContext* c, Kind kind, int n, const void* in, void* out)
{
    const int log2n = ilog2(n);
    if (log2n > max_log2n(kind)) return 0;
    return record(staged_transform(c->staging, kind, log2n, in, in_bytes(kind, n), out, out_bytes(kind, n), context_tables(c)),
                  "transform")
               ? 1
               : 0;
}

int run_batch(CkFftContext* c, Kind kind, int n, int count, const void* in, void* out, void* stream)
{
    if (!size_ok(c, n) || count < 0) return 0;
    if (!in || !out || in == out) return 0;
    const int log2n = ilog2(n);
    if (log2n > max_log2n(kind)) return 0;
    if (count == 0) return 1;
    if (!is_device_pointer(in) || !is_device_pointer(out)) return 0;
    hipError_t e = launch_transform(kind, log2n, count, in, out, context_tables(c), (hipStream_t)stream);
    return record(e, "launch") ? 1 : 0;
}

// CkFftTransformBatchNoAlloc: run_batch's rules, with the multi-pass intermediate in the
// fixed per-device scratch array.  Nothing on this path allocates: validation reads
// pointer attributes, and launch_transform is given the scratch.
int run_batch_noalloc(CkFftContext* c, Kind kind, int n, int count, const void* in, void* out, void* stream)
{
    if (!size_ok(c, n) || count < 0) return 0;
    if (!in || !out || in == out) return 0;
    const int log2n = ilog2(n);
    if (log2n > max_log2n(kind)) return 0;
    if (count == 0) return 1;
    if (!is_device_pointer(in) || !is_device_pointer(out)) return 0;
    const DeviceTables tb = context_tables(c);
    const size_t need = transform_scratch_bytes(kind, log2n, count, tb);
    void* scratch = nullptr;
    if (need > 0) {
        if (need > CKFFT_NOALLOC_SCRATCH_BYTES) {
            std::snprintf(g_last_error, sizeof(g_last_error),
                          "no-alloc call: intermediate of %zu bytes exceeds the fixed scratch of %zu bytes", need,
                          (size_t)CKFFT_NOALLOC_SCRATCH_BYTES);
            return 0;
        }
        scratch = noalloc_scratch(c->device);
        if (!scratch) {
            std::snprintf(g_last_error, sizeof(g_last_error),
                          "no-alloc call: the fixed scratch of device %d is unavailable (its address lookup at "
                          "CkFftInit failed)",
                          c->device);
            return 0;
        }
    }
    hipError_t e = launch_transform(kind, log2n, count, in, out, tb, (hipStream_t)stream, scratch,
                                    scratch ? (size_t)CKFFT_NOALLOC_SCRATCH_BYTES : 0);
    return record(e, "launch") ? 1 : 0;
}

// The context's device tables, in one block: the pyramid, the fine tables (levels above
// 2^16) and the transposed inter-pass table, each 256-byte aligned.
struct TableLayout {
    int log_pyr, log_max, log_xtw;
    size_t pyr_off, fine_off, xtw_off, total;
};

TableLayout table_layout(int nMax)
{
    TableLayout t;
    t.log_pyr = std::max(1, std::min(ilog2(nMax), kLogMaxPyramid));
    t.log_max = std::min(ilog2(nMax), kLogMaxSupported);
    t.log_xtw = t.log_pyr >= kLogMinXtw ? t.log_pyr : 0;
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    t.pyr_off = 0;
    t.fine_off = al(sizeof(float2) * ((size_t)2 << t.log_pyr));
    const size_t fine_bytes =
        t.log_max > kLogMaxPyramid ? sizeof(float2) * ((size_t)(t.log_max - kLogMaxPyramid) << kLogFine) : 0;
    t.xtw_off = t.fine_off + al(fine_bytes);
    t.total = t.xtw_off + al(t.log_xtw ? sizeof(float2) << t.log_xtw : 0);
    return t;
}

bool valid_init_args(int nMax, CkFftDirection direction)
{
    return pow2_positive(nMax) && (direction == kCkFftDirection_Forward || direction == kCkFftDirection_Inverse ||
                                   direction == kCkFftDirection_Both);
}

// CkFftInit with the device tables in `arena` (caller's device memory) or, when arena is
// null, in one block the library allocates.
CkFftContext* init_context(int nMax, CkFftDirection direction, void* buf, size_t* bufSize, void* arena,
                           size_t arena_bytes)
{
    if (!valid_init_args(nMax, direction)) return nullptr;
    if (buf && !bufSize) return nullptr;
    const size_t need = host_context_bytes();
    if (bufSize && (!buf || *bufSize < need)) {
        *bufSize = need;
        return nullptr;
    }

    int dev = 0;
    if (!record(hipGetDevice(&dev), "hipGetDevice")) {
        (void)hipGetLastError();
        return nullptr;
    }
    const TableLayout L = table_layout(nMax);
    char* block = (char*)arena;
    if (arena) {
        if (arena_bytes < L.total || ((uintptr_t)arena & 255) || !is_device_pointer(arena)) {
            std::snprintf(g_last_error, sizeof(g_last_error), "device arena: need %zu bytes of 256-aligned device memory",
                          L.total);
            return nullptr;
        }
    } else if (!record(hipMalloc((void**)&block, L.total), "hipMalloc(twiddles)")) {
        return nullptr;
    }
    float2* pyr = (float2*)(block + L.pyr_off);
    float2* fine = L.log_max > kLogMaxPyramid ? (float2*)(block + L.fine_off) : nullptr;
    float2* xtw = L.log_xtw ? (float2*)(block + L.xtw_off) : nullptr;
    const int log_pyr = L.log_pyr, log_max = L.log_max, log_xtw = L.log_xtw;
    auto release = [&]() {
        if (!arena) (void)hipFree(block);
    };
    if (!record(launch_twiddle_pyramid(pyr, log_pyr, nullptr), "twiddle launch") ||
        !record(launch_twiddle_fine(fine, log_max, nullptr), "fine twiddle launch") ||
        !record(launch_twiddle_transposed(xtw, pyr, log_xtw, nullptr), "pass twiddle launch") ||
        !record(hipStreamSynchronize(nullptr), "twiddle synchronize")) {
        release();
        return nullptr;
    }

    // the no-allocation entry point's scratch array: its address on this (current) device,
    // looked up once per device; a failure only disables that entry point
    (void)record(noalloc_scratch_setup(dev), "no-alloc scratch lookup");

    void* mem = buf ? buf : std::malloc(need);
    if (!mem) {
        release();
        return nullptr;
    }
    StagingPool* pool = staging_create();
    // the intermediates' pool is an optimisation: without one (creation refused), they come
    // from the device's default pool
    hipMemPool_t scratch = nullptr;
    if (pool && !arena) {
        hipMemPoolProps props{};
        props.allocType = hipMemAllocationTypePinned;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        uint64_t keep = UINT64_MAX;
        if (hipMemPoolCreate(&scratch, &props) != hipSuccess) {
            (void)hipGetLastError();
            scratch = nullptr;
        } else if (hipMemPoolSetAttribute(scratch, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipMemPoolDestroy(scratch);
            scratch = nullptr;
        }
    }
    if (!pool) {
        release();
        if (!buf) std::free(mem);
        return nullptr;
    }
    CkFftContext* c = new (mem) CkFftContext();
    c->magic = kContextMagic;
    c->max_count = nMax;
    c->direction = (int)direction;
    c->own_host = buf ? 0 : 1;
    c->device = dev;
    c->log_pyramid = log_pyr;
    c->d_pyramid = pyr;
    c->d_fine = fine;
    c->d_xtw = xtw;
    c->log_xtw = log_xtw;
    c->d_tables = block;
    c->own_device = arena ? 0 : 1;
    c->log_max = log_max;
    std::memset(c->tune, 0, sizeof(c->tune));
    c->precision = kCkFftPrecision_Fp32;
    c->staging = pool;
    c->pool = scratch;
    return c;
}

}  // namespace

extern "C" {

// src/ckfft/ckfft.cpp:14-34 and src/ckfft/context.cpp:24-114
CkFftContext* CkFftInit(int nMax, CkFftDirection direction, void* buf, size_t* bufSize)
{
    return init_context(nMax, direction, buf, bufSize, nullptr, 0);
}

size_t CkFftDeviceTableBytes(int nMax, CkFftDirection direction)
{
    return valid_init_args(nMax, direction) ? table_layout(nMax).total : 0;
}

CkFftContext* CkFftInitArena(int nMax, CkFftDirection direction, void* buf, size_t* bufSize, void* deviceArena,
                             size_t deviceArenaBytes)
{
    if (!deviceArena && !(bufSize && !buf)) return nullptr;   // an arena is required (query mode aside)
    return init_context(nMax, direction, buf, bufSize, deviceArena, deviceArenaBytes);
}

// src/ckfft/ckfft.cpp:36-53
int CkFftRealForward(CkFftContext* context, int n, const float* input, CkFftComplex* output)
{
    if (!context_ok(context, kCkFftDirection_Forward)) return 0;
    if (!size_ok(context, n)) return 0;
    if (!input || !output || (const void*)input == (const void*)output) return 0;
    return run_plain(context, kRealForward, n, input, output);
}

// src/ckfft/ckfft.cpp:55-76 (tmpBuf is checked first, as there)
int CkFftRealInverse(CkFftContext* context, int n, const CkFftComplex* input, float* output, CkFftComplex* tmpBuf)
{
    if (!tmpBuf) return 0;
    if (!context_ok(context, kCkFftDirection_Inverse)) return 0;
    if (!size_ok(context, n)) return 0;
    if (!input || !output || (const void*)input == (const void*)output) return 0;
    return run_plain(context, kRealInverse, n, input, output);
}

// src/ckfft/ckfft.cpp:78-95
int CkFftComplexForward(CkFftContext* context, int n, const CkFftComplex* input, CkFftComplex* output)
{
    if (!context_ok(context, kCkFftDirection_Forward)) return 0;
    if (!size_ok(context, n)) return 0;
    if (!input || !output || input == output) return 0;
    return run_plain(context, kComplexForward, n, input, output);
}

// src/ckfft/ckfft.cpp:97-114
int CkFftComplexInverse(CkFftContext* context, int n, const CkFftComplex* input, CkFftComplex* output)
{
    if (!context_ok(context, kCkFftDirection_Inverse)) return 0;
    if (!size_ok(context, n)) return 0;
    if (!input || !output || input == output) return 0;
    return run_plain(context, kComplexInverse, n, input, output);
}

// src/ckfft/ckfft.cpp:116-119 and src/ckfft/context.cpp:116-122
void CkFftShutdown(CkFftContext* context)
{
    if (!context || context->magic != kContextMagic) return;
    if (context->own_device) (void)hipFree(context->d_tables);
    staging_destroy(context->staging);
    // returns at once; memory still held by unfinished frees is released when they finish
    if (context->pool) (void)hipMemPoolDestroy(context->pool);
    const int own = context->own_host;
    context->magic = 0;
    context->~CkFftContext();
    if (own) std::free(context);
}

int CkFftComplexForwardBatch(CkFftContext* context, int n, int count, const CkFftComplex* input,
                             CkFftComplex* output, void* stream)
{
    if (!context_ok(context, kCkFftDirection_Forward)) return 0;
    return run_batch(context, kComplexForward, n, count, input, output, stream);
}

int CkFftComplexInverseBatch(CkFftContext* context, int n, int count, const CkFftComplex* input,
                             CkFftComplex* output, void* stream)
{
    if (!context_ok(context, kCkFftDirection_Inverse)) return 0;
    return run_batch(context, kComplexInverse, n, count, input, output, stream);
}

int CkFftRealForwardBatch(CkFftContext* context, int n, int count, const float* input, CkFftComplex* output,
                          void* stream)
{
    if (!context_ok(context, kCkFftDirection_Forward)) return 0;
    return run_batch(context, kRealForward, n, count, input, output, stream);
}

int CkFftRealInverseBatch(CkFftContext* context, int n, int count, const CkFftComplex* input, float* output,
                          void* stream)
{
    if (!context_ok(context, kCkFftDirection_Inverse)) return 0;
    return run_batch(context, kRealInverse, n, count, input, output, stream);
}

int CkFftTransformBatchNoAlloc(CkFftContext* context, int kind, int n, int count, const void* input, void* output,
                               void* stream)
{
    if (kind < kCkFftKind_ComplexForward || kind > kCkFftKind_RealInverse) return 0;
    const int dir = (kind == kCkFftKind_ComplexForward || kind == kCkFftKind_RealForward) ? kCkFftDirection_Forward
                                                                                          : kCkFftDirection_Inverse;
    if (!context_ok(context, dir)) return 0;
    return run_batch_noalloc(context, (Kind)kind, n, count, input, output, stream);
}

int CkFftGetTwiddles(CkFftContext* context, int L, CkFftComplex* out)
{
    if (!context || context->magic != kContextMagic || !out) return 0;
    if (!pow2_positive(L) || ilog2(L) > context->log_pyramid) return 0;
    return record(hipMemcpy(out, context->d_pyramid + L, sizeof(float2) * (size_t)L, hipMemcpyDeviceToHost),
                  "copy twiddles")
               ? 1
               : 0;
}

int CkFftSetPrecision(CkFftContext* context, int precision)
{
    if (!context || context->magic != kContextMagic) return 0;
    if (precision != kCkFftPrecision_Fp32 && precision != kCkFftPrecision_SplitBf16x3) return 0;
    context->precision = precision;
    return 1;
}

int CkFftGetPrecision(const CkFftContext* context)
{
    if (!context || context->magic != kContextMagic) return -1;
    return context->precision;
}

const char* CkFftGpuLastError(void) { return g_last_error; }

const char* CkFftGpuVersion(void)
{
    return "ckfft_amd 0.1 (gfx950 HIP; drop-in for ckfft 1.3 C API)";
}

}  // extern "C"
