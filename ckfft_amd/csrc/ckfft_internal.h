// ckfft_internal.h -- host runtime internals shared by the C-ABI shim and the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>

#include "ckfft/ckfft_gpu.h"

namespace ckfft_gpu {

// Transform kinds.  The real kinds are computed through a complex transform of
// half the size with the pack (forward) or unpack (inverse) pass fused in.
enum Kind : int {
    kComplexForward = 0,
    kComplexInverse = 1,
    kRealForward = 2,
    kRealInverse = 3,
};

// Largest complex size handled by one workgroup (registers + LDS exchange).
constexpr int kLogMaxLds = 15;
// Largest level of the twiddle pyramid kept on the device.
constexpr int kLogMaxPyramid = 16;
// Fine twiddle tables for levels above the pyramid hold 2^kLogFine entries each.
constexpr int kLogFine = 12;
// Largest supported transform (complex points or real samples): 2^28.  Twiddles of a
// level L > 2^16 are a coarse pyramid entry (level L / 2^12 <= 2^16) times a fine one.
constexpr int kLogMaxSupported = kLogMaxPyramid + kLogFine;
// Transposed inter-pass twiddles: for the two-pass transforms whose second pass has span
// 2^kLogXtwCols, table[r * 2^kLogXtwCols + j] = W_L^{r j} (L = 2^log_xtw, the context's
// largest pyramid level when it is >= 2^kLogMinXtw), so that the columns j a tile reads
// side by side find their twiddles side by side.
constexpr int kLogXtwCols = 8;
constexpr int kLogMinXtw = 15;
// Radix range of one multi-pass stage.
constexpr int kLogMinPass = 6;
constexpr int kLogMaxPass = 12;

// Device twiddle tables of a context.
//  * pyramid: level L (a power of two <= 2^log_pyramid) occupies entries [L, 2L);
//    entry L + k holds exp(-2 pi i k / L).  Entry 0 is unused (set to 1).
//  * fine: for each level L = 2^l, l in (16, log_max], 2^12 entries exp(-2 pi i m / L),
//    m < 2^12, at offset (l - 17) * 2^12.
// Every entry is evaluated in FP64 with an exact argument reduction and rounded once.
struct DeviceTables {
    const float2* pyramid;
    int log_pyramid;
    const float2* fine;      // null when log_max <= 16
    int log_max;             // largest log2(n) the tables serve
    int precision;           // kCkFftPrecision_* of the context
    const float2* xtw;       // transposed inter-pass twiddles (null when log_xtw == 0)
    int log_xtw;
    hipMemPool_t pool;       // host-side handle: the pool of multi-pass intermediates (null:
                             // the device's default pool)
};

// The tables (and pool) of a context, as the launchers take them.
DeviceTables context_tables(const CkFftContext* c);

// Launchers (return hipSuccess or an error).
hipError_t launch_twiddle_pyramid(float2* table, int log_lmax, hipStream_t stream);
hipError_t launch_twiddle_fine(float2* table, int log_max, hipStream_t stream);
hipError_t launch_twiddle_transposed(float2* table, const float2* pyramid, int log_l, hipStream_t stream);

// Batched transform of `count` signals of size n (n >= 1, power of two, n <= the
// kind's limit).  Pointers are device memory with the layouts of ckfft_gpu.h.  Sizes
// above one workgroup need an intermediate: `scratch` (device memory of scratch_bytes)
// when it is given, else a stream-ordered allocation.
hipError_t launch_transform(Kind kind, int log2n, int count, const void* in, void* out,
                            const DeviceTables& tables, hipStream_t stream, void* scratch = nullptr,
                            size_t scratch_bytes = 0);

// Bytes of intermediate launch_transform needs for this transform (0: none).
size_t transform_scratch_bytes(Kind kind, int log2n, int count, const DeviceTables& tables);

// The fixed per-device scratch array of CkFftTransformBatchNoAlloc (fft_pass.hip):
// noalloc_scratch_setup looks up its address on `device`, which must be the current
// device, once per device; noalloc_scratch returns that address (null before a successful
// setup on that device).
hipError_t noalloc_scratch_setup(int device);
void* noalloc_scratch(int device);

// Largest log2(n) launch_transform supports for a kind.
int max_log2n(Kind kind);

// Split-BF16x3 matrix-core path (fft_mfma.hip): complex kinds, 2^4 <= n <= 2^16.  The
// dispatch above routes there when the tables' precision asks for it.
bool mfma_supported(Kind kind, int log2n);
hipError_t launch_mfma(Kind kind, int log2n, int count, const void* in, void* out,
                       const DeviceTables& tables, hipStream_t stream);

// Host-pointer staging (fft_staging.cpp): a pool of slots (stream, device buffer, pinned
// bounce buffer) per context; one synchronous transform with any mix of host and device
// buffers, pipelining the copies.
struct StagingPool;
StagingPool* staging_create();
void staging_destroy(StagingPool* pool);
hipError_t staged_transform(StagingPool* pool, Kind kind, int log2n, const void* in, size_t in_bytes, void* out,
                            size_t out_bytes, const DeviceTables& tables);

// Multi-pass path (fft_pass.hip) for sizes above one workgroup.
// `scratch`: the intermediate (multipass_scratch_bytes), or null to allocate it on the
// stream from the tables' pool.
hipError_t launch_multipass(Kind kind, int log2n, int count, const void* in, void* out,
                            const DeviceTables& tables, hipStream_t stream, void* scratch = nullptr,
                            size_t scratch_bytes = 0);
size_t multipass_scratch_bytes(Kind kind, int log2n, int count);

}  // namespace ckfft_gpu

// Host part of a context.  Lives in library memory or in the caller's buffer.
struct _CkFftContext {
    uint64_t magic;          // kContextMagic while alive
    int max_count;           // nMax given to CkFftInit
    int direction;           // CkFftDirection bits
    int own_host;            // 1 if the library allocated this struct
    int device;              // HIP device the tables live on
    int log_pyramid;         // levels 1..2^log_pyramid are on the device
    float2* d_pyramid;       // device twiddle pyramid, 2 << log_pyramid entries
    float2* d_fine;          // fine twiddle tables for levels above 2^16 (or null)
    float2* d_xtw;           // transposed inter-pass twiddles (or null)
    int log_xtw;             // their level, 0 when absent
    void* d_tables;          // the one device block holding the three tables above
    int own_device;          // 1 if the library allocated d_tables (else a caller's arena)
    int log_max;             // largest log2(n) the tables serve
    // Batch-plan layout per [kind][log2 complex size] (CkFftAutotune, CkFftAutotuneMix);
    // every layout has the same arithmetic.  Sizes 2^1..2^15: 0 = the default, else
    // 1 + var + 2 * wide, var = 32 instead of 16 points per thread (2^5..2^13), wide = the
    // one-workgroup-per-CU persistent kernel instead of the two-per-CU one.  [kind][16]
    // (complex 2^16): 0 = the table-driven passes over all such transforms, 2 = an ordinary
    // two-pass launch each.
    unsigned char tune[4][17];
    // kCkFftPrecision_*: FP32 (default) or split-BF16x3 matrix-core radix-16 stages for
    // the complex sizes that mode covers.
    int precision;
    // Staging slots for host-pointer calls.
    ckfft_gpu::StagingPool* staging;
    // Stream-ordered pool of the multi-pass intermediates; it keeps freed memory for reuse
    // until CkFftShutdown.  Null for a context in a caller's device arena, whose
    // intermediates are allocated and released per call.
    hipMemPool_t pool;
};

constexpr uint64_t kContextMagic = 0x434b46465447505bull;  // "CKFFTGP["
