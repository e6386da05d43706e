// ckfft_internal.h -- host runtime internals shared by the C-ABI shim and the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>

#include "ckfft/ckfft.h"

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

// Twiddle pyramid layout: level L (a power of two) occupies entries [L, 2L) and
// entry L + k holds exp(-2 pi i k / L) in FP32, rounded from an FP64 evaluation.
// Entry 0 is unused (set to 1).

// Launchers (return hipSuccess or an error; never allocate).
hipError_t launch_twiddle_pyramid(float2* table, int log_lmax, hipStream_t stream);

// Batched transform of `count` signals of size n (n >= 1, power of two, n <= the
// kind's limit).  Pointers are device memory with the layouts of ckfft_gpu.h.
hipError_t launch_transform(Kind kind, int log2n, int count, const void* in, void* out,
                            const float2* pyramid, int log_pyramid, hipStream_t stream);

// Largest log2(n) launch_transform supports for a kind.
int max_log2n(Kind kind);

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
    // Staging for host-pointer calls (grown on demand, guarded by staging_mutex).
    std::mutex* staging_mutex;
    void* d_staging;
    size_t staging_bytes;
};

constexpr uint64_t kContextMagic = 0x434b46465447505bull;  // "CKFFTGP["
