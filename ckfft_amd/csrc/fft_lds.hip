// fft_lds.hip -- batch kernels for complex sizes up to 2^15 (real up to 2^16): one
// workgroup holds one or more whole transforms.  The transform body is in fft_lds.h.
#include <utility>

#include "fft_lds.h"

namespace ckfft_gpu {

// ---- the kernel ----------------------------------------------------------------------

// `count` contiguous signals; workgroup slot g handles signal blockIdx.x * TPW + slot.
template <int LOGN, int LOGP, int WG, int KIND, bool SPLIT>
__global__ __launch_bounds__(WG) void fft_lds_kernel(const float2* __restrict__ in, float2* __restrict__ out,
                                                     const float2* __restrict__ tw, int count)
{
    using SH = LdsShape<LOGN, LOGP, WG, KIND, SPLIT>;
    __shared__ __align__(16) float lds[SH::LDS_WORDS];   // 16-byte LDS reads of the pack staging
    // one transform per workgroup: its index is uniform, so the signal's base address
    // stays in scalar registers
    const int tr = SH::TPW == 1 ? 0 : (int)threadIdx.x / SH::T;
    const long long g = (long long)blockIdx.x * SH::TPW + tr;
    const bool active = g < count;
    // slots past the last signal (only in the last workgroup) compute on the last
    // signal and store nothing
    const long long gl = active ? g : (long long)count - 1;
    lds_transform<LOGN, LOGP, WG, KIND, SPLIT>(in + gl * SH::IN_STRIDE, out + gl * SH::OUT_STRIDE, active, tw, lds,
                                               threadIdx.x);
}

// One transform per workgroup with the exchange one LDS plane at a time, held to MINW
// waves per SIMD so that several workgroups share a CU: one workgroup's loads and stores
// then overlap another's arithmetic.  The grid is exactly `count` workgroups.
template <int LOGN, int LOGP, int KIND, int MINW>
__global__ __launch_bounds__((1 << LOGN) >> LOGP, MINW) void fft_lds_shared_cu_kernel(
    const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ tw, int count)
{
    constexpr int WG = (1 << LOGN) >> LOGP;
    using SH = LdsShape<LOGN, LOGP, WG, KIND, true>;
    __shared__ __align__(16) float lds[SH::LDS_WORDS];
    (void)count;
    const long long g = (long long)blockIdx.x;
    lds_transform<LOGN, LOGP, WG, KIND, true>(in + g * SH::IN_STRIDE, out + g * SH::OUT_STRIDE, true, tw, lds,
                                              threadIdx.x);
}

// ---- host-side dispatch ------------------------------------------------------------

using KernelFn = void (*)(const float2*, float2*, const float2*, int);

struct LaunchCfg {
    KernelFn fn;
    int wg;        // threads per workgroup
    int tpw;       // transforms per workgroup
};

template <int LOGN, int KIND>
constexpr LaunchCfg make_cfg()
{
    // points per thread: whole transform for N <= 16, 16 up to 2^13, 32 at 2^14, 64 at
    // 2^15 (512 threads x 64 points measured faster than 1024 x 32 on MI355X).
    // 2^13 and 2^14 exchange one LDS plane at a time (34 / 68 KiB), so that up to four /
    // two workgroups share a CU: batches of 256 MiB per buffer took 98-112 us against
    // 111-122 us (complex 2^13) and 109-124 us against 124-137 us (complex 2^14) with both
    // planes in LDS and one or two workgroups per CU.  The radix plan, and so every bit of
    // the result, is the same.
    if constexpr (LOGN == 13) {
        return LaunchCfg{fft_lds_shared_cu_kernel<13, 4, KIND, 6>, 512, 1};
    } else if constexpr (LOGN == 14) {
        return LaunchCfg{fft_lds_shared_cu_kernel<14, 5, KIND, 4>, 512, 1};
    } else {
        constexpr int LOGP = LOGN <= 4 ? LOGN : (LOGN <= 14 ? 4 : 6);
        constexpr int T = (1 << LOGN) >> LOGP;
        constexpr int WG = T >= 256 ? T : 256;
        constexpr bool SPLIT = LOGN >= 15;
        return LaunchCfg{fft_lds_kernel<LOGN, LOGP, WG, KIND, SPLIT>, WG, WG / T};
    }
}

template <int KIND, int... L>
constexpr void fill_table(LaunchCfg* tab, std::integer_sequence<int, L...>)
{
    ((tab[L + 1] = make_cfg<L + 1, KIND>()), ...);
}

struct LdsTables {
    LaunchCfg t[4][kLogMaxLds + 1];
    LdsTables()
    {
        for (auto& row : t)
            for (auto& c : row) c = LaunchCfg{nullptr, 0, 0};
        fill_table<kComplexForward>(t[kComplexForward], std::make_integer_sequence<int, kLogMaxLds>{});
        fill_table<kComplexInverse>(t[kComplexInverse], std::make_integer_sequence<int, kLogMaxLds>{});
        fill_table<kRealForward>(t[kRealForward], std::make_integer_sequence<int, kLogMaxLds>{});
        fill_table<kRealInverse>(t[kRealInverse], std::make_integer_sequence<int, kLogMaxLds>{});
    }
};

static const LdsTables& lds_tables()
{
    static const LdsTables tables;
    return tables;
}

// Launch the one-workgroup kernel for complex size 2^logc (real kinds: half size).
hipError_t launch_lds(Kind kind, int logc, int count, const void* in, void* out, const float2* pyramid,
                      hipStream_t stream)
{
    if (logc < 1 || logc > kLogMaxLds) return hipErrorInvalidValue;
    const LaunchCfg& c = lds_tables().t[kind][logc];
    if (!c.fn) return hipErrorInvalidValue;
    if (count <= 0) return hipSuccess;
    const unsigned grid = (unsigned)((count + c.tpw - 1) / c.tpw);
    hipLaunchKernelGGL(c.fn, dim3(grid), dim3(c.wg), 0, stream, (const float2*)in, (float2*)out, pyramid,
                       count);
    return hipGetLastError();
}

}  // namespace ckfft_gpu
