// fft_mfma.hip -- complex transforms in the split-BF16x3 precision mode (opt-in per
// context): radix-16 stages on the matrix cores (fft_mfma.h).
//
// Sizes 2^4 .. 2^14: one workgroup holds 4096 points or one whole signal (16384 points at
// most, 139 KB of LDS); the first stage reads global memory, the last writes it, the
// stages between exchange through LDS.
// Sizes 2^15 and 2^16: two Stockham passes N = R1 * R2 (R1 = 256) through global memory,
// each pass a tile of CB butterflies x R points in LDS (8192 points, 70 KB):
//   pass 1 (span 1):   column j reads x[j + r N/R1],              writes y[j R1 + k]
//   pass 2 (span R1):  column j reads y[j + r R1] * W_N^{r j},    writes y[j + k R1]
// Pass 2 reads exactly the points it writes within its tile and tiles are disjoint, so it
// runs in place in the output buffer and no scratch is needed.
#include "ckfft_internal.h"
#include "fft_mfma.h"
#include "fft_pass.h"

namespace ckfft_gpu {

constexpr int kLogMinMfma = 4;
constexpr int kLogMaxMfma = 16;
constexpr int kLogMfmaPassR1 = 8;

template <int LOGN>
struct WholeShape {
    static constexpr int LOGB = LOGN >= 12 ? 0 : 12 - LOGN;   // signals per workgroup (log2)
    static constexpr int LOGT = LOGN + LOGB;                   // points per workgroup (log2)
    static constexpr int WG = LOGT >= 13 ? 512 : 256;
    using TL = MfmaTile<LOGN, LOGB>;
};

// The first stage reads the signals straight from global memory and the last writes
// them back; the stages between exchange through the LDS tile.
template <int LOGN, bool INV>
__global__ __launch_bounds__(WholeShape<LOGN>::WG) void fft_mfma_whole_kernel(const float2* __restrict__ in,
                                                                              float2* __restrict__ out,
                                                                              const float2* __restrict__ pyr,
                                                                              long long count)
{
    using SH = WholeShape<LOGN>;
    using TL = typename SH::TL;
    constexpr int STAGES = (LOGN % 4 ? 1 : 0) + LOGN / 4;
    __shared__ float2 lds[STAGES > 1 ? TL::WORDS : 1];
    const SignalIO<LOGN> io{in, out, (long long)blockIdx.x << SH::LOGB, count};
    mfma_columns_fft<TL, SH::WG, INV>(io, LdsIO<TL>{lds}, io, pyr, threadIdx.x);
}

template <int LOGR>
struct MfmaPassShape {
    static constexpr int LOGCB = 13 - LOGR;                    // 8192 points per tile
    static constexpr int WG = 512;
    using TL = MfmaTile<LOGR, LOGCB>;
};

// One pass tile; FIRST: span 1 (no twiddle, contiguous output block), else span N/R with
// the inter-pass twiddle W_N^{r j} on load.  In and out may alias in the second pass.
template <int LOGR, bool INV, bool FIRST>
__global__ __launch_bounds__(512) void fft_mfma_pass_kernel(const float2* in, float2* out,
                                                            const float2* __restrict__ pyr,
                                                            const float2* __restrict__ xtw, int log_xtw, int logn,
                                                            long long tiles_per_signal, long long total_tiles)
{
    using SH = MfmaPassShape<LOGR>;
    using TL = typename SH::TL;
    constexpr int R = 1 << LOGR, CB = 1 << SH::LOGCB, T = R * CB, WG = SH::WG;
    __shared__ float2 lds[TL::WORDS];
    const int tid = threadIdx.x;
    const long long tile = xcd_tile(blockIdx.x, total_tiles);
    const long long g = tile / tiles_per_signal;
    const long long j0 = (tile - g * tiles_per_signal) * CB;
    const long long N = 1ll << logn;
    const long long stride = N >> LOGR;                        // N/R: point stride of a column
    const float2* src = in + g * N;
    float2* dst = out + g * N;
#pragma unroll 4
    for (int e = tid; e < T; e += WG) {
        const int r = e >> SH::LOGCB, b = e & (CB - 1);
        const long long j = j0 + b;
        float2 v = ld_in<true>(&src[j + r * stride]);   // read once
        if constexpr (!FIRST) {
            // W_N^{r j}: the transposed copy (columns side by side) when the context has it
            const float2 w = xtw ? xtw[((long long)r << (log_xtw - logn)) * (1 << kLogXtwCols) + j] : pyr[N + r * j];
            v = cmul(v, dirw<INV>(w));
        }
        lds[tile_at<TL>(b, r)] = v;
    }
    __syncthreads();
    const LdsIO<TL> lio{lds};
    mfma_columns_fft<TL, WG, INV>(lio, lio, lio, pyr, tid);
    if constexpr (FIRST) {
#pragma unroll 4
        for (int e = tid; e < T; e += WG) {
            const int b = e >> LOGR, k = e & (R - 1);
            dst[(j0 + b) * R + k] = lds[tile_at<TL>(b, k)];
        }
    } else {
#pragma unroll 4
        for (int e = tid; e < T; e += WG) {
            const int k = e >> SH::LOGCB, b = e & (CB - 1);
            st_out<true>(&dst[j0 + b + k * stride], lds[tile_at<TL>(b, k)]);   // the transform's output
        }
    }
}

// ---- host side ----------------------------------------------------------------------

using WholeFn = void (*)(const float2*, float2*, const float2*, long long);

template <int LOGN>
constexpr WholeFn whole_fn(bool inv)
{
    return inv ? fft_mfma_whole_kernel<LOGN, true> : fft_mfma_whole_kernel<LOGN, false>;
}

static WholeFn whole_kernel(int logn, bool inv, int* wg, int* logb)
{
    switch (logn) {
#define CKFFT_WHOLE(L)                       \
    case L:                                  \
        *wg = WholeShape<L>::WG;             \
        *logb = WholeShape<L>::LOGB;         \
        return whole_fn<L>(inv);
        CKFFT_WHOLE(4) CKFFT_WHOLE(5) CKFFT_WHOLE(6) CKFFT_WHOLE(7) CKFFT_WHOLE(8) CKFFT_WHOLE(9)
        CKFFT_WHOLE(10) CKFFT_WHOLE(11) CKFFT_WHOLE(12) CKFFT_WHOLE(13) CKFFT_WHOLE(14)
#undef CKFFT_WHOLE
    default:
        return nullptr;
    }
}

using PassFnM = void (*)(const float2*, float2*, const float2*, const float2*, int, int, long long, long long);

static PassFnM pass_kernel(int logr, bool inv, bool first)
{
    if (logr == 8)
        return first ? (inv ? fft_mfma_pass_kernel<8, true, true> : fft_mfma_pass_kernel<8, false, true>)
                     : (inv ? fft_mfma_pass_kernel<8, true, false> : fft_mfma_pass_kernel<8, false, false>);
    if (logr == 7 && !first)
        return inv ? fft_mfma_pass_kernel<7, true, false> : fft_mfma_pass_kernel<7, false, false>;
    return nullptr;
}

bool mfma_supported(Kind kind, int log2n)
{
    return (kind == kComplexForward || kind == kComplexInverse) && log2n >= kLogMinMfma && log2n <= kLogMaxMfma;
}

hipError_t launch_mfma(Kind kind, int log2n, int count, const void* in, void* out, const DeviceTables& tb,
                       hipStream_t s)
{
    if (!mfma_supported(kind, log2n) || log2n > tb.log_pyramid || count < 0) return hipErrorInvalidValue;
    if (count == 0) return hipSuccess;
    const bool inv = kind == kComplexInverse;
    if (log2n <= 14) {
        int wg = 0, logb = 0;
        const WholeFn fn = whole_kernel(log2n, inv, &wg, &logb);
        const long long groups = ((long long)count + (1ll << logb) - 1) >> logb;
        if (!fn || groups > 0x7fffffffll) return hipErrorInvalidValue;
        hipLaunchKernelGGL(fn, dim3((unsigned)groups), dim3(wg), 0, s, (const float2*)in, (float2*)out, tb.pyramid,
                           (long long)count);
        return hipGetLastError();
    }
    const int logr1 = kLogMfmaPassR1, logr2 = log2n - kLogMfmaPassR1;
    const PassFnM p1 = pass_kernel(logr1, inv, true), p2 = pass_kernel(logr2, inv, false);
    if (!p1 || !p2) return hipErrorInvalidValue;
    const long long tiles1 = ((1ll << log2n) >> logr1) >> MfmaPassShape<kLogMfmaPassR1>::LOGCB;
    const long long tiles2 = ((1ll << log2n) >> logr2) >> (13 - logr2);
    if (tiles1 * count > 0x7fffffffll || tiles2 * count > 0x7fffffffll) return hipErrorInvalidValue;
    // pass 2 has span 2^kLogMfmaPassR1 = 256 columns: the transposed twiddles apply
    static_assert(kLogMfmaPassR1 == kLogXtwCols, "pass-2 span must match the transposed table");
    const float2* xtw = tb.log_xtw >= log2n ? tb.xtw : nullptr;
    hipLaunchKernelGGL(p1, dim3((unsigned)(tiles1 * count)), dim3(512), 0, s, (const float2*)in, (float2*)out,
                       tb.pyramid, xtw, tb.log_xtw, log2n, tiles1, tiles1 * count);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(p2, dim3((unsigned)(tiles2 * count)), dim3(512), 0, s, (const float2*)out, (float2*)out,
                       tb.pyramid, xtw, tb.log_xtw, log2n, tiles2, tiles2 * count);
    return hipGetLastError();
}

}  // namespace ckfft_gpu
