// fft_misc.hip -- twiddle generation, closed forms for the smallest sizes, dispatch.
#include <utility>

#include "ckfft_internal.h"
#include "fft_device.h"

namespace ckfft_gpu {

hipError_t launch_lds(Kind kind, int logc, int count, const void* in, void* out, const float2* pyramid,
                      hipStream_t stream);

// ---- twiddle pyramid -------------------------------------------------------------
//
// Replaces the reference's table loop (src/ckfft/context.cpp:90-105), which forms
// the angle in FP32 and calls cosf/sinf.  Here every entry is evaluated in FP64 with
// an exact argument reduction and rounded once to FP32, so each stored value is the
// correctly rounded exp(-2 pi i k / L) (up to FP64's own sub-ulp error):
//   x = 2k/L (exact, L a power of two), n = round(2x), d = x - n/2 in [-1/4, 1/4]
//   (exact), sincospi(d) in FP64, then a rotation by n quarter turns (exact).
// Quarter-turn entries such as L/4 come out exactly (0, -1).
__device__ float2 twiddle_fp64(long long k, int lg)
{
    const double x = ldexp((double)k, 1 - lg);                   // 2k/L in [0, 2), L = 2^lg
    const double nq = rint(2.0 * x);                             // quarter turns
    const double d = x - 0.5 * nq;                               // |d| <= 1/4, exact
    double s, c;
    sincospi(d, &s, &c);
    double sn, cs;                                               // sin(pi x), cos(pi x)
    switch (((int)nq) & 3) {
    case 0: sn = s; cs = c; break;
    case 1: sn = c; cs = -s; break;
    case 2: sn = -s; cs = -c; break;
    default: sn = -c; cs = s; break;
    }
    return make_float2((float)cs, (float)(-sn));
}

__global__ void twiddle_pyramid_kernel(float2* __restrict__ table, int log_lmax)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = 2ll << log_lmax;
    if (e >= total) return;
    if (e == 0) {
        table[0] = make_float2(1.f, 0.f);
        return;
    }
    const int lg = 63 - __builtin_clzll((unsigned long long)e);   // level L = 2^lg
    table[e] = twiddle_fp64(e - (1ll << lg), lg);
}

// Fine tables: level 2^l for l in (16, log_max], entries m < 2^12.
__global__ void twiddle_fine_kernel(float2* __restrict__ table, int log_max)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)(log_max - kLogMaxPyramid) << kLogFine;
    if (e >= total) return;
    const int lg = kLogMaxPyramid + 1 + (int)(e >> kLogFine);
    table[e] = twiddle_fp64(e & ((1 << kLogFine) - 1), lg);
}

hipError_t launch_twiddle_fine(float2* table, int log_max, hipStream_t stream)
{
    if (log_max <= kLogMaxPyramid) return hipSuccess;
    const long long total = (long long)(log_max - kLogMaxPyramid) << kLogFine;
    hipLaunchKernelGGL(twiddle_fine_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, table,
                       log_max);
    return hipGetLastError();
}

// Copies of pyramid entries (bit-identical values), laid out for the second pass of the
// two-pass transforms: table[r * 256 + j] = pyramid level L entry r*j.
__global__ void twiddle_transposed_kernel(float2* __restrict__ table, const float2* __restrict__ pyr, int log_l)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (1ll << log_l)) return;
    const long long r = e >> kLogXtwCols, j = e & ((1 << kLogXtwCols) - 1);
    table[e] = pyr[(1ll << log_l) + r * j];
}

hipError_t launch_twiddle_transposed(float2* table, const float2* pyramid, int log_l, hipStream_t stream)
{
    if (!table) return hipSuccess;
    hipLaunchKernelGGL(twiddle_transposed_kernel, dim3((unsigned)(((1ll << log_l) + 255) / 256)), dim3(256), 0,
                       stream, table, pyramid, log_l);
    return hipGetLastError();
}

hipError_t launch_twiddle_pyramid(float2* table, int log_lmax, hipStream_t stream)
{
    const long long total = 2ll << log_lmax;
    const int block = 256;
    const unsigned grid = (unsigned)((total + block - 1) / block);
    hipLaunchKernelGGL(twiddle_pyramid_kernel, dim3(grid), dim3(block), 0, stream, table, log_lmax);
    return hipGetLastError();
}

// ---- closed forms for complex n <= 2 and real n <= 4 --------------------------------
//
// Same formulas and operation order as the reference's special cases
// (src/ckfft/fft.cpp:20-31, src/ckfft/fft_real.cpp:20-48 and 69-93), so these
// sizes are bit-identical with it.  One thread per signal.
template <int KIND, int N>
__global__ __launch_bounds__(256) void tiny_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                   int count)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= count) return;
    if constexpr (KIND == kComplexForward || KIND == kComplexInverse) {
        const float2* x = (const float2*)in + g * N;
        float2* y = (float2*)out + g * N;
        if constexpr (N == 1) {
            y[0] = x[0];
        } else {
            const float2 a = x[0], b = x[1];
            y[0] = cadd(a, b);
            y[1] = csub(a, b);
        }
    } else if constexpr (KIND == kRealForward) {
        const float* x = in + g * N;
        float2* y = (float2*)out + g * (N / 2 + 1);
        if constexpr (N == 1) {
            y[0] = make_float2(x[0] * 2.0f, 0.0f);
        } else if constexpr (N == 2) {
            y[0] = make_float2((x[0] + x[1]) * 2.0f, 0.0f);
            y[1] = make_float2((x[0] - x[1]) * 2.0f, 0.0f);
        } else {
            const float s02 = (x[0] + x[2]) * 2.0f, d02 = (x[0] - x[2]) * 2.0f;
            const float s13 = (x[1] + x[3]) * 2.0f, d13 = (x[1] - x[3]) * 2.0f;
            y[0] = make_float2(s02 + s13, 0.0f);
            y[1] = make_float2(d02, -d13);
            y[2] = make_float2(s02 - s13, 0.0f);
        }
    } else {
        const float2* y = (const float2*)in + g * (N / 2 + 1);
        float* x = out + g * N;
        if constexpr (N == 1) {
            x[0] = y[0].x;
        } else if constexpr (N == 2) {
            x[0] = y[0].x + y[1].x;
            x[1] = y[0].x - y[1].x;
        } else {
            const float s02 = y[0].x + y[2].x;
            const float s13 = 2.0f * y[1].x;
            const float d02 = y[0].x - y[2].x;
            const float d13 = 2.0f * y[1].y;
            x[0] = s02 + s13;
            x[1] = d02 - d13;
            x[2] = s02 - s13;
            x[3] = d02 + d13;
        }
    }
}

template <int KIND, int N>
static hipError_t launch_tiny_t(int count, const void* in, void* out, hipStream_t stream)
{
    const unsigned grid = (unsigned)((count + 255) / 256);
    hipLaunchKernelGGL((tiny_kernel<KIND, N>), dim3(grid), dim3(256), 0, stream, (const float*)in, (float*)out,
                       count);
    return hipGetLastError();
}

static hipError_t launch_tiny(Kind kind, int log2n, int count, const void* in, void* out, hipStream_t stream)
{
    switch (kind) {
    case kComplexForward:
        return log2n == 0 ? launch_tiny_t<kComplexForward, 1>(count, in, out, stream)
                          : launch_tiny_t<kComplexForward, 2>(count, in, out, stream);
    case kComplexInverse:
        return log2n == 0 ? launch_tiny_t<kComplexInverse, 1>(count, in, out, stream)
                          : launch_tiny_t<kComplexInverse, 2>(count, in, out, stream);
    case kRealForward:
        return log2n == 0 ? launch_tiny_t<kRealForward, 1>(count, in, out, stream)
               : log2n == 1 ? launch_tiny_t<kRealForward, 2>(count, in, out, stream)
                            : launch_tiny_t<kRealForward, 4>(count, in, out, stream);
    default:
        return log2n == 0 ? launch_tiny_t<kRealInverse, 1>(count, in, out, stream)
               : log2n == 1 ? launch_tiny_t<kRealInverse, 2>(count, in, out, stream)
                            : launch_tiny_t<kRealInverse, 4>(count, in, out, stream);
    }
}

// ---- dispatch ---------------------------------------------------------------------------

int max_log2n(Kind) { return kLogMaxSupported; }

// The route launch_transform takes; true when it is the multi-pass one (which needs an
// intermediate).
static bool takes_multipass(Kind kind, int log2n, const DeviceTables& tb)
{
    const bool real = (kind == kRealForward || kind == kRealInverse);
    if (log2n <= (real ? 2 : 1)) return false;
    if (tb.precision == kCkFftPrecision_SplitBf16x3 && mfma_supported(kind, log2n)) return false;
    return (real ? log2n - 1 : log2n) > kLogMaxLds;
}

size_t transform_scratch_bytes(Kind kind, int log2n, int count, const DeviceTables& tb)
{
    return takes_multipass(kind, log2n, tb) ? multipass_scratch_bytes(kind, log2n, count) : 0;
}

hipError_t launch_transform(Kind kind, int log2n, int count, const void* in, void* out, const DeviceTables& tb,
                            hipStream_t stream, void* scratch, size_t scratch_bytes)
{
    if (log2n < 0 || log2n > max_log2n(kind) || log2n > tb.log_max || count < 0) return hipErrorInvalidValue;
    if (count == 0) return hipSuccess;
    const bool real = (kind == kRealForward || kind == kRealInverse);
    if (log2n <= (real ? 2 : 1)) return launch_tiny(kind, log2n, count, in, out, stream);
    // complex size 2^logc; the real kinds run a half-size complex transform and read
    // the twiddles of level 2^log2n for their pack/unpack
    const int logc = real ? log2n - 1 : log2n;
    if (tb.precision == kCkFftPrecision_SplitBf16x3 && mfma_supported(kind, log2n))
        return launch_mfma(kind, log2n, count, in, out, tb, stream);
    if (logc <= kLogMaxLds) {
        if (tb.log_pyramid < log2n) return hipErrorInvalidValue;
        return launch_lds(kind, logc, count, in, out, tb.pyramid, stream);
    }
    return launch_multipass(kind, log2n, count, in, out, tb, stream, scratch, scratch_bytes);
}

}  // namespace ckfft_gpu
