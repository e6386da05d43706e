// fft_pass.hip -- multi-pass transforms for sizes above one workgroup.
//
// Replaces the reference's recursion (src/ckfft/fft_default.cpp:168-266) for sizes a
// workgroup cannot hold; the reference has no such split.  The transform is a
// Stockham FFT over global memory with large radices: a size N = R1 * R2 (* R3) runs
// as 2 (3) passes, pass p being one Stockham stage of radix R_p (64..4096) and span
// NS_p = R_1 * ... * R_{p-1}:
//   butterfly j (0 <= j < N/R) reads x[j + r*N/R], multiplies point r by
//   W_{NS*R}^{r*(j mod NS)} (none in the first pass), takes an R-point DFT and writes
//   y[(j/NS)*NS*R + j%NS + r*NS].
// A workgroup owns CB consecutive butterflies (a "tile") and runs their R-point DFTs
// with the register/LDS Stockham of fft_stockham.h.  Loads read CB-point contiguous
// runs from R rows; the first pass writes one contiguous block of CB*R points, later
// passes write CB-point runs.  Tiles are dealt to the 8 XCDs in contiguous ranges so
// that tiles sharing cache lines share an L2.
//
// Real transforms above 2^16 samples run the half-size complex transform here plus a
// separate pack (forward) or unpack (inverse) pass; below that they stay fused in the
// one-workgroup kernels.
#include <utility>

#include "fft_pass.h"

namespace ckfft_gpu {

template <int LOGR, bool INV, bool FIRST>
__global__ __launch_bounds__(PassShape<LOGR>::WG) void fft_pass_kernel(
    const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ pyr,
    const float2* __restrict__ fine, const float2* __restrict__ xtw, int log_xtw, int logn, int logns,
    long long tiles_per_signal, long long total_tiles)
{
    using SH = PassShape<LOGR>;
    __shared__ float lds[SH::CB * SH::RSTRIDE];
    const long long tile = xcd_tile(blockIdx.x, total_tiles);
    const long long g = tile / tiles_per_signal;
    const long long q = tile - g * tiles_per_signal;
    const long long N = 1ll << logn;
    // the transposed twiddles cover span 2^kLogXtwCols at levels <= 2^16: radix <= 2^8
    constexpr int XTW = LOGR + kLogXtwCols <= kLogMaxPyramid ? 1 : 0;
    pass_tile<LOGR, SH::LOGCB, INV, FIRST, XTW>(in + g * N, out + g * N, q, logn, logns, pyr, fine, xtw, log_xtw,
                                                lds, threadIdx.x);
}

// ---- real pack / unpack for sizes above the one-workgroup limit ----------------------

// Real forward: Z (h complex points per signal, contiguous) -> out (h+1 per signal).
// Thread k packs the pair (k, h-k), k in [0, h/2], as the fused epilogue does.
__global__ __launch_bounds__(256) void real_pack_kernel(const float2* __restrict__ z, float2* __restrict__ out,
                                                        const float2* __restrict__ pyr,
                                                        const float2* __restrict__ fine, int logn, long long pairs,
                                                        long long count)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= pairs * count) return;
    const long long g = e / pairs;
    const long long k = e - g * pairs;
    const long long h = 1ll << (logn - 1);
    const float2* zs = z + g * h;
    float2* o = out + g * (h + 1);
    const float2 zk = zs[k & (h - 1)], zm = zs[(h - k) & (h - 1)];
    const float2 a = make_float2(zk.x + zm.x, zk.y - zm.y);
    const float2 d = make_float2(zk.x - zm.x, zk.y + zm.y);
    const float2 c = cmul(twiddle_at(pyr, fine, logn, k), d);
    o[k] = make_float2(a.x + c.y, a.y - c.x);
    if (k != h - k) o[h - k] = make_float2(a.x - c.y, -(a.y + c.x));
}

// Real inverse: Y (h+1 per signal) -> t (h complex points per signal).  Thread k unpacks
// the pair (k, h-k), k in [0, h/2], reading each Y value once; each point is formed from
// its own value, its mirror and its own twiddle.
__device__ __forceinline__ float2 unpack_point(float2 yk, float2 ym, float2 w)
{
    const float2 a = make_float2(yk.x + ym.x, yk.y - ym.y);
    const float2 d = make_float2(yk.x - ym.x, yk.y + ym.y);
    const float2 c = cmul(cconj(w), d);
    return make_float2(a.x - c.y, a.y + c.x);
}

__global__ __launch_bounds__(256) void real_unpack_kernel(const float2* __restrict__ y, float2* __restrict__ t,
                                                          const float2* __restrict__ pyr,
                                                          const float2* __restrict__ fine, int logn, long long pairs,
                                                          long long count)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= pairs * count) return;
    const long long g = e / pairs;
    const long long k = e - g * pairs;
    const long long h = 1ll << (logn - 1);
    const float2* ys = y + g * (h + 1);
    float2* ts = t + g * h;
    const float2 yk = ys[k], ym = ys[h - k];
    ts[k] = unpack_point(yk, ym, twiddle_at(pyr, fine, logn, k));
    if (k != 0 && k != h - k) ts[h - k] = unpack_point(ym, yk, twiddle_at(pyr, fine, logn, h - k));
}

// ---- host side ----------------------------------------------------------------------

using PassFn = void (*)(const float2*, float2*, const float2*, const float2*, const float2*, int, int, int, long long,
                        long long);

struct PassCfg {
    PassFn fn;
    int wg;
    int cb;
};

template <int LOGR, bool INV, bool FIRST>
constexpr PassCfg pass_cfg()
{
    return PassCfg{fft_pass_kernel<LOGR, INV, FIRST>, PassShape<LOGR>::WG, PassShape<LOGR>::CB};
}

template <bool INV, bool FIRST, int... L>
constexpr void fill_pass(PassCfg* tab, std::integer_sequence<int, L...>)
{
    ((tab[L + kLogMinPass] = pass_cfg<L + kLogMinPass, INV, FIRST>()), ...);
}

struct PassTables {
    PassCfg t[2][2][kLogMaxPass + 1];     // [inverse][first][log2 radix]
    PassTables()
    {
        constexpr int NP = kLogMaxPass - kLogMinPass + 1;
        fill_pass<false, false>(t[0][0], std::make_integer_sequence<int, NP>{});
        fill_pass<false, true>(t[0][1], std::make_integer_sequence<int, NP>{});
        fill_pass<true, false>(t[1][0], std::make_integer_sequence<int, NP>{});
        fill_pass<true, true>(t[1][1], std::make_integer_sequence<int, NP>{});
    }
};

static const PassTables& pass_tables()
{
    static const PassTables tables;
    return tables;
}

// Split log2 N into pass radices (first pass first), each in [kLogMinPass, kLogMaxPass].
// Two passes up to 2^22; above that three passes of radix <= 2^10 measured faster on
// MI355X (2^24: 8+8+8 in 210 us against 12+12 in 239 us): radix-4096 tiles hold only
// 4 columns, so their loads use 32-byte segments.
int plan_passes(int logc, int* logr)
{
    const int np = logc <= 22 ? 2 : 3;
    if (logc < np * kLogMinPass || logc > np * kLogMaxPass) return 0;
    for (int p = 0; p < np; ++p) logr[p] = logc / np + (p < logc % np ? 1 : 0);
    return np;
}

static hipError_t launch_pass(int logr, bool inv, bool first, const float2* in, float2* out, const DeviceTables& tb,
                              int logc, int logns, long long count, hipStream_t s)
{
    const PassCfg& c = pass_tables().t[inv][first][logr];
    const long long tiles = ((1ll << logc) >> logr) / c.cb;
    const long long total = tiles * count;
    if (total > 0x7fffffffll) return hipErrorInvalidValue;
    hipLaunchKernelGGL(c.fn, dim3((unsigned)total), dim3(c.wg), 0, s, in, out, tb.pyramid, tb.fine, tb.xtw, tb.log_xtw,
                       logc, logns, tiles, total);
    return hipGetLastError();
}

// Complex transform of `count` signals of 2^logc points through `np` passes, from `in`
// to `out`, using `tmp` (same size) for the intermediate(s).
static hipError_t run_passes(bool inv, int logc, long long count, const float2* in, float2* out, float2* tmp,
                             const DeviceTables& tb, hipStream_t s)
{
    int logr[3];
    const int np = plan_passes(logc, logr);
    if (!np) return hipErrorInvalidValue;
    // alternate buffers so that the last pass lands in `out`
    const float2* src = in;
    int logns = 0;
    for (int p = 0; p < np; ++p) {
        float2* dst = ((np - 1 - p) % 2 == 0) ? out : tmp;
        hipError_t e = launch_pass(logr[p], inv, p == 0, src, dst, tb, logc, logns, count, s);
        if (e != hipSuccess) return e;
        src = dst;
        logns += logr[p];
    }
    return hipSuccess;
}

hipError_t launch_multipass(Kind kind, int log2n, int count, const void* in, void* out, const DeviceTables& tb,
                            hipStream_t s)
{
    const bool real = (kind == kRealForward || kind == kRealInverse);
    const int logc = real ? log2n - 1 : log2n;
    int logr[3];
    if (!plan_passes(logc, logr)) return hipErrorInvalidValue;
    if (log2n > tb.log_max) return hipErrorInvalidValue;
    const long long pts = (1ll << logc) * (long long)count;
    float2* tmp = nullptr;
    // real kinds need one more signal-sized buffer for the packed/unpacked half spectrum
    const size_t tmp_bytes = sizeof(float2) * pts * (real ? 2 : 1);
    hipError_t e = tb.pool ? hipMallocFromPoolAsync((void**)&tmp, tmp_bytes, tb.pool, s)
                           : hipMallocAsync((void**)&tmp, tmp_bytes, s);
    if (e != hipSuccess) return e;
    const int threads = 256;
    if (kind == kComplexForward || kind == kComplexInverse) {
        e = run_passes(kind == kComplexInverse, logc, count, (const float2*)in, (float2*)out, tmp, tb, s);
    } else if (kind == kRealForward) {
        float2* z = tmp + pts;
        e = run_passes(false, logc, count, (const float2*)in, z, tmp, tb, s);
        if (e == hipSuccess) {
            const long long pairs = (1ll << (logc - 1)) + 1;
            const long long n_thr = pairs * count;
            hipLaunchKernelGGL(real_pack_kernel, dim3((unsigned)((n_thr + threads - 1) / threads)), dim3(threads), 0,
                               s, z, (float2*)out, tb.pyramid, tb.fine, log2n, pairs, (long long)count);
            e = hipGetLastError();
        }
    } else {
        float2* t = tmp + pts;
        const long long pairs = (1ll << (logc - 1)) + 1;    // k = 0 .. h/2
        const long long n_thr = pairs * count;
        hipLaunchKernelGGL(real_unpack_kernel, dim3((unsigned)((n_thr + threads - 1) / threads)), dim3(threads), 0, s,
                           (const float2*)in, t, tb.pyramid, tb.fine, log2n, pairs, (long long)count);
        e = hipGetLastError();
        if (e == hipSuccess) e = run_passes(true, logc, count, t, (float2*)out, tmp, tb, s);
    }
    hipError_t e2 = hipFreeAsync(tmp, s);
    return e != hipSuccess ? e : e2;
}

}  // namespace ckfft_gpu
