// fft_lds.hip -- one-workgroup Stockham transforms for complex sizes up to 2^15.
//
// Replaces the reference's recursive radix-4 decimation-in-time transform
// (src/ckfft/fft_default.cpp:12-266) and its real pack/unpack passes
// (src/ckfft/fft_real_default.cpp:121-222) for every size that fits one workgroup.
//
// Design (gfx950, wave64):
//  * Each transform of N complex points is owned by T = N/P threads; every thread
//    keeps P points in registers for the whole transform.  Stages are radix 16
//    (the first stage takes the remaining radix 2/4/8), so a 4096-point transform is
//    three register-resident radix-16 stages with two LDS exchanges between them.
//  * Stockham autosort indexing: stage s with radix R and span Ns reads element
//    j + r*N/R and writes (j/Ns)*Ns*R + j%Ns + r*Ns, so the first stage reads global
//    memory and the last stage writes it in natural order, both fully coalesced
//    (consecutive lanes touch consecutive 8-byte elements).
//  * LDS holds real and imaginary parts in separate padded planes (one pad word per
//    16) so the strided first-stage writes do not collide on LDS banks.  When both
//    planes do not fit (N = 2^15: 2 x 136 KiB), the exchange runs plane by plane
//    through one 136 KiB buffer.
//  * Twiddles come from the FP64-generated pyramid table (level L holds
//    exp(-2 pi i k/L)); stage s uses level Ns*R, read through the caches.
//  * Real forward (2N reals): an N-point complex transform of the samples viewed as
//    complex, then the pack pass out[k] = A - i W^k D, out[N-k] = conj(A + i W^k D)
//    with A = Z[k] + conj Z[N-k], D = Z[k] - conj Z[N-k], W = exp(-2 pi i/2N), fused
//    after the last stage.  This is the reference's post-pass
//    (src/ckfft/fft_real_default.cpp:129-170) written for one (k, N-k) pair per step.
//  * Real inverse: the unpack pre-pass t[k] = (Y[k] + conj Y[N-k]) +
//    i conj(W^k) (Y[k] - conj Y[N-k]) (src/ckfft/fft_real_default.cpp:180-219) is
//    fused into the first stage's global loads, then the N-point inverse transform
//    writes the 2N reals as N complex values.
#include "ckfft_internal.h"
#include "fft_device.h"

namespace ckfft_gpu {

// Padded LDS index: one unused word after every 16.
__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 4); }
constexpr int lds_span(int n) { return n + (n >> 4); }

template <int LOGN, int LOGP>
struct Plan {
    static constexpr int N = 1 << LOGN;
    static constexpr int P = 1 << LOGP;          // points per thread
    static constexpr int T = N / P;              // threads per transform
    static constexpr int RMAX_LOG = LOGP < 4 ? LOGP : 4;
    static constexpr int NSTAGES = (LOGN + RMAX_LOG - 1) / RMAX_LOG;
    static constexpr int FIRST_LOG = LOGN - (NSTAGES - 1) * RMAX_LOG;
    static constexpr int rlog(int s) { return s == 0 ? FIRST_LOG : RMAX_LOG; }
    static constexpr int nslog(int s) { return s == 0 ? 0 : FIRST_LOG + (s - 1) * RMAX_LOG; }
    static_assert(LOGP <= LOGN, "more points per thread than points");
    static_assert(FIRST_LOG >= 1 && FIRST_LOG <= RMAX_LOG, "bad radix plan");
};

// ---- stage arithmetic --------------------------------------------------------

template <class PL, int S, bool INV>
__device__ __forceinline__ void stage_compute(float2* v, int t, const float2* __restrict__ tw)
{
    constexpr int R = 1 << PL::rlog(S);
    constexpr int NS = 1 << PL::nslog(S);
    constexpr int B = PL::P / R;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        if constexpr (NS > 1) {
            const int k = (t + b * PL::T) & (NS - 1);
            const float2* lvl = tw + NS * R;     // level NS*R of the pyramid
#pragma unroll
            for (int r = 1; r < R; ++r)
                v[b * R + r] = cmul(v[b * R + r], dirw<INV>(lvl[r * k]));
        }
        dft<R, INV>(v + b * R);
    }
}

// Index of element r of butterfly j at the input of stage S.
template <class PL, int S>
__device__ __forceinline__ int in_index(int j, int r)
{
    return j + r * (PL::N >> PL::rlog(S));
}

// Index of element r of butterfly j at the output of stage S.
template <class PL, int S>
__device__ __forceinline__ int out_index(int j, int r)
{
    constexpr int R = 1 << PL::rlog(S);
    constexpr int NS = 1 << PL::nslog(S);
    return (j / NS) * NS * R + (j & (NS - 1)) + r * NS;
}

// COMP: 1 = real plane, 2 = imaginary plane, 3 = both.
template <class PL, int S, int COMP>
__device__ __forceinline__ void lds_write(const float2* v, int t, float* re, float* im)
{
    constexpr int R = 1 << PL::rlog(S);
    constexpr int B = PL::P / R;
#pragma unroll
    for (int b = 0; b < B; ++b)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = lds_pad(out_index<PL, S>(t + b * PL::T, r));
            if constexpr (COMP & 1) re[i] = v[b * R + r].x;
            if constexpr (COMP & 2) im[i] = v[b * R + r].y;
        }
}

template <class PL, int S, int COMP>
__device__ __forceinline__ void lds_read(float2* v, int t, const float* re, const float* im)
{
    constexpr int R = 1 << PL::rlog(S);
    constexpr int B = PL::P / R;
#pragma unroll
    for (int b = 0; b < B; ++b)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = lds_pad(in_index<PL, S>(t + b * PL::T, r));
            if constexpr (COMP & 1) v[b * R + r].x = re[i];
            if constexpr (COMP & 2) v[b * R + r].y = im[i];
        }
}

// Move data from the output layout of stage SW to the input layout of stage SR.
template <class PL, int SW, int SR, bool SPLIT>
__device__ __forceinline__ void exchange(float2* v, int t, float* re, float* im)
{
    if constexpr (!SPLIT) {
        lds_write<PL, SW, 3>(v, t, re, im);
        __syncthreads();
        lds_read<PL, SR, 3>(v, t, re, im);
        __syncthreads();
    } else {
        lds_write<PL, SW, 1>(v, t, re, im);
        __syncthreads();
        lds_read<PL, SR, 1>(v, t, re, im);
        __syncthreads();
        lds_write<PL, SW, 2>(v, t, re, im);
        __syncthreads();
        lds_read<PL, SR, 2>(v, t, re, im);
        __syncthreads();
    }
}

template <class PL, int S, bool INV, bool SPLIT>
__device__ __forceinline__ void run_stages(float2* v, int t, float* re, float* im,
                                           const float2* __restrict__ tw)
{
    if constexpr (S < PL::NSTAGES) {
        exchange<PL, S - 1, S, SPLIT>(v, t, re, im);
        stage_compute<PL, S, INV>(v, t, tw);
        run_stages<PL, S + 1, INV, SPLIT>(v, t, re, im, tw);
    }
}

// ---- first-stage loads ---------------------------------------------------------

template <class PL>
__device__ __forceinline__ void load_first_complex(float2* v, int t, const float2* __restrict__ src,
                                                   bool active)
{
    constexpr int R = 1 << PL::rlog(0);
    constexpr int B = PL::P / R;
#pragma unroll
    for (int b = 0; b < B; ++b)
#pragma unroll
        for (int r = 0; r < R; ++r)
            v[b * R + r] = active ? src[in_index<PL, 0>(t + b * PL::T, r)] : make_float2(0.f, 0.f);
}

// Real inverse: unpack the N+1 half-spectrum values into the N-point complex input.
template <class PL>
__device__ __forceinline__ void load_first_unpack(float2* v, int t, const float2* __restrict__ y,
                                                  const float2* __restrict__ tw, bool active)
{
    constexpr int N = PL::N;
    constexpr int R = 1 << PL::rlog(0);
    constexpr int B = PL::P / R;
    const float2* lvl = tw + 2 * N;             // level 2N: exp(-2 pi i k / 2N)
#pragma unroll
    for (int b = 0; b < B; ++b)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float2 out = make_float2(0.f, 0.f);
            if (active) {
                const int k = in_index<PL, 0>(t + b * PL::T, r);
                const float2 yk = y[k], ym = y[N - k];
                const float2 a = make_float2(yk.x + ym.x, yk.y - ym.y);     // Y[k] + conj Y[N-k]
                const float2 d = make_float2(yk.x - ym.x, yk.y + ym.y);     // Y[k] - conj Y[N-k]
                const float2 c = cmul(cconj(lvl[k]), d);                    // exp(+2 pi i k/2N) d
                out = make_float2(a.x - c.y, a.y + c.x);                    // a + i c
            }
            v[b * R + r] = out;
        }
}

// ---- last-stage stores -----------------------------------------------------------

template <class PL>
__device__ __forceinline__ void store_last(const float2* v, int t, float2* __restrict__ dst, bool active)
{
    constexpr int S = PL::NSTAGES - 1;
    constexpr int R = 1 << PL::rlog(S);
    constexpr int B = PL::P / R;
    if (!active) return;
#pragma unroll
    for (int b = 0; b < B; ++b)
#pragma unroll
        for (int r = 0; r < R; ++r)
            dst[out_index<PL, S>(t + b * PL::T, r)] = v[b * R + r];
}

// Pack one (k, N-k) pair of the real forward transform; zk = Z[k], zm = Z[N-k mod N].
template <int N>
__device__ __forceinline__ void pack_pair(int k, float2 zk, float2 zm, const float2* __restrict__ lvl,
                                          float2* __restrict__ dst)
{
    const float2 a = make_float2(zk.x + zm.x, zk.y - zm.y);
    const float2 d = make_float2(zk.x - zm.x, zk.y + zm.y);
    const float2 c = cmul(lvl[k], d);                       // W^k D
    dst[k] = make_float2(a.x + c.y, a.y - c.x);             // a - i c
    if (k != N - k) dst[N - k] = make_float2(a.x - c.y, -(a.y + c.x));   // conj(a + i c)
}

// Real forward epilogue when one transform spans several threads: Z goes through
// LDS, then thread t packs pairs k = t + q*T for k in [0, N/2].
template <class PL, bool SPLIT>
__device__ __forceinline__ void pack_epilogue(float2* v, int t, float* re, float* im,
                                              const float2* __restrict__ tw, float2* __restrict__ dst,
                                              bool active)
{
    constexpr int N = PL::N;
    constexpr int T = PL::T;
    constexpr int S = PL::NSTAGES - 1;
    constexpr int Q = (N / 2) / T + 1;                      // pair slots per thread
    const float2* lvl = tw + 2 * N;
    float2 zk[Q], zm[Q];
    if constexpr (!SPLIT) {
        lds_write<PL, S, 3>(v, t, re, im);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int k = t + q * T;
            if (k <= N / 2) {
                const int ik = lds_pad(k & (N - 1)), im_ = lds_pad((N - k) & (N - 1));
                zk[q] = make_float2(re[ik], im[ik]);
                zm[q] = make_float2(re[im_], im[im_]);
            }
        }
    } else {
        lds_write<PL, S, 1>(v, t, re, im);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int k = t + q * T;
            if (k <= N / 2) {
                zk[q].x = re[lds_pad(k & (N - 1))];
                zm[q].x = re[lds_pad((N - k) & (N - 1))];
            }
        }
        __syncthreads();
        lds_write<PL, S, 2>(v, t, re, im);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int k = t + q * T;
            if (k <= N / 2) {
                zk[q].y = im[lds_pad(k & (N - 1))];
                zm[q].y = im[lds_pad((N - k) & (N - 1))];
            }
        }
    }
    if (!active) return;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int k = t + q * T;
        if (k <= N / 2) pack_pair<N>(k, zk[q], zm[q], lvl, dst);
    }
}

// ---- the kernel ----------------------------------------------------------------------

// KIND: kComplexForward / kComplexInverse / kRealForward / kRealInverse.
// For the real kinds N = 2^LOGN is half the real length.
template <int LOGN, int LOGP, int WG, int KIND, bool SPLIT>
__global__ __launch_bounds__(WG) void fft_lds_kernel(const float2* __restrict__ in, float2* __restrict__ out,
                                                     const float2* __restrict__ tw, int count)
{
    using PL = Plan<LOGN, LOGP>;
    constexpr bool INV = (KIND == kComplexInverse || KIND == kRealInverse);
    constexpr int N = PL::N;
    constexpr int T = PL::T;
    constexpr int TPW = WG / T;                 // transforms per workgroup
    static_assert(TPW >= 1 && TPW * T == WG, "workgroup must hold whole transforms");
    static_assert(!SPLIT || TPW == 1, "plane-by-plane exchange is for one transform per workgroup");
    constexpr bool NEED_LDS = (T > 1);
    constexpr int SPAN = lds_span(N);
    constexpr int LDS_WORDS = NEED_LDS ? (SPLIT ? SPAN : 2 * SPAN * TPW) : 1;
    __shared__ float lds[LDS_WORDS];

    const int tr = threadIdx.x / T;
    const int t = threadIdx.x - tr * T;
    const long long g = (long long)blockIdx.x * TPW + tr;
    const bool active = g < count;
    float* re = lds + (SPLIT ? 0 : tr * 2 * SPAN);
    float* im = SPLIT ? lds : re + SPAN;

    float2 v[PL::P];
    if constexpr (KIND == kRealInverse)
        load_first_unpack<PL>(v, t, in + g * (N + 1), tw, active);
    else
        load_first_complex<PL>(v, t, in + g * N, active);

    stage_compute<PL, 0, INV>(v, t, tw);
    run_stages<PL, 1, INV, SPLIT>(v, t, re, im, tw);

    if constexpr (KIND == kRealForward) {
        float2* dst = out + g * (N + 1);
        if constexpr (T == 1) {
            // whole transform in this thread's registers, in natural order
            if (active) {
                const float2* lvl = tw + 2 * N;
#pragma unroll
                for (int k = 0; k <= N / 2; ++k) pack_pair<N>(k, v[k & (N - 1)], v[(N - k) & (N - 1)], lvl, dst);
            }
        } else {
            pack_epilogue<PL, SPLIT>(v, t, re, im, tw, dst, active);
        }
    } else {
        store_last<PL>(v, t, out + g * N, active);
    }
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
    // points per thread: whole transform for N <= 16, 16 up to 2^14, 32 at 2^15
    constexpr int LOGP = LOGN <= 4 ? LOGN : (LOGN <= 14 ? 4 : 5);
    constexpr int T = (1 << LOGN) >> LOGP;
    constexpr int WG = T >= 256 ? T : 256;
    constexpr bool SPLIT = LOGN >= 15;
    return LaunchCfg{fft_lds_kernel<LOGN, LOGP, WG, KIND, SPLIT>, WG, WG / T};
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
