// fft_mfma.h -- radix-16 DFT stages as matrix products on the gfx950 matrix cores, with
// split-BF16x3 operands and FP32 accumulation (the opt-in "split-BF16x3" precision mode).
//
// The reference computes every stage with radix-2/4 butterflies in FP32
// (src/ckfft/fft_default.cpp:168-266); this mode computes the same transform with the
// radix-16 stages as dense products instead.  A radix-16 DFT over 32 butterflies is
//     Y (32 x 32) = G (32 x 32) * X (32 x 32)
// with rows/columns of G and X ordered [re 0..15, im 0..15]:
//     G[k][r] = Re w,  G[k][16+r] = -Im w,  G[16+k][r] = Im w,  G[16+k][16+r] = Re w,
//     w = exp(-+2 pi i k r / 16)
// and runs as two v_mfma_f32_32x32x16_bf16 (K = 16 each: the real, then the imaginary
// inputs).  Every FP32 operand is split exactly into three BF16 parts
//     x = hi + mid + lo,  hi = bf16(x), mid = bf16(x - hi), lo = bf16(x - hi - mid)
// (round to nearest even, so |mid| <= 2^-9 |x| and |lo| <= 2^-18 |x|; each residual is
// exact in FP32 and the split is exact) and the six largest partial products are
// accumulated in FP32, smallest first:
//     lo*hi + mid*mid + hi*lo + mid*hi + hi*mid + hi*hi
// The dropped terms (mid*lo, lo*mid, lo*lo) are below 2^-27 relative, so the result has
// FP32-level error.  Twiddles are applied in FP32 (one fused complex product per point)
// before the split.
//
// Operand layouts of v_mfma_f32_32x32x16_bf16 (64 lanes; checked on gfx950):
//   A[m][k] (32 x 16): lane m + 32*(k/8), element k%8
//   B[k][n] (16 x 32): lane n + 32*(k/8), element k%8
//   D[row][col] (32 x 32 FP32): lane col + 32*((row/4)%2), register (row/8)*4 + row%4
// So lane n + 32h holds inputs r = 8h..8h+7 of butterfly n (real parts for the first
// product, imaginary for the second) and receives outputs
//     k_u = (u/4)*8 + 4h + u%4   (u = 0..7): real part in register u, imaginary in u+8.
#pragma once
#include "fft_device.h"

namespace ckfft_gpu {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// Exact three-way split of two FP32 values; each result word packs the two BF16 parts
// (a in the low half, b in the high half), the layout of two operand elements.
__device__ __forceinline__ void split3_pair(float a, float b, unsigned& hi, unsigned& mid, unsigned& lo)
{
    const unsigned h = __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){a, b}, bf16x2));
    const float ra = a - __uint_as_float(h << 16), rb = b - __uint_as_float(h & 0xffff0000u);
    const unsigned m = __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){ra, rb}, bf16x2));
    const float sa = ra - __uint_as_float(m << 16), sb = rb - __uint_as_float(m & 0xffff0000u);
    hi = h;
    mid = m;
    lo = __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){sa, sb}, bf16x2));
}

// ---- the DFT matrix operand, split at compile time -------------------------------------

// BF16 of x (round to nearest even, finite x) as the high half of an FP32 bit pattern.
constexpr unsigned bf16_bits(float x)
{
    const unsigned u = __builtin_bit_cast(unsigned, x);
    return (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
}

// Lane slices of G for both directions: w[inv][lane][c][part] holds 8 BF16 elements
// (element i in half i%2 of word i/2) of the A operand for the real (c = 0) or imaginary
// (c = 1) inputs, part 0 hi, 1 mid, 2 lo.
struct DftTable {
    alignas(16) unsigned w[2][64][2][3][4];
    constexpr DftTable() : w{}
    {
        for (int inv = 0; inv < 2; ++inv)
            for (int lane = 0; lane < 64; ++lane) {
                const int row = lane & 31, h = lane >> 5, k = row & 15;
                for (int i = 0; i < 8; ++i) {
                    const int q = (4 * k * (8 * h + i)) & 63;           // exp(-+2 pi i k r / 16)
                    const float wr = kW64Re[q], wi = inv ? -kW64Im[q] : kW64Im[q];
                    const float g[2] = {row < 16 ? wr : wi, row < 16 ? -wi : wr};
                    for (int c = 0; c < 2; ++c) {
                        const unsigned p0 = bf16_bits(g[c]);
                        const float r1 = g[c] - __builtin_bit_cast(float, p0);
                        const unsigned p1 = bf16_bits(r1);
                        const unsigned p2 = bf16_bits(r1 - __builtin_bit_cast(float, p1));
                        const unsigned parts[3] = {p0, p1, p2};
                        for (int part = 0; part < 3; ++part)
                            w[inv][lane][c][part][i / 2] |= (parts[part] >> 16) << (16 * (i & 1));
                    }
                }
            }
    }
};
__device__ constexpr DftTable kDftTable{};

// This lane's slice of G: a[c][p] as above.
struct DftOperand {
    bf16x8 a[2][3];
};

template <bool INV>
__device__ __forceinline__ DftOperand load_dft_operand(int lane)
{
    const u32x4* t = reinterpret_cast<const u32x4*>(&kDftTable.w[INV ? 1 : 0][lane][0][0][0]);
    DftOperand A;
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int p = 0; p < 3; ++p) A.a[c][p] = __builtin_bit_cast(bf16x8, t[c * 3 + p]);
    return A;
}

__device__ __forceinline__ f32x16 mfma_bf16(bf16x8 a, bf16x8 b, f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// Radix-16 DFT of the wave's 32 butterflies; x = this lane's 8 inputs (already
// twiddled).  Returns the accumulator in the D layout above.
__device__ __forceinline__ f32x16 dft16_mfma(const float2 (&x)[8], const DftOperand& A)
{
    unsigned re[3][4], im[3][4];
#pragma unroll
    for (int e = 0; e < 8; e += 2) {
        split3_pair(x[e].x, x[e + 1].x, re[0][e / 2], re[1][e / 2], re[2][e / 2]);
        split3_pair(x[e].y, x[e + 1].y, im[0][e / 2], im[1][e / 2], im[2][e / 2]);
    }
    bf16x8 rb[3], ib[3];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        rb[p] = __builtin_bit_cast(bf16x8, (u32x4){re[p][0], re[p][1], re[p][2], re[p][3]});
        ib[p] = __builtin_bit_cast(bf16x8, (u32x4){im[p][0], im[p][1], im[p][2], im[p][3]});
    }
    f32x16 acc = {};
    // the smallest products first, the leading one last
    acc = mfma_bf16(A.a[0][2], rb[0], acc);
    acc = mfma_bf16(A.a[1][2], ib[0], acc);
    acc = mfma_bf16(A.a[0][1], rb[1], acc);
    acc = mfma_bf16(A.a[1][1], ib[1], acc);
    acc = mfma_bf16(A.a[0][0], rb[2], acc);
    acc = mfma_bf16(A.a[1][0], ib[2], acc);
    acc = mfma_bf16(A.a[0][1], rb[0], acc);
    acc = mfma_bf16(A.a[1][1], ib[0], acc);
    acc = mfma_bf16(A.a[0][0], rb[1], acc);
    acc = mfma_bf16(A.a[1][0], ib[1], acc);
    acc = mfma_bf16(A.a[0][0], rb[0], acc);
    acc = mfma_bf16(A.a[1][0], ib[0], acc);
    return acc;
}

// Output j of a lane's 8 (u = 0..7): point offset (u/4)*8 + u%4 from the lane's first
// output 4h (in units of the stage span).
__device__ constexpr int kOutRun[8] = {0, 1, 2, 3, 8, 9, 10, 11};

// ---- where a stage reads and writes ----------------------------------------------------
//
// LDS tile of B columns x L points: column b, point idx at b * COLS + idx + idx/16
// (float2 units) -- one pad word per 16 points, and an odd column stride so that lanes
// walking across columns at one point index fall on different banks.
template <int LOGL, int LOGB>
struct MfmaTile {
    static constexpr int L = 1 << LOGL;
    static constexpr int B = 1 << LOGB;
    static constexpr int COLS = L + L / 16 + 1;
    static constexpr int WORDS = B * COLS;                 // float2 entries
};

template <class TL>
__device__ __forceinline__ int tile_at(int b, int idx)
{
    return b * TL::COLS + idx + (idx >> 4);
}

// Every IO type offers
//   load<CNT, STRIDE>(b, idx0, x):   x[e] = z[b][idx0 + e STRIDE], e < CNT
//   store_run<CNT>(b, idx0, v):      z[b][idx0 + e] = v[e]      (idx0 % CNT == 0, CNT | 16)
//   store_out<SPAN>(b, idx0, v):     z[b][idx0 + kOutRun[u] SPAN] = v[u], u < 8
template <class TL>
struct LdsIO {
    static constexpr bool kLds = true;
    float2* lds;

    template <int CNT, int STRIDE>
    __device__ __forceinline__ void load(int b, int idx0, float2* x) const
    {
        if constexpr (STRIDE % 16 == 0) {
            const int p = tile_at<TL>(b, idx0);
#pragma unroll
            for (int e = 0; e < CNT; ++e) x[e] = lds[p + e * (STRIDE + STRIDE / 16)];
        } else {
#pragma unroll
            for (int e = 0; e < CNT; ++e) x[e] = lds[tile_at<TL>(b, idx0 + e * STRIDE)];
        }
    }
    template <int CNT>
    __device__ __forceinline__ void store_run(int b, int idx0, const float2* v) const
    {
        static_assert(16 % CNT == 0, "a run must not cross a pad word");
        const int p = tile_at<TL>(b, idx0);
#pragma unroll
        for (int e = 0; e < CNT; ++e) lds[p + e] = v[e];
    }
    template <int SPAN>
    __device__ __forceinline__ void store_out(int b, int idx0, const float2* v) const
    {
        if constexpr (SPAN % 16 == 0) {
            const int p = tile_at<TL>(b, idx0);
#pragma unroll
            for (int u = 0; u < 8; ++u) lds[p + kOutRun[u] * (SPAN + SPAN / 16)] = v[u];
        } else if constexpr (SPAN == 1) {
            // idx0 = 16 i + 4h: the run stays inside one 16-point block
            const int p = tile_at<TL>(b, idx0);
#pragma unroll
            for (int u = 0; u < 8; ++u) lds[p + kOutRun[u]] = v[u];
        } else {
#pragma unroll
            for (int u = 0; u < 8; ++u) lds[tile_at<TL>(b, idx0 + kOutRun[u] * SPAN)] = v[u];
        }
    }
};

// Whole signals in global memory: column b is signal g0 + b of 2^LOGN points.  A short
// last group reads the last signal again and stores nothing for columns past `count`.
// Reads and writes carry the non-temporal hint (ld_in / st_out): each is done once.
template <int LOGN>
struct SignalIO {
    static constexpr bool kLds = false;
    static constexpr int N = 1 << LOGN;
    const float2* __restrict__ in;
    float2* __restrict__ out;
    long long g0, count;

    template <int CNT, int STRIDE>
    __device__ __forceinline__ void load(int b, int idx0, float2* x) const
    {
        const long long g = g0 + b < count ? g0 + b : count - 1;
        const float2* p = in + g * N + idx0;
#pragma unroll
        for (int e = 0; e < CNT; ++e) x[e] = ld_in<true>(&p[e * STRIDE]);
    }
    template <int CNT>
    __device__ __forceinline__ void store_run(int b, int idx0, const float2* v) const
    {
        if (g0 + b >= count) return;
        float2* p = out + (g0 + b) * N + idx0;
#pragma unroll
        for (int e = 0; e < CNT; ++e) st_out<true>(&p[e], v[e]);
    }
    template <int SPAN>
    __device__ __forceinline__ void store_out(int b, int idx0, const float2* v) const
    {
        if (g0 + b >= count) return;
        float2* p = out + (g0 + b) * N + idx0;
#pragma unroll
        for (int u = 0; u < 8; ++u) st_out<true>(&p[kOutRun[u] * SPAN], v[u]);
    }
};

// ---- stages -----------------------------------------------------------------------------
//
// A stage reads all its inputs (src), computes, and writes (dst).  When both are the LDS
// tile it works in place: every read completes (barrier) before the first write.  A
// stage writing LDS ends with a barrier so the next stage sees the whole tile.

// First Stockham stage of radix 2^LOGR0 (2, 4 or 8) on the VALU: span 1, no twiddles.
// Sub-butterfly i of column b reads z[i + q L/R], writes z[i R + k].
template <class TL, int WG, int LOGR0, bool INV, class SRC, class DST>
__device__ __forceinline__ void small_stage(const SRC& src, const DST& dst, int tid)
{
    constexpr int R = 1 << LOGR0, NBS = TL::L / R;
    constexpr int TOTAL = NBS * TL::B;
    static_assert(TOTAL % WG == 0, "butterflies must divide among the threads");
    constexpr int PER = TOTAL / WG;
    float2 x[PER][R];
#pragma unroll
    for (int c = 0; c < PER; ++c) {
        const int f = c * WG + tid;
        src.template load<R, NBS>(f / NBS, f % NBS, x[c]);
        dft<R, INV>(x[c]);
    }
    if constexpr (SRC::kLds && DST::kLds) __syncthreads();
#pragma unroll
    for (int c = 0; c < PER; ++c) {
        const int f = c * WG + tid;
        dst.template store_run<R>(f / NBS, (f % NBS) * R, x[c]);
    }
    if constexpr (DST::kLds) __syncthreads();
}

// One radix-16 Stockham stage of span 2^LOGNS on the matrix cores: sub-butterfly i of
// column b reads z[i + q L/16] (times W_{16 NS}^{q (i mod NS)}), writes
// z[(i / NS) 16 NS + i mod NS + k NS].  Each wave takes chunks of 32 butterflies.
template <class TL, int WG, int LOGNS, bool INV, class SRC, class DST>
__device__ __forceinline__ void mfma_stage(const SRC& src, const DST& dst, const float2* __restrict__ pyr,
                                           const DftOperand& A, int tid)
{
    constexpr int NB = TL::L / 16, NS = 1 << LOGNS;
    constexpr int CHUNKS = NB * TL::B / 32;
    constexpr int WAVES = WG / 64;
    static_assert(CHUNKS % WAVES == 0, "chunks must divide among the waves");
    constexpr int K = CHUNKS / WAVES;
    const int wave = tid >> 6, lane = tid & 63, n = lane & 31, h = lane >> 5;
    f32x16 acc[K];
#pragma unroll
    for (int c = 0; c < K; ++c) {
        const int f = (c * WAVES + wave) * 32 + n;
        const int i = f % NB;
        float2 x[8];
        src.template load<8, NB>(f / NB, i + 8 * h * NB, x);
        if constexpr (NS > 1) {
            const int m = i & (NS - 1);
            const float2* tw = pyr + 16 * NS + 8 * h * m;
#pragma unroll
            for (int e = 0; e < 8; ++e) x[e] = cmul(x[e], dirw<INV>(tw[e * m]));
        }
        acc[c] = dft16_mfma(x, A);
    }
    if constexpr (SRC::kLds && DST::kLds) __syncthreads();
#pragma unroll
    for (int c = 0; c < K; ++c) {
        const int f = (c * WAVES + wave) * 32 + n;
        const int i = f % NB;
        float2 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = make_float2(acc[c][u], acc[c][u + 8]);
        dst.template store_out<NS>(f / NB, (i >> LOGNS) * 16 * NS + (i & (NS - 1)) + 4 * h * NS, v);
    }
    if constexpr (DST::kLds) __syncthreads();
}

// Stage S of the column FFT: the first reads `first`, the last writes `last`, the others
// read and write the LDS tile.
template <class TL, int WG, bool INV, int S, class FIRST, class LAST>
__device__ __forceinline__ void column_stages(const FIRST& first, const LdsIO<TL>& mid, const LAST& last,
                                              const float2* __restrict__ pyr, const DftOperand& A, int tid)
{
    constexpr int LOGL = __builtin_ctz(TL::L);
    constexpr int LOGR0 = LOGL % 4;
    constexpr int NST = (LOGR0 ? 1 : 0) + LOGL / 4;
    if constexpr (S < NST) {
        auto body = [&](const auto& src, const auto& dst) {
            if constexpr (S == 0 && LOGR0 != 0)
                small_stage<TL, WG, LOGR0, INV>(src, dst, tid);
            else
                mfma_stage<TL, WG, LOGR0 + 4 * (S - (LOGR0 ? 1 : 0)), INV>(src, dst, pyr, A, tid);
        };
        if constexpr (S == 0 && S == NST - 1)
            body(first, last);
        else if constexpr (S == 0)
            body(first, mid);
        else if constexpr (S == NST - 1)
            body(mid, last);
        else
            body(mid, mid);
        column_stages<TL, WG, INV, S + 1>(first, mid, last, pyr, A, tid);
    }
}

// L-point DFT (natural order in and out) of each of the B columns: one VALU stage of
// radix 2^(LOGL mod 4) if needed, then LOGL/4 radix-16 matrix-core stages.  When `first`
// is the LDS tile it must be complete (barrier) on entry; when `last` is, it is complete
// on exit.
template <class TL, int WG, bool INV, class FIRST, class LAST>
__device__ __forceinline__ void mfma_columns_fft(const FIRST& first, const LdsIO<TL>& mid, const LAST& last,
                                                 const float2* __restrict__ pyr, int tid)
{
    const DftOperand A = load_dft_operand<INV>(tid & 63);
    column_stages<TL, WG, INV, 0>(first, mid, last, pyr, A, tid);
}

}  // namespace ckfft_gpu
