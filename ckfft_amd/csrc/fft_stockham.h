// fft_stockham.h -- register/LDS Stockham building blocks shared by the one-workgroup
// kernels (fft_lds.hip) and the multi-pass kernels (fft_pass.hip).
//
// A transform of N points is owned by T = N/P threads, each keeping P points in
// registers.  Stage s has radix R (16, except one smaller "odd" stage) and span Ns:
// butterfly j reads point j + r*N/R and writes (j/Ns)*Ns*R + j%Ns + r*Ns (Stockham
// autosort), so input and output are in natural order.  Between stages the data moves
// through LDS as separate real/imaginary planes, padded per exchange (ExchangeLayout).
#pragma once
#include "ckfft_internal.h"
#include "fft_device.h"

namespace ckfft_gpu {

// Radix plan: every stage is radix 2^RMAX_LOG except one smaller "odd" stage, placed
// first (SMALL_LAST = false) or last (SMALL_LAST = true).  The radix is 16, except for
// 2^15 points at 64 per thread: three radix-32 stages (two butterflies per thread), so
// two LDS exchanges instead of the three of 16 x 16 x 16 x 8.
template <int LOGN, int LOGP, bool SMALL_LAST = false>
struct Plan {
    static constexpr int N = 1 << LOGN;
    static constexpr int P = 1 << LOGP;          // points per thread
    static constexpr int T = N / P;              // threads per transform
    static constexpr int RMAX_LOG = (LOGN == 15 && LOGP == 6) ? 5 : (LOGP < 4 ? LOGP : 4);
    static constexpr int NSTAGES = (LOGN + RMAX_LOG - 1) / RMAX_LOG;
    static constexpr int SMALL_LOG = LOGN - (NSTAGES - 1) * RMAX_LOG;
    static constexpr int rlog(int s)
    {
        return (SMALL_LAST ? s == NSTAGES - 1 : s == 0) ? SMALL_LOG : RMAX_LOG;
    }
    static constexpr int nslog(int s)
    {
        int a = 0;
        for (int i = 0; i < s; ++i) a += rlog(i);
        return a;
    }
    static_assert(LOGP <= LOGN, "more points per thread than points");
    static_assert(SMALL_LOG >= 1 && SMALL_LOG <= RMAX_LOG, "bad radix plan");
};

// Butterflies per thread in stage S.
template <class PL, int S>
constexpr int bflies() { return PL::P >> PL::rlog(S); }

// LDS layout of the exchange that follows stage SW: logical index i is stored at word
// i + PAD * (i >> K) (a signed i shifts arithmetically).  A half-wave (32 lanes) addresses
// the 32 banks in one cycle, so the layout is chosen from the accesses of that exchange.
//  * A transform of >= 32 threads: K = 5.  The next stage reads runs of 32 consecutive
//    points from 32-aligned starts, which a pad every 32 words never splits.  Stage SW
//    writes runs of NS consecutive points NS*R apart: no pad when NS >= 32 (the runs are
//    whole rows), one word per 32 when NS*R < 32 (lane stride R), and otherwise 32/R words,
//    which moves each of the 32/NS runs of a half-wave by NS banks.
//  * Fewer threads (several transforms in a half-wave): one word per 16.
template <class PL, int SW>
struct ExchangeLayout {
    static constexpr int R = 1 << PL::rlog(SW);
    static constexpr int NS = 1 << PL::nslog(SW);
    static constexpr int K = PL::T >= 32 ? 5 : 4;
    static constexpr int PAD = PL::T >= 32 ? (NS >= 32 ? 0 : (NS * R < 32 ? 1 : 32 / R)) : 1;
    __host__ __device__ static constexpr int pad(int i) { return i + PAD * (i >> K); }
    static constexpr int SPAN = (PL::N - 1) + PAD * ((PL::N - 1) >> K) + 1;   // words per plane
};

// Words per LDS plane that every exchange of a plan fits in.
template <class PL, int S = 0>
constexpr int plan_span()
{
    if constexpr (S + 1 >= PL::NSTAGES) {
        return ExchangeLayout<PL, S>::SPAN;
    } else {
        constexpr int a = ExchangeLayout<PL, S>::SPAN, b = plan_span<PL, S + 1>();
        return a > b ? a : b;
    }
}

// Thread -> butterfly assignment of stage S.  Normally thread t owns butterflies
// t + b*T.  In a "mirror-closed" stage (MIRROR, needs an even count B >= 2) thread t owns
// j = t + b*T for b < B/2 and M - j for the other half (M = N/R butterflies; thread 0
// takes M/2 in place of M - 0), so that the points k and N - k of every pair live in the
// same thread: element (j, r) is point j + r*M and its mirror N - (j + r*M) is element
// (M - j, R-1-r).  The real pack/unpack then needs no exchange.
template <class PL, int S, bool MIRROR>
__device__ __forceinline__ int bfly(int t, int b)
{
    constexpr int B = bflies<PL, S>();
    constexpr int M = PL::N >> PL::rlog(S);
    if constexpr (!MIRROR) {
        return t + b * PL::T;
    } else {
        static_assert(B >= 2 && (B % 2) == 0, "mirror-closed stage needs an even butterfly count");
        if (b < B / 2) return t + b * PL::T;
        const int jl = t + (b - B / 2) * PL::T;
        return jl == 0 ? M / 2 : M - jl;
    }
}

// ---- stage arithmetic --------------------------------------------------------

// Multiply a[1..R-1] by w^r, w = lvl[k].  Only the power-of-two powers w, w^2, w^4, w^8
// are read from the table (exact FP64-rounded values); the others are products of at
// most four of them, which keeps the twiddle working set per butterfly small.  The reads
// (twiddle_load) are separate from the products (twiddle_apply) so that a stage can
// issue its reads early.
template <int R>
struct TwBase {
    float2 w[R >= 32 ? 5 : R >= 16 ? 4 : R >= 8 ? 3 : R >= 4 ? 2 : 1];   // w, w^2, w^4, w^8, w^16
};

template <int R, bool INV>
__device__ __forceinline__ TwBase<R> twiddle_load(const float2* __restrict__ lvl, int k)
{
    TwBase<R> b;
    b.w[0] = dirw<INV>(lvl[k]);
    if constexpr (R >= 4) b.w[1] = dirw<INV>(lvl[2 * k]);
    if constexpr (R >= 8) b.w[2] = dirw<INV>(lvl[4 * k]);
    if constexpr (R >= 16) b.w[3] = dirw<INV>(lvl[8 * k]);
    if constexpr (R >= 32) b.w[4] = dirw<INV>(lvl[16 * k]);
    return b;
}

// The powers w^0 .. w^(R-1) from the power-of-two ones in tb, formed as twiddle_apply
// forms them (w^0 = 1 exactly).
template <int R>
__device__ __forceinline__ void twiddle_powers(float2* p, const TwBase<R>& tb)
{
    p[0] = make_float2(1.0f, 0.0f);
    if constexpr (R >= 2) p[1] = tb.w[0];
    if constexpr (R >= 4) {
        p[2] = tb.w[1];
        p[3] = cmul(p[1], p[2]);
    }
    if constexpr (R >= 8) {
        p[4] = tb.w[2];
        p[5] = cmul(p[1], p[4]);
        p[6] = cmul(p[2], p[4]);
        p[7] = cmul(p[3], p[4]);
    }
    if constexpr (R >= 16) {
        p[8] = tb.w[3];
        p[9] = cmul(p[1], p[8]);
        p[10] = cmul(p[2], p[8]);
        p[11] = cmul(p[3], p[8]);
        p[12] = cmul(p[4], p[8]);
        p[13] = cmul(p[1], p[12]);
        p[14] = cmul(p[2], p[12]);
        p[15] = cmul(p[3], p[12]);
    }
    if constexpr (R == 32) {
        p[16] = tb.w[4];
#pragma unroll
        for (int m = 1; m < 16; ++m) p[16 + m] = cmul(p[16], p[m]);
    }
}

template <int R>
__device__ __forceinline__ void twiddle_apply(float2* a, const TwBase<R>& tb)
{
    if constexpr (R == 32) {
        // w^0..w^15 as twiddle_apply<16> forms them, w^(16+m) = w^16 w^m
        float2 p[16];
        TwBase<16> lo;
#pragma unroll
        for (int i = 0; i < 4; ++i) lo.w[i] = tb.w[i];
        twiddle_powers<16>(p, lo);
        const float2 w16 = tb.w[4];
#pragma unroll
        for (int m = 1; m < 16; ++m) a[m] = cmul(a[m], p[m]);
        a[16] = cmul(a[16], w16);
#pragma unroll
        for (int m = 1; m < 16; ++m) a[16 + m] = cmul(a[16 + m], cmul(w16, p[m]));
    } else if constexpr (R >= 2) {
        const float2 w1 = tb.w[0];
        if constexpr (R == 2) {
            a[1] = cmul(a[1], w1);
        } else {
            const float2 w2 = tb.w[1];
            const float2 w3 = cmul(w1, w2);
            a[1] = cmul(a[1], w1);
            a[2] = cmul(a[2], w2);
            a[3] = cmul(a[3], w3);
            if constexpr (R >= 8) {
                const float2 w4 = tb.w[2];
                a[4] = cmul(a[4], w4);
                a[5] = cmul(a[5], cmul(w1, w4));
                a[6] = cmul(a[6], cmul(w2, w4));
                const float2 w7 = cmul(w3, w4);
                a[7] = cmul(a[7], w7);
                if constexpr (R == 16) {
                    const float2 w8 = tb.w[3];
                    a[8] = cmul(a[8], w8);
                    a[9] = cmul(a[9], cmul(w1, w8));
                    a[10] = cmul(a[10], cmul(w2, w8));
                    a[11] = cmul(a[11], cmul(w3, w8));
                    const float2 w12 = cmul(w4, w8);
                    a[12] = cmul(a[12], w12);
                    a[13] = cmul(a[13], cmul(w1, w12));
                    a[14] = cmul(a[14], cmul(w2, w12));
                    a[15] = cmul(a[15], cmul(w3, w12));
                }
            }
        }
    }
}

template <int R, bool INV>
__device__ __forceinline__ void twiddle_mul(float2* a, const float2* __restrict__ lvl, int k)
{
    twiddle_apply<R>(a, twiddle_load<R, INV>(lvl, k));
}

// The twiddle reads of every butterfly of stage S (stages after the first), issued ahead
// of the stage so that waiting for them never waits for later memory operations.
template <class PL, int S, bool INV, bool MIRROR = false>
struct StageTwiddles {
    static constexpr int R = 1 << PL::rlog(S);
    static constexpr int NS = 1 << PL::nslog(S);
    TwBase<R> b[PL::P / R];
    __device__ __forceinline__ StageTwiddles(int t, const float2* __restrict__ tw)
    {
        static_assert(NS > 1, "the first stage has no twiddles");
#pragma unroll
        for (int i = 0; i < PL::P / R; ++i)
            b[i] = twiddle_load<R, INV>(tw + NS * R, bfly<PL, S, MIRROR>(t, i) & (NS - 1));
    }
    // Have every read land here.  A stage that stores under a branch needs this before its
    // first store: the compiler's wait for a later read, placed after stores it cannot
    // prove were issued, would otherwise count those stores and wait for them as well.
    __device__ __forceinline__ void settle()
    {
#pragma unroll
        for (int i = 0; i < PL::P / R; ++i)
#pragma unroll
            for (int j = 0; j < (int)(sizeof(b[i].w) / sizeof(float2)); ++j)
                asm volatile("" : "+v"(b[i].w[j].x), "+v"(b[i].w[j].y));
    }
};

// The same for a plain array of values read from memory.
template <int K>
__device__ __forceinline__ void settle_reads(float2 (&w)[K])
{
#pragma unroll
    for (int i = 0; i < K; ++i) asm volatile("" : "+v"(w[i].x), "+v"(w[i].y));
}

// Butterfly b of stage S with twiddles read beforehand.
template <class PL, int S, bool INV, bool MIRROR>
__device__ __forceinline__ void butterfly_compute(float2* v, const StageTwiddles<PL, S, INV, MIRROR>& tws, int b)
{
    constexpr int R = 1 << PL::rlog(S);
    twiddle_apply<R>(v + b * R, tws.b[b]);
    dft<R, INV>(v + b * R);
}

// Butterfly b of stage S: its twiddles, then its DFT.
template <class PL, int S, bool INV, bool MIRROR = false>
__device__ __forceinline__ void butterfly_compute(float2* v, int t, const float2* __restrict__ tw, int b)
{
    constexpr int R = 1 << PL::rlog(S);
    constexpr int NS = 1 << PL::nslog(S);
    if constexpr (NS > 1) {
        const int k = bfly<PL, S, MIRROR>(t, b) & (NS - 1);
        twiddle_mul<R, INV>(v + b * R, tw + NS * R, k);     // level NS*R of the pyramid
    }
    dft<R, INV>(v + b * R);
}

template <class PL, int S, bool INV, bool MIRROR = false>
__device__ __forceinline__ void stage_compute(float2* v, int t, const float2* __restrict__ tw)
{
#pragma unroll
    for (int b = 0; b < bflies<PL, S>(); ++b) butterfly_compute<PL, S, INV, MIRROR>(v, t, tw, b);
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

// ---- LDS addresses: one per-thread base plus compile-time offsets ----------------
//
// Every exchange index here -- in_index / out_index of butterfly bfly(t, b), element r --
// splits as U(t, g) + C(b, r): U depends on the thread and on the butterfly's group g
// (0: j = t + bT; in a mirror-closed stage 1: the first mirrored butterfly, where thread
// 0 takes M/2, and 2: the other mirrored ones), C is a compile-time constant.  The low
// K bits of U and C (K of the exchange's layout) never carry into each other, so the
// padded index is pad(U) + pad(C): each thread computes at most three bases once and
// every access is an immediate offset.  ok() proves the split at compile time; the stages
// where it cannot (transforms of few threads) use the plain formula.

template <class PL, int S, bool MIRROR, bool OUT>
struct IndexSplit {
    // writes of stage S use the layout after S, reads of stage S the layout after S-1
    using LY = ExchangeLayout<PL, OUT ? S : (S > 0 ? S - 1 : 0)>;
    static constexpr int R = 1 << PL::rlog(S);
    static constexpr int NS = 1 << PL::nslog(S);
    static constexpr int B = PL::P / R;
    static constexpr int M = PL::N >> PL::rlog(S);
    static constexpr int T = PL::T;
    static_assert(!(MIRROR && OUT) || NS == 1, "mirror-closed stages are edge stages");

    static constexpr int group(int b) { return (!MIRROR || b < B / 2) ? 0 : (b == B / 2 ? 1 : 2); }
    static constexpr int butterfly(int t, int b)
    {
        if (group(b) == 0) return t + b * T;
        const int jl = t + (b - B / 2) * T;
        return jl == 0 ? M / 2 : M - jl;
    }
    static constexpr int reference(int t, int b, int r)
    {
        const int j = butterfly(t, b);
        return OUT ? (j / NS) * NS * R + (j & (NS - 1)) + r * NS : j + r * M;
    }
    static constexpr int U(int t, int g)
    {
        const int scale = OUT ? R : 1;                       // out index of span 1 is j*R + r
        if (g == 1) return (t == 0 ? -(M / 2) : -t) * scale;
        if (g == 2) return -t * scale;
        if (!OUT) return t;
        return NS >= T ? t : (t / NS) * NS * R + (t & (NS - 1));
    }
    static constexpr int C(int b, int r)
    {
        const int g = group(b);
        if (!OUT) return g == 0 ? b * T + r * M : M - (b - B / 2) * T + r * M;
        if (g != 0) return (M - (b - B / 2) * T) * R + r;
        const int bt = b * T;
        return NS >= T ? (bt / NS) * NS * R + (bt % NS) + r * NS : (bt / NS) * NS * R + r * NS;
    }
    // U + C equals the index by construction (T and NS are powers of two, one divides the
    // other; a sample of threads re-checks it).  pad(U + C) = pad(U) + pad(C) exactly when
    // (U & L) + (C & L) <= L, L = 2^K - 1 (floor division by 2^K on both sides), or when
    // the layout has no pad; the bound below, the largest low part of U over all threads
    // plus that of C over the group's elements, makes the split exact for every t, b, r.
    static constexpr bool ok()
    {
        constexpr int L = (1 << LY::K) - 1;
        const int probe[6] = {0, 1, 15, 16, T / 2, T - 1};
        for (int k = 0; k < 6; ++k) {
            const int t = probe[k] < T ? probe[k] : T - 1;
            for (int b = 0; b < B; ++b)
                for (int r = 0; r < R; ++r)
                    if (U(t, group(b)) + C(b, r) != reference(t, b, r)) return false;
        }
        for (int g = 0; g < 3; ++g) {
            int mu = -1, mc = -1;
            for (int b = 0; b < B; ++b)
                if (group(b) == g)
                    for (int r = 0; r < R; ++r) mc = (C(b, r) & L) > mc ? (C(b, r) & L) : mc;
            if (mc < 0) continue;                                   // no butterfly in this group
            for (int t = 0; t < T; ++t) mu = (U(t, g) & L) > mu ? (U(t, g) & L) : mu;
            if (LY::PAD != 0 && mu + mc > L) return false;
        }
        return true;
    }
};

// The check, evaluated once per stage shape.
template <class PL, int S, bool MIRROR, bool OUT>
inline constexpr bool kIndexSplits = IndexSplit<PL, S, MIRROR, OUT>::ok();

// LDS instructions take an unsigned immediate offset below 64 KiB, so a plane larger than
// that is addressed through one base pointer per 16384-word window: the window is chosen
// at compile time from the offset, and every access is base + immediate.
constexpr int kLdsWindow = 16384;

template <int NW>
struct LdsWindows {
    float* p[NW];
    // plane + u is the thread's base; each window's offset is made opaque so that the
    // compiler keeps one register per window instead of re-adding large constants
    __device__ __forceinline__ LdsWindows(float* plane, int u)
    {
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            int off = u + w * kLdsWindow;
            if (w > 0) asm volatile("" : "+v"(off));
            p[w] = plane + off;
        }
    }
    __device__ __forceinline__ float& at(int c) const { return p[c / kLdsWindow][c % kLdsWindow]; }
};

// COMP: 1 = real plane, 2 = imaginary plane, 3 = both.
template <class PL, int S, int COMP, bool MIRROR = false>
__device__ __forceinline__ void lds_write(const float2* v, int t, float* re, float* im)
{
    using SP = IndexSplit<PL, S, MIRROR, true>;
    constexpr int R = SP::R, B = SP::B;
    if constexpr (kIndexSplits<PL, S, MIRROR, true>) {
        constexpr int NW = plan_span<PL>() / kLdsWindow + 1;
        using LY = typename SP::LY;
        const int ub[3] = {LY::pad(SP::U(t, 0)), MIRROR ? LY::pad(SP::U(t, 1)) : 0, MIRROR ? LY::pad(SP::U(t, 2)) : 0};
#pragma unroll
        for (int g = 0; g < (MIRROR ? 3 : 1); ++g) {
            const LdsWindows<NW> wr(re, ub[g]), wi(im, ub[g]);
#pragma unroll
            for (int b = 0; b < B; ++b) {
                if (SP::group(b) != g) continue;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int c = LY::pad(SP::C(b, r));
                    if constexpr (COMP & 1) wr.at(c) = v[b * R + r].x;
                    if constexpr (COMP & 2) wi.at(c) = v[b * R + r].y;
                }
            }
        }
    } else {   // small transforms (fewer than 16 threads) where the split carries
#pragma unroll
        for (int b = 0; b < B; ++b)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = ExchangeLayout<PL, S>::pad(out_index<PL, S>(bfly<PL, S, MIRROR>(t, b), r));
                if constexpr (COMP & 1) re[i] = v[b * R + r].x;
                if constexpr (COMP & 2) im[i] = v[b * R + r].y;
            }
    }
}

template <class PL, int S, int COMP, bool MIRROR = false>
__device__ __forceinline__ void lds_read(float2* v, int t, const float* re, const float* im)
{
    using SP = IndexSplit<PL, S, MIRROR, false>;
    constexpr int R = SP::R, B = SP::B;
    if constexpr (kIndexSplits<PL, S, MIRROR, false>) {
        constexpr int NW = plan_span<PL>() / kLdsWindow + 1;
        using LY = typename SP::LY;
        const int ub[3] = {LY::pad(SP::U(t, 0)), MIRROR ? LY::pad(SP::U(t, 1)) : 0, MIRROR ? LY::pad(SP::U(t, 2)) : 0};
#pragma unroll
        for (int g = 0; g < (MIRROR ? 3 : 1); ++g) {
            const LdsWindows<NW> wr(const_cast<float*>(re), ub[g]), wi(const_cast<float*>(im), ub[g]);
#pragma unroll
            for (int b = 0; b < B; ++b) {
                if (SP::group(b) != g) continue;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int c = LY::pad(SP::C(b, r));
                    if constexpr (COMP & 1) v[b * R + r].x = wr.at(c);
                    if constexpr (COMP & 2) v[b * R + r].y = wi.at(c);
                }
            }
        }
    } else {   // small transforms (fewer than 16 threads) where the split carries
#pragma unroll
        for (int b = 0; b < B; ++b)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = ExchangeLayout<PL, S - 1>::pad(in_index<PL, S>(bfly<PL, S, MIRROR>(t, b), r));
                if constexpr (COMP & 1) v[b * R + r].x = re[i];
                if constexpr (COMP & 2) v[b * R + r].y = im[i];
            }
    }
}

// The order of one transform's LDS writes and reads across its threads.  A transform whose
// T threads lie in one wavefront (T <= 64, so threads tr*T .. tr*T + T - 1 share a wave)
// exchanges only within that wave: the wave's LDS operations execute in order, so a
// compiler fence and a wave barrier suffice, and the other waves of the workgroup go on.
// Used for 32 <= T <= 64: complex 1024 x 8192 25.3-26.2 us against 27.3-27.8 us with the
// workgroup barrier; complex 64 (T = 4) measured 0.8 us slower with it, 256 (T = 16) the same.
template <class PL>
__device__ __forceinline__ void exchange_sync()
{
    if constexpr (PL::T >= 32 && PL::T <= 64) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    } else {
        __syncthreads();
    }
}

// Move data from the output layout of stage SW to the input layout of stage SR.
template <class PL, int SW, int SR, bool SPLIT, bool MW, bool MR>
__device__ __forceinline__ void exchange(float2* v, int t, float* re, float* im)
{
    if constexpr (!SPLIT) {
        lds_write<PL, SW, 3, MW>(v, t, re, im);
        exchange_sync<PL>();
        lds_read<PL, SR, 3, MR>(v, t, re, im);
        exchange_sync<PL>();
    } else {
        lds_write<PL, SW, 1, MW>(v, t, re, im);
        __syncthreads();
        lds_read<PL, SR, 1, MR>(v, t, re, im);
        __syncthreads();
        // The imaginary plane recomputes its LDS addresses from an opaque copy of t.
        // Otherwise the compiler keeps every address of the real-plane pass live across
        // the barriers; mirror-closed stages have one address register per point, and
        // at 64 points per thread that spills.
        int t2 = t;
        asm volatile("" : "+v"(t2));
        lds_write<PL, SW, 2, MW>(v, t2, re, im);
        __syncthreads();
        lds_read<PL, SR, 2, MR>(v, t2, re, im);
        __syncthreads();
    }
}

}  // namespace ckfft_gpu
