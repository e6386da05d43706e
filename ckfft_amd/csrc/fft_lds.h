// fft_lds.h -- the one-workgroup transform body (register/LDS Stockham, fused real
// pack/unpack), shared by the batch kernels (fft_lds.hip) and the persistent kernel
// (fft_batch.hip).
//
// Replaces the reference's recursive radix-4 decimation-in-time transform
// (src/ckfft/fft_default.cpp:12-266) and its real pack/unpack passes
// (src/ckfft/fft_real_default.cpp:13-114) for every size that fits one workgroup.
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
//    (src/ckfft/fft_real_default.cpp:30-62) written for one (k, N-k) pair per step.
//  * Real inverse: the unpack pre-pass t[k] = (Y[k] + conj Y[N-k]) +
//    i conj(W^k) (Y[k] - conj Y[N-k]) (src/ckfft/fft_real_default.cpp:79-111) is
//    fused into the first stage's global loads, then the N-point inverse transform
//    writes the 2N reals as N complex values.
#pragma once
#include "fft_stockham.h"

namespace ckfft_gpu {

// Which stage (if any) uses the mirror-closed assignment: the first stage of a real
// inverse (unpack fused into its loads) and the last stage of a real forward (pack fused
// into its stores), when that edge stage's radix is at most 8.  The rule looks only at
// the radix plan, never at the thread layout: every layout of a size (any points per
// thread >= 16) then has >= 2 butterflies per thread there and takes the same path, so
// a transform rounds identically in every kernel that runs it.
// Radix-32 edge stages occur only in the 2^15 plan, and every kernel runs 2^15 at 64
// points per thread (two butterflies per thread), so the rule admits them too.
constexpr bool mirror_radix(int rlog) { return rlog <= 3 || rlog == 5; }

template <class PL, int KIND>
constexpr int mirror_stage()
{
    if constexpr (PL::T == 1) {
        return -1;
    } else if constexpr (KIND == kRealInverse) {
        return (mirror_radix(PL::rlog(0)) && bflies<PL, 0>() >= 2) ? 0 : -1;
    } else if constexpr (KIND == kRealForward) {
        return (mirror_radix(PL::rlog(PL::NSTAGES - 1)) && bflies<PL, PL::NSTAGES - 1>() >= 2) ? PL::NSTAGES - 1
                                                                                                 : -1;
    } else {
        return -1;
    }
}

// exp(-2 pi i (2(R-1-r)+1) / 4R): the pack/unpack twiddle of point M/2 + (R-1-r)*M of an
// edge stage of radix R <= 32 (M = N/R), a constant (kW128 entry 2m equals kW64 entry m,
// so the radices up to 16 get the values they had from kW64).
template <int R>
__device__ __forceinline__ float2 half_step_twiddle(int r)
{
    static_assert(R >= 1 && R <= 32, "edge radix");
    return w128((2 * (R - 1 - r) + 1) * (32 / R));
}

// Stages S .. END-1, each after the exchange that brings its inputs.
template <class PL, int S, int END, bool INV, bool SPLIT, int MS>
__device__ __forceinline__ void run_stages(float2* v, int t, float* re, float* im,
                                           const float2* __restrict__ tw)
{
    if constexpr (S < END) {
        exchange<PL, S - 1, S, SPLIT, MS == S - 1, MS == S>(v, t, re, im);
        stage_compute<PL, S, INV, MS == S>(v, t, tw);
        run_stages<PL, S + 1, END, INV, SPLIT, MS>(v, t, re, im, tw);
    }
}

// ---- canonical pack/unpack twiddles --------------------------------------------------

// exp(-2 pi i k / 2N) for point k (0 <= k <= N) of the real pack/unpack of an N-point half
// transform whose edge stage (the one fused with the pack or unpack) has radix R: with
// k = j + r*M, M = N/R, it is the table value W^j times the constant exp(-2 pi i r/2R).
// The non-mirror paths form every twiddle this way; the mirror-closed paths form the
// lower half this way and take each upper point's as -conj of its partner's.  Which of
// the two a size uses depends only on its radix plan (mirror_stage), so every kernel
// computing a given size rounds the same way.
template <int N, int R>
__device__ __forceinline__ float2 edge_twiddle(const float2* __restrict__ lvl, int k)
{
    constexpr int M = N / R;
    static_assert(R >= 1 && R <= 32 && M * R == N, "edge radix");
    return cmul(lvl[k & (M - 1)], w64((k / M) * (32 / R)));
}

// ---- first-stage loads ---------------------------------------------------------

// First-stage loads.  They carry no per-element predicate: a workgroup slot past the
// last signal reads a valid signal instead (see the kernel) and never stores, so every
// load issues back to back.
template <class PL, bool NT>
__device__ __forceinline__ void load_first_complex(float2* v, int t, const float2* __restrict__ src)
{
    constexpr int R = 1 << PL::rlog(0);
    constexpr int B = PL::P / R;
#pragma unroll
    for (int b = 0; b < B; ++b)
#pragma unroll
        for (int r = 0; r < R; ++r) v[b * R + r] = ld_in<NT>(&src[in_index<PL, 0>(t + b * PL::T, r)]);
}

// Real-inverse unpack of one point: t[k] = (Y[k] + conj Y[N-k]) + i e^{+2 pi i k/2N} (Y[k] - conj Y[N-k]),
// with w = e^{-2 pi i k/2N} from the table.
__device__ __forceinline__ float2 unpack_one(float2 yk, float2 ym, float2 w)
{
    const float2 a = make_float2(yk.x + ym.x, yk.y - ym.y);     // Y[k] + conj Y[N-k]
    const float2 d = make_float2(yk.x - ym.x, yk.y + ym.y);     // Y[k] - conj Y[N-k]
    const float2 c = cmul(cconj(w), d);
    return make_float2(a.x - c.y, a.y + c.x);                   // a + i c
}

// Real inverse, general assignment: every point reads Y[k] and Y[N-k] itself (the second
// read hits the caches; used when the first stage has one butterfly per thread).
template <class PL>
__device__ __forceinline__ void load_first_unpack(float2* v, int t, const float2* __restrict__ y,
                                                  const float2* __restrict__ tw)
{
    constexpr int N = PL::N;
    constexpr int R = 1 << PL::rlog(0);
    constexpr int B = PL::P / R;
    const float2* lvl = tw + 2 * N;             // level 2N: exp(-2 pi i k / 2N)
#pragma unroll
    for (int b = 0; b < B; ++b) {
        // point j + r*N/R: W^(j + r N/R) = W^j * exp(-2 pi i r / 2R), the second a constant
        const int j = t + b * PL::T;
        const float2 wj = lvl[j];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int k = in_index<PL, 0>(j, r);
            v[b * R + r] = unpack_one(y[k], y[N - k], cmul(wj, w64(r * (32 / R))));   // = edge_twiddle<N, R>(k)
        }
    }
}

// Both unpacked values of a mirror pair (k, N-k): t[k] = unpack_one(Y[k], Y[N-k], w) and
// t[N-k] = unpack_one(Y[N-k], Y[k], -conj w), sharing their sums, differences and product.
// Bit-identical to the two unpack_one calls: the partner's sum and difference are the
// negations (exact) of these, and cmul(conj(-conj w), (-d.x, d.y)) = (c.x, -c.y) exactly.
__device__ __forceinline__ void unpack_pair(float2& yk, float2& ym, float2 w)
{
    const float2 a = make_float2(yk.x + ym.x, yk.y - ym.y);
    const float2 d = make_float2(yk.x - ym.x, yk.y + ym.y);
    const float2 c = cmul(cconj(w), d);
    yk = make_float2(a.x - c.y, a.y + c.x);
    ym = make_float2(a.x + c.y, c.x - a.y);
}

// Orders LDS accesses of different lanes of one wave: every access before it is complete
// and visible to the wave's lanes after it (no barrier: the other waves do not take part).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// The mirror-closed unpack for an edge radix above 8 (the radix-32 first stage of 2^15,
// one butterfly pair per thread).  Every thread but 0 holds mirror pairs in its two
// butterflies and unpacks them in place with unpack_pair.  Thread 0 holds the
// self-mirrored butterflies 0 and M/2, whose partners are other elements of the same
// butterflies: element (0, r), point r*M, pairs with element (0, R-r) (with Y[N] for
// r = 0), and element (1, q), point M/2 + q*M, with element (1, R-1-q); in every case the
// partner of point k is point N - k.  The 2R lanes of thread 0's wave unpack one element
// each: lane e reads its point and the partner from memory itself, first of all its reads,
// puts the result in `part` (LDS outside the exchange planes), and thread 0 takes the 2R
// results after the in-place pass.  Earlier, thread 0 wrote its 2R loaded values to
// `part` for the lanes to read, so its wave waited for all of thread 0's reads and two
// more LDS round trips before its own unpack: 2.7 us of the real 2^16 inverse.
template <class PL, bool INV, bool NT>
__device__ __forceinline__ void first_stage_unpack_mirror_wide(float2* v, int t, const float2* __restrict__ y,
                                                               const float2* __restrict__ tw, float* part)
{
    constexpr int N = PL::N;
    constexpr int R = 1 << PL::rlog(0);
    constexpr int B = PL::P / R;
    constexpr int M = N / R;
    static_assert(B == 2 && 2 * R == 64, "one mirror pair of butterflies per thread, 2R lanes in a wave");
    const float2* lvl = tw + 2 * N;
    // lane e < 2R of wave 0: element e of thread 0's butterflies, its partner and twiddle
    float2 own = make_float2(0.0f, 0.0f), partner = own, lane_w = own;
    if (t < 2 * R) {
        const int q = t & (R - 1);
        const int k = t < R ? q * M : M / 2 + q * M;
        own = y[k];
        partner = y[N - k];
        lane_w = t < R ? cmul(lvl[0], w64(q * (32 / R))) : half_step_twiddle<R>(R - 1 - q);
    }
    const float2 wl = lvl[t];                           // W^t; W^(t + rM) = W^t exp(-2 pi i r/2R)
#pragma unroll
    for (int r = 0; r < R; ++r) {
        v[r] = ld_in<NT>(&y[in_index<PL, 0>(bfly<PL, 0, true>(t, 0), r)]);
        v[R + r] = ld_in<NT>(&y[in_index<PL, 0>(bfly<PL, 0, true>(t, 1), r)]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) unpack_pair(v[r], v[R + R - 1 - r], cmul(wl, w64(r * (32 / R))));
    if (t < 2 * R) {                                    // thread 0's wave
        float2* pp = reinterpret_cast<float2*>(part);
        pp[t] = unpack_one(own, partner, lane_w);
        wave_lds_sync();
        if (t == 0) {
#pragma unroll
            for (int e = 0; e < 2 * R; ++e) v[e] = pp[e];
        }
    }
    butterfly_compute<PL, 0, INV, true>(v, t, tw, 0);
    butterfly_compute<PL, 0, INV, true>(v, t, tw, 1);
}

// Real inverse with a mirror-closed first stage, fused with that stage: each Y value is
// read once, by the thread that also holds its mirror, and the pairs are unpacked in
// registers.  Element (b, r) of a lower slot b < H is point k = jl + r*M and pairs with
// element (b+H, R-1-r), point N - k.  Thread 0's slot pair holds the self-mirrored
// butterflies 0 and M/2 instead; it fetches its partners from memory (loads in one lane).
// The loads are issued pair by pair (slots b and b+H, then thread 0's partners after the
// first pair), ahead of the arithmetic; each pair is then unpacked and its two
// butterflies computed.
template <class PL, bool INV, bool NT>
__device__ __forceinline__ void first_stage_unpack_mirror(float2* v, int t, const float2* __restrict__ y,
                                                          const float2* __restrict__ tw, float* part)
{
    constexpr int N = PL::N;
    constexpr int R = 1 << PL::rlog(0);
    constexpr int B = PL::P / R;
    constexpr int H = B / 2;
    constexpr int M = N / R;
    if constexpr (R > 8) {
        first_stage_unpack_mirror_wide<PL, INV, NT>(v, t, y, tw, part);
        return;
    }
    const float2* lvl = tw + 2 * N;
    float2 q0[R], q1[R];                                // thread 0: partners of its first pair
    float2 wl[H];                                       // W^jl; W^(jl + rM) = W^jl exp(-2 pi i r/2R)
#pragma unroll
    for (int b = 0; b < H; ++b) wl[b] = lvl[t + b * PL::T];
#pragma unroll
    for (int b = 0; b < H; ++b) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            v[b * R + r] = ld_in<NT>(&y[in_index<PL, 0>(bfly<PL, 0, true>(t, b), r)]);
            v[(b + H) * R + r] = ld_in<NT>(&y[in_index<PL, 0>(bfly<PL, 0, true>(t, b + H), r)]);
        }
        if (b == 0 && t == 0) {
            // butterflies 0 and M/2: point r*M pairs with N - r*M, point M/2 + (R-1-r)*M
            // with M/2 + r*M
#pragma unroll
            for (int r = 0; r < R; ++r) {
                q0[r] = y[N - r * M];
                q1[r] = y[M / 2 + r * M];
            }
        }
    }
#pragma unroll
    for (int b = 0; b < H; ++b) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float2& lo = v[b * R + r];
            float2& hi = v[(b + H) * R + R - 1 - r];
            const float2 w0 = cmul(wl[b], w64(r * (32 / R)));
            float2 p0 = hi, p1 = lo;                     // Y[N - k0], Y[N - k1] when paired
            float2 w1 = make_float2(-w0.x, w0.y);        // W^(N-k0) = -conj(W^k0)
            if (b == 0 && t == 0) {
                // W^(M/2 + (R-1-r)M) = exp(-2 pi i (2(R-1-r)+1) / 4R)
                p0 = q0[r];
                p1 = q1[r];
                w1 = half_step_twiddle<R>(r);
            }
            const float2 y0 = lo, y1 = hi;
            lo = unpack_one(y0, p0, w0);
            hi = unpack_one(y1, p1, w1);
        }
        butterfly_compute<PL, 0, INV, true>(v, t, tw, b);
        butterfly_compute<PL, 0, INV, true>(v, t, tw, b + H);
    }
}

// ---- last-stage stores -----------------------------------------------------------

template <class PL, bool NT>
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
            st_out<NT>(&dst[out_index<PL, S>(t + b * PL::T, r)], v[b * R + r]);
}

// One real-forward output: out[k] = A - i W^k D, A = Z[k] + conj Z[N-k], D = Z[k] - conj Z[N-k].
__device__ __forceinline__ float2 pack_one(float2 zk, float2 zm, float2 w)
{
    const float2 a = make_float2(zk.x + zm.x, zk.y - zm.y);
    const float2 d = make_float2(zk.x - zm.x, zk.y + zm.y);
    const float2 c = cmul(w, d);
    return make_float2(a.x + c.y, a.y - c.x);
}

// Both outputs of the pair (k, N-k), each formed exactly as pack_one forms it;
// zk = Z[k], zm = Z[N-k mod N].  R is the edge radix of edge_twiddle.
template <int N, int R, bool NT>
__device__ __forceinline__ void pack_pair(int k, float2 zk, float2 zm, const float2* __restrict__ lvl,
                                          float2* __restrict__ dst)
{
    st_out<NT>(&dst[k], pack_one(zk, zm, edge_twiddle<N, R>(lvl, k)));
    if (k != N - k) st_out<NT>(&dst[N - k], pack_one(zm, zk, edge_twiddle<N, R>(lvl, N - k)));
}

// The signal base of a workgroup that owns one transform, in scalar registers.
__device__ __forceinline__ float2* uniform_base(float2* p)
{
    const unsigned long long a = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
    return (float2*)(((unsigned long long)hi << 32) | lo);
}

typedef unsigned StoreBits __attribute__((ext_vector_type(2)));
// Buffer resource word 3 for untyped accesses: 32-bit data format.
constexpr int kRawBufferFormat = 0x00020000;

// Real forward: the mirror-closed last stage fused with the pack.  Every output point k
// is written once, by the thread that holds Z[k], from Z[k] and its mirror Z[N-k] in the
// same thread's registers.  Thread 0's first slot pair holds the self-mirrored
// butterflies 0 and M/2, whose partners sit at other register positions; it also writes
// out[N].  Pair by pair (slots b and b+H): the two butterflies compute, then their
// outputs are packed and stored.  The stage's twiddles (tws) and the pack twiddles W^jl
// of the lower slots (wl) are read beforehand.
// The pack's store indices as per-thread parts plus non-negative compile-time parts, and
// inside the N+1 outputs: k0 = t + (b T + r M); k1 = (M - t - (H-1) T) + ((H-1-b) T +
// (R-1-r) M) for b > 0 and jh0 + (R-1-r) M for b == 0.  Checked for every thread.
template <class PL>
constexpr bool pack_split_exact()
{
    constexpr int S = PL::NSTAGES - 1, N = PL::N, R = 1 << PL::rlog(S), B = PL::P / R, H = B / 2, M = N / R;
    constexpr int T = PL::T;
    for (int t = 0; t < T; ++t) {
        const int jh0 = t == 0 ? M / 2 : M - t;
        const int um = M - t - (H - 1) * T;
        if (um < 0) return false;
        for (int b = 0; b < H; ++b)
            for (int r = 0; r < R; ++r) {
                const int jl = t + b * T, jh = b == 0 ? jh0 : M - jl;
                const int k0 = jl + r * M, k1 = jh + (R - 1 - r) * M;
                const int c0 = b * T + r * M;
                const int c1 = (H - 1 - b) * T + (R - 1 - r) * M, c1b0 = (R - 1 - r) * M;
                if (t + c0 != k0 || k0 < 0 || k0 > N || c0 < 0) return false;
                if (b == 0 && (jh0 + c1b0 != k1 || c1b0 < 0)) return false;
                if (b > 0 && (um + c1 != k1 || c1 < 0)) return false;
                if (k1 < 0 || k1 > N) return false;
            }
    }
    return true;
}

// UNIFORM (one transform per workgroup): the stores go through a buffer resource over the
// N+1 outputs, per-thread byte offsets in VGPRs and the compile-time parts as scalar
// offsets (pack_split_exact), so no store needs 64-bit address arithmetic.
// Both outputs of a mirror pair (k, N-k) of the real forward: out[k] = pack_one(Z[k],
// Z[N-k], w), out[N-k] = pack_one(Z[N-k], Z[k], -conj w), sharing their sums, differences
// and product (bit-identical to the two pack_one calls, as for unpack_pair).
__device__ __forceinline__ void pack_pair_cse(float2 zk, float2 zm, float2 w, float2& ok, float2& om)
{
    const float2 a = make_float2(zk.x + zm.x, zk.y - zm.y);
    const float2 d = make_float2(zk.x - zm.x, zk.y + zm.y);
    const float2 c = cmul(w, d);
    ok = make_float2(a.x + c.y, a.y - c.x);
    om = make_float2(a.x - c.y, -(a.y + c.x));
}

template <class PL, bool INV, bool UNIFORM, bool NT>
__device__ __forceinline__ void last_stage_pack_mirror(float2* v, int t, const float2* __restrict__ tw,
                                                       StageTwiddles<PL, PL::NSTAGES - 1, INV, true>& tws,
                                                       float2 (&wl)[PL::P / (1 << PL::rlog(PL::NSTAGES - 1)) / 2],
                                                       float2* __restrict__ dst, bool active, float* part,
                                                       float* stage)
{
    constexpr int S = PL::NSTAGES - 1;
    constexpr int N = PL::N;
    constexpr int R = 1 << PL::rlog(S);
    constexpr int B = PL::P / R;
    constexpr int H = B / 2;
    constexpr int M = N / R;                    // the last stage writes point j + r*M
    const float2* lvl = tw + 2 * N;
    const int jh0 = bfly<PL, S, true>(t, H);    // M - t, or M/2 for t == 0
    if constexpr (R > 8) {
        // One mirror pair of butterflies per thread (the radix-32 last stage of 2^15).
        // Every thread but 0 packs its pairs with pack_pair_cse; thread 0's butterflies 0
        // and M/2 pair within themselves (element (0, q), point q*M, with element (0, R-q);
        // element (1, q), point M/2 + q*M, with element (1, R-1-q)), so thread 0 puts their
        // 2R values in `part` and the 2R lanes of its wave pack one point each.
        // A row of N + 1 outputs that does not start on a 128-byte line (batched rows
        // follow each other with a stride of N + 1 values) goes out through LDS in two
        // halves, [0, N/2) and [N/2, N], each stored with every wave writing whole lines:
        // stores that cover lines only partly cost far more HBM time than whole-line
        // stores (complex 2^15 x 256 measured 36 us with its output on a line boundary,
        // 45 us with it 8 bytes off).  A row on a line boundary is stored directly.
        static_assert(H == 1 && 2 * R == 64, "one mirror pair of butterflies per thread, 2R lanes in a wave");
        tws.settle();
        settle_reads(wl);
        butterfly_compute<PL, S, INV, true>(v, tws, 0);
        butterfly_compute<PL, S, INV, true>(v, tws, 1);
        float2 sp = make_float2(0.0f, 0.0f), nyq = make_float2(0.0f, 0.0f);
        int spk = 0;                              // the point lane t of thread 0's wave packs
        if (t < 2 * R) {
            float2* pp = reinterpret_cast<float2*>(part);
            if (t == 0) {
#pragma unroll
                for (int e = 0; e < 2 * R; ++e) pp[e] = v[e];
                nyq = pack_one(v[0], v[0], edge_twiddle<N, R>(lvl, N));   // the Nyquist output
            }
            wave_lds_sync();
            const int q = t & (R - 1);
            if (t < R) {
                spk = q * M;
                sp = pack_one(pp[q], pp[(R - q) & (R - 1)], cmul(lvl[0], w64(q * (32 / R))));
            } else {
                spk = M / 2 + q * M;
                sp = pack_one(pp[R + q], pp[R + R - 1 - q], half_step_twiddle<R>(R - 1 - q));
            }
        }
        // in place: v[r] = out[t + rM], v[2R-1-r] = out[(M - t) + (R-1-r)M]
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float2 o0, o1;
            pack_pair_cse(v[r], v[R + R - 1 - r], cmul(wl[0], w64(r * (32 / R))), o0, o1);
            v[r] = o0;
            v[R + R - 1 - r] = o1;
        }
        if (((unsigned long long)dst & 127u) == 0) {
            // a row that starts on a line: the direct stores already cover whole lines
            if (t != 0 && active) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    st_out<NT>(&dst[t + r * M], v[r]);
                    st_out<NT>(&dst[(M - t) + (R - 1 - r) * M], v[R + R - 1 - r]);
                }
            }
            if (t < 2 * R && active) st_out<NT>(&dst[spk], sp);
            if (t == 0 && active) st_out<NT>(&dst[N], nyq);
            return;
        }
        float2* stg = reinterpret_cast<float2*>(stage);   // the exchange planes are free now
        // points [c0, c0 + lead) share a line with the part before them; from c0 + lead on,
        // every 64 points a wave stores are 4 whole lines (dst + c0 is equally far from a
        // line in both halves: N/2 points are whole lines).  The staged half sits `sh`
        // slots up so that the pairs read back from c0 + lead start on 16 bytes and load
        // as single 16-byte LDS reads (8-byte-aligned pairs read as two halves that
        // conflict on the banks: 247 conflict cycles per wave).
        const unsigned mis = (unsigned)(((unsigned long long)dst) & 127u);
        const int lead = mis ? (int)((128u - mis) >> 3) : 0;
        const int sh = lead & 1;
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int c0 = half ? N / 2 : 0;
            const int c1 = half ? N + 1 : N / 2;
            float2* sg = stg + sh - c0;                    // sg[k]: point k of this half
            // point t + rM lies in the upper half for r >= R/2; its mirror, for r < R/2
            if (t != 0 && active) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if ((r >= R / 2) == (half == 1)) sg[t + r * M] = v[r];
                    else sg[(M - t) + (R - 1 - r) * M] = v[R + R - 1 - r];
                }
            }
            if (t < 2 * R && active && (spk >= N / 2) == (half == 1)) sg[spk] = sp;
            if (half == 1 && t == 0 && active) sg[N] = nyq;
            __syncthreads();
            if (active) {
                const int a = c0 + lead;
                if (t < lead) st_out<NT>(&dst[c0 + t], sg[c0 + t]);
                const int cnt = c1 - a;
                // pairs of points as 16-byte stores (dst + a is on a line): half the store
                // instructions of one point per lane, forward 41.4 -> 40.7 us from HBM
                float4* dst4 = reinterpret_cast<float4*>(dst + a);
                const float4* sg4 = reinterpret_cast<const float4*>(sg + a);   // sh + lead is even
#pragma unroll
                for (int j = 0; j < (N / 2 + 1 + 2 * PL::T - 1) / (2 * PL::T); ++j) {
                    const int i = t + j * PL::T;
                    if (2 * i + 1 < cnt) {
                        st_out4<NT>(&dst4[i], sg4[i]);
                    } else if (2 * i < cnt) {
                        st_out<NT>(&dst[a + 2 * i], sg[a + 2 * i]);
                    }
                }
            }
            if (half == 0) __syncthreads();
        }
        return;
    }
    __amdgpu_buffer_rsrc_t rs;
    unsigned vo0 = 0, voh0 = 0, vom = 0;
    if constexpr (UNIFORM) {
        static_assert(pack_split_exact<PL>(), "pack store index split");
        rs = __builtin_amdgcn_make_buffer_rsrc(uniform_base(dst), (short)0, (N + 1) * (int)sizeof(float2),
                                               kRawBufferFormat);
        vo0 = (unsigned)t * (unsigned)sizeof(float2);
        voh0 = (unsigned)jh0 * (unsigned)sizeof(float2);
        vom = (unsigned)(M - t - (H - 1) * PL::T) * (unsigned)sizeof(float2);
    }
    tws.settle();
    settle_reads(wl);
#pragma unroll
    for (int b = 0; b < H; ++b) {
        butterfly_compute<PL, S, INV, true>(v, tws, b);
        butterfly_compute<PL, S, INV, true>(v, tws, b + H);
        if (!active) continue;
        const int jl = t + b * PL::T;
        const int jh = b == 0 ? jh0 : M - jl;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const float2 lo = v[b * R + r];
            const float2 hi = v[(b + H) * R + R - 1 - r];
            const int k0 = jl + r * M, k1 = jh + (R - 1 - r) * M;
            const float2 w0 = cmul(wl[b], w64(r * (32 / R)));   // W^(jl + rM) = W^jl exp(-2 pi i r/2R)
            float2 p0 = hi, p1 = lo;
            float2 w1 = make_float2(-w0.x, w0.y);   // W^(N-k0) = -conj(W^k0)
            if (b == 0) {
                // thread 0: point r*M pairs with (R-r)*M, point M/2 + (R-1-r)*M with M/2 + r*M,
                // whose twiddle is exp(-2 pi i (2(R-1-r)+1) / 4R)
                const bool z = (t == 0);
                p0 = z ? v[(R - r) & (R - 1)] : p0;
                p1 = z ? v[H * R + r] : p1;
                w1 = z ? half_step_twiddle<R>(r) : w1;
            }
            if constexpr (UNIFORM) {
                const int c0 = b * PL::T + r * M;
                const int c1 = b == 0 ? (R - 1 - r) * M : (H - 1 - b) * PL::T + (R - 1 - r) * M;
                __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(StoreBits, pack_one(lo, p0, w0)), rs, vo0,
                                                      c0 * (int)sizeof(float2), kStoreAux<NT>);
                __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(StoreBits, pack_one(hi, p1, w1)), rs,
                                                      b == 0 ? voh0 : vom, c1 * (int)sizeof(float2), kStoreAux<NT>);
            } else {
                st_out<NT>(&dst[k0], pack_one(lo, p0, w0));
                st_out<NT>(&dst[k1], pack_one(hi, p1, w1));
            }
        }
    }
    if (active && t == 0) st_out<NT>(&dst[N], pack_one(v[0], v[0], edge_twiddle<N, R>(lvl, N)));   // the Nyquist output
}

// Every output index of stage S (no mirror) equals IndexSplit's U(t) + C(b, r) and lies in
// [0, N): checked for every thread, butterfly and element at compile time.
template <class PL, int S>
constexpr bool store_split_exact()
{
    using IS = IndexSplit<PL, S, false, true>;
    for (int t = 0; t < PL::T; ++t)
        for (int b = 0; b < IS::B; ++b)
            for (int r = 0; r < IS::R; ++r) {
                const int i = IS::U(t, 0) + IS::C(b, r);
                if (i != IS::reference(t, b, r) || i < 0 || i >= PL::N) return false;
            }
    return true;
}

// The last stage of a complex or real-inverse transform fused with its stores, butterfly
// by butterfly.  The stage's twiddles (tws) are read beforehand: a wait for a read issued
// after the first stores would also wait for those stores.
// UNIFORM (one transform per workgroup): the stores go through a buffer resource over the
// signal, each index split as IndexSplit's per-thread part (a VGPR byte offset, computed
// once) plus its compile-time part (a scalar offset), so no store needs 64-bit address
// arithmetic.  The resource spans the signal's N points; every index is below N.
template <class PL, bool INV, bool UNIFORM, bool NT>
__device__ __forceinline__ void last_stage_store(float2* v, int t,
                                                 StageTwiddles<PL, PL::NSTAGES - 1, INV, false>& tws,
                                                 float2* __restrict__ dst, bool active)
{
    constexpr int S = PL::NSTAGES - 1;
    constexpr int R = 1 << PL::rlog(S);
    constexpr int B = PL::P / R;
    tws.settle();
    if constexpr (UNIFORM) {
        using IS = IndexSplit<PL, S, false, true>;
        static_assert(store_split_exact<PL, S>(), "store index split");
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc(uniform_base(dst), (short)0, PL::N * (int)sizeof(float2), kRawBufferFormat);
        const unsigned vo = (unsigned)IS::U(t, 0) * (unsigned)sizeof(float2);
#pragma unroll
        for (int b = 0; b < B; ++b) {
            butterfly_compute<PL, S, INV, false>(v, tws, b);
            if (active) {
#pragma unroll
                for (int r = 0; r < R; ++r)
                    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(StoreBits, v[b * R + r]), rs, vo,
                                                          IS::C(b, r) * (int)sizeof(float2), kStoreAux<NT>);
            }
        }
    } else {
#pragma unroll
        for (int b = 0; b < B; ++b) {
            butterfly_compute<PL, S, INV, false>(v, tws, b);
            if (active) {
#pragma unroll
                for (int r = 0; r < R; ++r) st_out<NT>(&dst[out_index<PL, S>(t + b * PL::T, r)], v[b * R + r]);
            }
        }
    }
}

// Real forward epilogue when the last stage cannot be mirror-closed: Z goes through
// LDS, then thread t packs pairs k = t + q*T for k in [0, N/2].
template <class PL, bool SPLIT, bool NT>
__device__ __forceinline__ void pack_epilogue(float2* v, int t, float* re, float* im,
                                              const float2* __restrict__ tw, float2* __restrict__ dst,
                                              bool active)
{
    constexpr int N = PL::N;
    constexpr int T = PL::T;
    constexpr int S = PL::NSTAGES - 1;
    constexpr int Q = (N / 2) / T + 1;                      // pair slots per thread
    using LY = ExchangeLayout<PL, S>;                       // the layout stage S writes
    const float2* lvl = tw + 2 * N;
    float2 zk[Q], zm[Q];
    if constexpr (!SPLIT) {
        lds_write<PL, S, 3>(v, t, re, im);
        exchange_sync<PL>();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int k = t + q * T;
            if (k <= N / 2) {
                const int ik = LY::pad(k & (N - 1)), im_ = LY::pad((N - k) & (N - 1));
                zk[q] = make_float2(re[ik], im[ik]);
                zm[q] = make_float2(re[im_], im[im_]);
            }
        }
    } else {
        lds_write<PL, S, 1>(v, t, re, im);
        exchange_sync<PL>();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int k = t + q * T;
            if (k <= N / 2) {
                zk[q].x = re[LY::pad(k & (N - 1))];
                zm[q].x = re[LY::pad((N - k) & (N - 1))];
            }
        }
        exchange_sync<PL>();
        lds_write<PL, S, 2>(v, t, re, im);
        exchange_sync<PL>();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int k = t + q * T;
            if (k <= N / 2) {
                zk[q].y = im[LY::pad(k & (N - 1))];
                zm[q].y = im[LY::pad((N - k) & (N - 1))];
            }
        }
    }
    if (!active) return;
    constexpr int RE = 1 << PL::rlog(S);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int k = t + q * T;
        if (k <= N / 2) pack_pair<N, RE, NT>(k, zk[q], zm[q], lvl, dst);
    }
}

// ---- one workgroup's transform body ---------------------------------------------------

// Shape of a transform class in a workgroup of WG threads.  KIND: kComplexForward /
// kComplexInverse / kRealForward / kRealInverse; for the real kinds N = 2^LOGN is half
// the real length.  The real forward plan puts its small radix last so that the last
// stage can be mirror-closed.
template <int LOGN, int LOGP, int WG, int KIND, bool SPLIT>
struct LdsShape {
    using PL = Plan<LOGN, LOGP, KIND == kRealForward>;
    static constexpr bool INV = (KIND == kComplexInverse || KIND == kRealInverse);
    static constexpr int N = PL::N;
    static constexpr int T = PL::T;
    static constexpr int TPW = WG / T;                // transforms per workgroup
    static_assert(TPW >= 1 && TPW * T == WG, "workgroup must hold whole transforms");
    static_assert(!SPLIT || TPW == 1, "plane-by-plane exchange is for one transform per workgroup");
    static constexpr int SPAN = plan_span<PL>();
    // LDS holds the exchanges (any plan of more than one stage, which includes every plan
    // of more than one thread)
    static constexpr int MS = mirror_stage<PL, KIND>();
    // a mirror-closed unpack of edge radix above 8 keeps thread 0's 2R unpacked values in
    // LDS after the exchange planes (one transform per workgroup)
    // a mirror-closed pack or unpack of edge radix above 8 passes thread 0's 2R points
    // through LDS after the exchange planes (one transform per workgroup)
    static constexpr int EDGE_LOG = MS < 0 ? 0 : PL::rlog(MS);
    static constexpr bool WIDE_EDGE = MS >= 0 && EDGE_LOG > 3;
    static_assert(!WIDE_EDGE || TPW == 1, "the wide pack/unpack keeps one transform's values");
    static constexpr int PART_OFFSET = SPLIT ? SPAN : 2 * SPAN * TPW;
    static constexpr int LDS_WORDS = PL::NSTAGES > 1 ? PART_OFFSET + (WIDE_EDGE ? 4 << EDGE_LOG : 0) : 1;
    // the wide pack stages each half of the N+1 outputs (N + 2 floats, one slot up at most)
    // below PART_OFFSET
    static_assert(!(WIDE_EDGE && KIND == kRealForward) || N + 4 <= PART_OFFSET, "pack staging fits the planes");
    // complex elements per signal on each side
    static constexpr int IN_STRIDE = KIND == kRealInverse ? N + 1 : N;
    static constexpr int OUT_STRIDE = KIND == kRealForward ? N + 1 : N;
};

// The transform owned by this thread's slot (tid / T) of a workgroup: read `src`, write
// `dst` (that signal's buffers), with `lds` holding LdsShape::LDS_WORDS floats.  A slot
// with `active` false computes (on a valid signal) but stores nothing, so every load
// issues unpredicated.  Used by the batch kernel below and by the persistent kernel.
// NT: the transform's input reads and output writes carry the non-temporal hint (see
// ld_in / st_out).
template <int LOGN, int LOGP, int WG, int KIND, bool SPLIT, bool NT = true>
__device__ __forceinline__ void lds_transform(const float2* __restrict__ src, float2* __restrict__ dst, bool active,
                                              const float2* __restrict__ tw, float* lds, int tid)
{
    using SH = LdsShape<LOGN, LOGP, WG, KIND, SPLIT>;
    using PL = typename SH::PL;
    constexpr bool INV = SH::INV;
    constexpr int N = SH::N;
    constexpr int T = SH::T;
    constexpr int SPAN = SH::SPAN;
    constexpr int MS = SH::MS;

    const int tr = tid / T;
    const int t = tid - tr * T;
    float* re = lds + (SPLIT ? 0 : tr * 2 * SPAN);
    float* im = SPLIT ? lds : re + SPAN;

    float2 v[PL::P];
    if constexpr (T == 1) {
        // whole transform in this thread's registers, one stage (n <= 16 points): its
        // outputs sit in natural order, v[k] = Z[k], which the pack below relies on
        static_assert(PL::NSTAGES == 1, "a one-thread transform is a single stage");
        if constexpr (KIND == kRealInverse)
            load_first_unpack<PL>(v, t, src, tw);
        else
            load_first_complex<PL, NT>(v, t, src);
        stage_compute<PL, 0, INV>(v, t, tw);
        if constexpr (KIND == kRealForward) {
            if (active) {
                const float2* lvl = tw + 2 * N;
#pragma unroll
                for (int k = 0; k <= N / 2; ++k)
                    pack_pair<N, N, NT>(k, v[k & (N - 1)], v[(N - k) & (N - 1)], lvl, dst);
            }
        } else {
            store_last<PL, NT>(v, t, dst, active);
        }
    } else {
        static_assert(PL::NSTAGES >= 2, "a transform over several threads has at least two stages");
        constexpr int SL = PL::NSTAGES - 1;
        // first stage, fused with the loads (and the real-inverse unpack)
        if constexpr (KIND == kRealInverse && MS == 0) {
            first_stage_unpack_mirror<PL, INV, NT>(v, t, src, tw, lds + SH::PART_OFFSET);
        } else if constexpr (KIND == kRealInverse) {
            load_first_unpack<PL>(v, t, src, tw);
            stage_compute<PL, 0, INV>(v, t, tw);
        } else {
            load_first_complex<PL, NT>(v, t, src);
            stage_compute<PL, 0, INV>(v, t, tw);
        }
        // middle stages, then the exchange into the last one
        run_stages<PL, 1, SL, INV, SPLIT, MS>(v, t, re, im, tw);
        // last stage, fused with the stores (and the real-forward pack); its twiddle reads
        // are issued before the exchange that brings its inputs
        if constexpr (KIND == kRealForward && MS == SL) {
            constexpr int H = bflies<PL, SL>() / 2;
            StageTwiddles<PL, SL, INV, true> tws(t, tw);
            float2 wl[H];
#pragma unroll
            for (int b = 0; b < H; ++b) wl[b] = tw[2 * N + t + b * T];
            exchange<PL, SL - 1, SL, SPLIT, MS == SL - 1, true>(v, t, re, im);
            last_stage_pack_mirror<PL, INV, SH::TPW == 1, NT>(v, t, tw, tws, wl, dst, active, lds + SH::PART_OFFSET,
                                                          lds);
        } else if constexpr (KIND == kRealForward) {
            exchange<PL, SL - 1, SL, SPLIT, MS == SL - 1, false>(v, t, re, im);
            stage_compute<PL, SL, INV>(v, t, tw);
            pack_epilogue<PL, SPLIT, NT>(v, t, re, im, tw, dst, active);
        } else if constexpr (T < 32) {
            // several transforms per half-wave (small sizes): reading the twiddles ahead
            // measured slower here
            exchange<PL, SL - 1, SL, SPLIT, MS == SL - 1, false>(v, t, re, im);
            stage_compute<PL, SL, INV>(v, t, tw);
            store_last<PL, NT>(v, t, dst, active);
        } else {
            StageTwiddles<PL, SL, INV, false> tws(t, tw);
            exchange<PL, SL - 1, SL, SPLIT, MS == SL - 1, false>(v, t, re, im);
            last_stage_store<PL, INV, SH::TPW == 1, NT>(v, t, tws, dst, active);
        }
    }
}

}  // namespace ckfft_gpu
