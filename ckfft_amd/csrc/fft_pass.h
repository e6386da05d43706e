// fft_pass.h -- one tile of a multi-pass (global-memory Stockham) stage, shared by the
// pass kernels (fft_pass.hip) and the persistent kernel (fft_batch.hip).
#pragma once
#include "fft_stockham.h"

namespace ckfft_gpu {

// exp(-2 pi i m / 2^loglev) from the device tables: the pyramid up to 2^16, above that
// a coarse pyramid entry times a fine-table entry (both FP64-rounded; one product).
__device__ __forceinline__ float2 twiddle_at(const float2* __restrict__ pyr, const float2* __restrict__ fine,
                                             int loglev, long long m)
{
    if (loglev <= kLogMaxPyramid) return pyr[(1ll << loglev) + m];
    const float2 c = pyr[(1ll << (loglev - kLogFine)) + (m >> kLogFine)];
    const float2 f = fine[((long long)(loglev - kLogMaxPyramid - 1) << kLogFine) + (m & ((1 << kLogFine) - 1))];
    return cmul(c, f);
}

// Tile order: contiguous tile ranges per XCD (workgroup i runs on XCD i mod 8).
__device__ __forceinline__ long long xcd_tile(long long bid, long long nblocks)
{
    const long long per = nblocks >> 3;
    if (bid >= (per << 3)) return bid;
    return (bid & 7) * per + (bid >> 3);
}

// Tile width of the stand-alone pass kernels: 16 butterflies (128-byte runs) up to radix
// 2^10, fewer above so that a tile holds <= 16384 points, more below so that a
// workgroup has >= 256 threads.
constexpr int default_logcb(int logr) { return logr >= 8 ? (logr <= 10 ? 4 : 14 - logr) : 12 - logr; }

template <int LOGR, int LOGCB_ = default_logcb(LOGR)>
struct PassShape {
    static constexpr int LOGP = 4;                                        // 16 points per thread
    static constexpr int R = 1 << LOGR;
    static constexpr int T = R >> LOGP;                                   // threads per DFT
    static constexpr int LOGCB = LOGCB_;
    static constexpr int CB = 1 << LOGCB;
    static constexpr int WG = CB * T;
    static constexpr int SPAN = plan_span<Plan<LOGR, LOGP>>();
    static constexpr int RSTRIDE = 2 * SPAN + 1;   // per-DFT LDS region; odd so regions start on different banks
};

template <class PL, int S, bool INV, bool FIRST, class SH>
__device__ __forceinline__ void pass_stages(float2* v, int tA, float* reA, float* imA, int tB, float* reB,
                                            float* imB, const float2* __restrict__ pyr)
{
    if constexpr (S < PL::NSTAGES) {
        // the first pass stores with "point-fastest" lanes (contiguous output block); so the
        // last sub-stage switches mapping B; every other sub-stage uses "DFT-fastest" lanes (A)
        constexpr bool B_LAST = FIRST && (S == PL::NSTAGES - 1);
        lds_write<PL, S - 1, 3>(v, tA, reA, imA);
        __syncthreads();
        if constexpr (B_LAST) {
            lds_read<PL, S, 3>(v, tB, reB, imB);
        } else {
            lds_read<PL, S, 3>(v, tA, reA, imA);
        }
        __syncthreads();
        stage_compute<PL, S, INV>(v, B_LAST ? tB : tA, pyr);
        pass_stages<PL, S + 1, INV, FIRST, SH>(v, tA, reA, imA, tB, reB, imB, pyr);
    }
}

// Inter-pass twiddles of butterfly column jm in a pass of span 2^logns: the thread's point r
// (r = in_index of its first-stage element) is multiplied by W_L^(r jm), L = 2^logns * R.
// XTW: 0 = twiddles from the pyramid / fine tables; 2 = from the transposed copy, which
// the caller has checked covers the pass (span 2^kLogXtwCols, level <= 2^log_xtw).  The
// host picks one per kernel, so a kernel has a single code path.
// fetch() issues the table reads and apply() forms the twiddles and multiplies the points,
// so that a tile can issue the reads before its data loads and their latency overlaps
// (at level 2^24 the reads cost as much as 40% of the pass when issued after the data).
// The values and products are those of reading each twiddle where it is used.
template <class PL, bool INV, int XTW>
struct PassTwiddles {
    static_assert(XTW == 0 || XTW == 2, "twiddle source");
    static constexpr int LOGR = PL::nslog(PL::NSTAGES);       // log2 of the pass radix
    static constexpr int R = PL::N, T = PL::T;
    static constexpr int LR0 = PL::rlog(0);
    static constexpr int R0 = 1 << LR0;
    static constexpr int B0 = PL::P / R0;
    static constexpr int M0 = R / R0;
    // XTW 2: one value per point.  XTW 0: the thread's points are r = tA + b*T + rr*M0, so
    // W^(r jm) = W^((tA + b T) jm) * (W^(M0 jm))^rr: log2(R0) base powers, then B0 row
    // values, read once per thread instead of once per point.  Above 2^16 each value is a
    // coarse pyramid entry times a fine-table entry; gathering one per point made the pass
    // bound by the texture-address unit.
    static constexpr int NV = XTW == 2 ? PL::P : LR0 + B0;
    float2 f[NV];                       // pyramid, fine-table or transposed entries
    float2 c[XTW == 2 ? 1 : NV];        // coarse pyramid entries (levels above 2^16)
    bool coarse;

    __device__ __forceinline__ void fetch(int tA, long long jm, int logns, const float2* __restrict__ pyr,
                                          const float2* __restrict__ fine, const float2* __restrict__ xtw,
                                          int log_xtw)
    {
        const int loglev = logns + LOGR;
        if constexpr (XTW == 2) {
            // neighbouring columns, neighbouring words; point r's row starts at word
            // r << sh, and r = tA + in_index(b*T, rr) (in_index is linear in the butterfly)
            const int sh = log_xtw - loglev + kLogXtwCols;
            const float2* tcol = xtw + jm + ((long long)tA << sh);
#pragma unroll
            for (int b = 0; b < B0; ++b)
#pragma unroll
                for (int rr = 0; rr < R0; ++rr) f[b * R0 + rr] = tcol[in_index<PL, 0>(b * T, rr) << sh];
        } else {
            // every exponent is below the level L <= 2^28: 32-bit arithmetic
            const unsigned jmu = (unsigned)jm;
            unsigned m[NV];
#pragma unroll
            for (int k = 0; k < LR0; ++k) m[k] = (unsigned)(M0 << k) * jmu;
#pragma unroll
            for (int b = 0; b < B0; ++b) m[LR0 + b] = (unsigned)(tA + b * T) * jmu;
            coarse = loglev > kLogMaxPyramid;
            if (coarse) {
                const float2* cl = pyr + (1ll << (loglev - kLogFine));
                const float2* fl = fine + ((long long)(loglev - kLogMaxPyramid - 1) << kLogFine);
#pragma unroll
                for (int i = 0; i < NV; ++i) {
                    c[i] = cl[m[i] >> kLogFine];
                    f[i] = fl[m[i] & ((1 << kLogFine) - 1)];
                }
            } else {
                const float2* lvl = pyr + (1ll << loglev);
#pragma unroll
                for (int i = 0; i < NV; ++i) f[i] = lvl[m[i]];
            }
        }
    }

    // entry i as twiddle_at forms it
    __device__ __forceinline__ float2 value(int i) const { return coarse ? cmul(c[i], f[i]) : f[i]; }

    __device__ __forceinline__ void apply(float2* v) const
    {
        if constexpr (XTW == 2) {
#pragma unroll
            for (int i = 0; i < PL::P; ++i) v[i] = cmul(v[i], dirw<INV>(f[i]));
        } else {
            TwBase<R0> sb;
#pragma unroll
            for (int k = 0; k < LR0; ++k) sb.w[k] = dirw<INV>(value(k));
            float2 step[R0];
            twiddle_powers<R0>(step, sb);
#pragma unroll
            for (int b = 0; b < B0; ++b) {
                const float2 wb = dirw<INV>(value(LR0 + b));
                v[b * R0] = cmul(v[b * R0], wb);
#pragma unroll
                for (int rr = 1; rr < R0; ++rr) v[b * R0 + rr] = cmul(v[b * R0 + rr], cmul(wb, step[rr]));
            }
        }
    }
};

// One tile of one pass: butterflies q*CB .. q*CB + CB-1 of the signal at `src` (2^logn
// points), written to `dst`; logns = log2 of the pass span (0 for the first pass).
// `lds` holds CB * RSTRIDE floats.  Used by the pass kernels here and by the batch
// plans' table-driven passes.  XTW as in PassTwiddles.
// The tile's input reads carry the non-temporal hint (each value is read once); its writes
// do when NT_OUT, for the pass that writes the transform's output.  The intermediate of a
// later pass is written normally: with the hint on it too, 2^24 took 187 us against 135.
template <int LOGR, int LOGCB, bool INV, bool FIRST, int XTW, bool NT_OUT>
__device__ __forceinline__ void pass_tile(const float2* __restrict__ src, float2* __restrict__ dst, long long q,
                                          int logn, int logns, const float2* __restrict__ pyr,
                                          const float2* __restrict__ fine, const float2* __restrict__ xtw,
                                          int log_xtw, float* lds, int tid)
{
    using SH = PassShape<LOGR, LOGCB>;
    using PL = Plan<LOGR, SH::LOGP>;
    static_assert(PL::NSTAGES >= 2, "a pass needs at least two sub-stages");
    constexpr int R = SH::R, T = SH::T, CB = SH::CB;
    const long long j0 = q * CB;

    const int cA = tid & (CB - 1), tA = tid >> SH::LOGCB;  // DFT-fastest lanes
    const int cB = tid / T, tB = tid - cB * T;             // point-fastest lanes
    float* reA = lds + cA * SH::RSTRIDE;
    float* imA = reA + SH::SPAN;
    float* reB = lds + cB * SH::RSTRIDE;
    float* imB = reB + SH::SPAN;

    float2 v[PL::P];
    constexpr int R0 = 1 << PL::rlog(0);
    constexpr int B0 = PL::P / R0;
    PassTwiddles<PL, INV, XTW> ptw;
    if constexpr (!FIRST) ptw.fetch(tA, (j0 + cA) & ((1ll << logns) - 1), logns, pyr, fine, xtw, log_xtw);
    // point r = tA + in_index(b*T, rr) of column j0 + cA: a per-thread base plus offsets
    // that are the same for every thread
    const int logstride = logn - LOGR;
    const float2* srct = src + j0 + cA + ((long long)tA << logstride);
#pragma unroll
    for (int b = 0; b < B0; ++b)
#pragma unroll
        for (int rr = 0; rr < R0; ++rr) v[b * R0 + rr] = ld_in<true>(&srct[(long long)in_index<PL, 0>(b * T, rr) << logstride]);
    if constexpr (!FIRST) ptw.apply(v);
    stage_compute<PL, 0, INV>(v, tA, pyr);
    pass_stages<PL, 1, INV, FIRST, SH>(v, tA, reA, imA, tB, reB, imB, pyr);

    constexpr int SL = PL::NSTAGES - 1;
    constexpr int RL = 1 << PL::rlog(SL);
    constexpr int BL = PL::P / RL;
    if constexpr (FIRST) {
        // span 1: butterfly j writes points j*R .. j*R + R-1, one contiguous block per tile
        float2* blk = dst + (j0 + cB) * R;
#pragma unroll
        for (int b = 0; b < BL; ++b)
#pragma unroll
            for (int rr = 0; rr < RL; ++rr) st_out<NT_OUT>(&blk[out_index<PL, SL>(tB + b * T, rr)], v[b * RL + rr]);
    } else {
        const long long ns = 1ll << logns;
        const long long j = j0 + cA;
        float2* base = dst + (j >> logns) * ns * R + (j & (ns - 1));
#pragma unroll
        for (int b = 0; b < BL; ++b)
#pragma unroll
            for (int rr = 0; rr < RL; ++rr)
                st_out<NT_OUT>(&base[(unsigned)out_index<PL, SL>(tA + b * T, rr) << logns], v[b * RL + rr]);   // < N <= 2^28
    }
}

}  // namespace ckfft_gpu
