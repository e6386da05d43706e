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
// Real transforms above 2^16 samples run the half-size complex transform here with the
// pack fused into its last pass (forward) or the unpack into its first (inverse); below
// that they stay fused in the one-workgroup kernels.
#include <utility>

#include "fft_pass.h"

namespace ckfft_gpu {

// XTW_PASS: the host has checked that the transposed twiddles cover this pass (span
// 2^kLogXtwCols, level <= 2^log_xtw; possible for radix <= 2^8 only); otherwise the
// twiddles come from the pyramid and fine tables.  One code path per kernel.
// LAST: the pass writes the transform's output (with the non-temporal hint).
template <int LOGR, bool INV, bool FIRST, bool XTW_PASS, bool LAST>
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
    static_assert(!XTW_PASS || LOGR + kLogXtwCols <= kLogMaxPyramid, "transposed twiddles cover radix <= 2^8");
    constexpr int XTW = XTW_PASS ? 2 : 0;
    pass_tile<LOGR, SH::LOGCB, INV, FIRST, XTW, LAST>(in + g * N, out + g * N, q, logn, logns, pyr, fine, xtw, log_xtw,
                                                lds, threadIdx.x);
}

// ---- real pack / unpack fused into the edge pass -------------------------------------
//
// The pack of a real forward pairs output point k of the half-size complex transform with
// h-k, the unpack of a real inverse input point k with h-k.  In a pass over `ncols`
// columns (butterflies), point c + r*ncols pairs with (ncols - c) + (R-1-r)*ncols for
// c >= 1, column 0 pairs with itself (element R-r) and so does column ncols/2 (element
// R-1-r).  A mirror tile takes CB/2 low columns in its first CB/2 lanes and their mirrors
// in the others (the slot of column 0's mirror takes column ncols/2), so every pair meets
// in one workgroup and goes through its LDS.  Fusing saves a read and a write of the
// whole half spectrum.  The pack / unpack twiddle of point c + r*ncols is read as
// W_2h^c * W_2R^r: one fine-table read per column (shared by the column's points) and a
// pyramid entry, instead of a fine-table gather per point.

// Real-inverse unpack of point k: t[k] = (Y[k] + conj Y[h-k]) + i conj(W^k) (Y[k] - conj Y[h-k]).
__device__ __forceinline__ float2 unpack_point(float2 yk, float2 ym, float2 w)
{
    const float2 a = make_float2(yk.x + ym.x, yk.y - ym.y);
    const float2 d = make_float2(yk.x - ym.x, yk.y + ym.y);
    const float2 c = cmul(cconj(w), d);
    return make_float2(a.x - c.y, a.y + c.x);
}

// Column of lane c (0 <= c < CB) of mirror tile q.
__device__ __forceinline__ long long mirror_col(long long q, int c, int cb, long long ncols)
{
    const long long l = q * (cb / 2) + (c < cb / 2 ? c : c - cb / 2);
    if (c < cb / 2) return l;
    return l == 0 ? ncols / 2 : ncols - l;
}

// Real inverse: the first pass of the inverse half-size transform, reading the unpacked
// input t[k] = unpack(Y[k], Y[h-k], W^k) directly from Y (h+1 values per signal).
template <int LOGR>
__global__ __launch_bounds__(PassShape<LOGR>::WG) void fft_pass_unpack_kernel(
    const float2* __restrict__ y, float2* __restrict__ out, const float2* __restrict__ pyr,
    const float2* __restrict__ fine, int logn, long long tiles_per_signal, long long total_tiles)
{
    using SH = PassShape<LOGR>;
    using PL = Plan<LOGR, SH::LOGP>;
    constexpr int R = SH::R, T = SH::T, CB = SH::CB;
    constexpr int R0 = 1 << PL::rlog(0), B0 = PL::P / R0;
    __shared__ float lds[CB * SH::RSTRIDE];
    const long long tile = xcd_tile(blockIdx.x, total_tiles);
    const long long g = tile / tiles_per_signal;
    const long long q = tile - g * tiles_per_signal;
    const long long h = 1ll << (logn - 1);
    const long long ncols = h >> LOGR;
    const float2* ys = y + g * (h + 1);
    const int tid = threadIdx.x;
    const int cA = tid & (CB - 1), tA = tid >> SH::LOGCB;
    const int cB = tid / T, tB = tid - cB * T;
    const long long col = mirror_col(q, cA, CB, ncols);
    const float2 wcol = twiddle_at(pyr, fine, logn, col);   // W_2h^col, read ahead of the data

    float2 v[PL::P];
#pragma unroll
    for (int b = 0; b < B0; ++b)
#pragma unroll
        for (int rr = 0; rr < R0; ++rr) v[b * R0 + rr] = ld_in<true>(&ys[col + (long long)in_index<PL, 0>(tA + b * T, rr) * ncols]);
    const float2 y_last = ys[h];   // the mirror of point 0
    // every lane's values into its column's LDS row, then each point's mirror back out
    float* row = lds + cA * SH::RSTRIDE;
#pragma unroll
    for (int b = 0; b < B0; ++b)
#pragma unroll
        for (int rr = 0; rr < R0; ++rr) {
            const int r = in_index<PL, 0>(tA + b * T, rr);
            row[r] = v[b * R0 + rr].x;
            row[R + r] = v[b * R0 + rr].y;
        }
    __syncthreads();
    const bool self = col == 0 || col == ncols / 2;
    const float* mrow = lds + (self ? cA : (cA < CB / 2 ? cA + CB / 2 : cA - CB / 2)) * SH::RSTRIDE;
#pragma unroll
    for (int b = 0; b < B0; ++b)
#pragma unroll
        for (int rr = 0; rr < R0; ++rr) {
            const int r = in_index<PL, 0>(tA + b * T, rr);
            const int rm = col == 0 ? ((R - r) & (R - 1)) : (R - 1 - r);
            const float2 ym = (col == 0 && r == 0) ? y_last : make_float2(mrow[rm], mrow[R + rm]);
            v[b * R0 + rr] = unpack_point(v[b * R0 + rr], ym, cmul(wcol, pyr[2 * R + r]));
        }
    __syncthreads();

    float* reA = lds + cA * SH::RSTRIDE;
    float* imA = reA + SH::SPAN;
    float* reB = lds + cB * SH::RSTRIDE;
    float* imB = reB + SH::SPAN;
    stage_compute<PL, 0, true>(v, tA, pyr);
    pass_stages<PL, 1, true, true, SH>(v, tA, reA, imA, tB, reB, imB, pyr);
    // span 1: column c writes points c*R .. c*R + R-1
    constexpr int SL = PL::NSTAGES - 1;
    constexpr int RL = 1 << PL::rlog(SL);
    constexpr int BL = PL::P / RL;
    float2* blk = out + g * h + mirror_col(q, cB, CB, ncols) * R;
#pragma unroll
    for (int b = 0; b < BL; ++b)
#pragma unroll
        for (int rr = 0; rr < RL; ++rr) blk[out_index<PL, SL>(tB + b * T, rr)] = v[b * RL + rr];
}

// One pair of the real-forward pack, k <= h/2: out[k] and out[h-k] from Z[k] and Z[h-k].
__device__ __forceinline__ void pack_out(float2* __restrict__ o, long long k, long long h, float2 zk, float2 zm,
                                         float2 w)
{
    const float2 a = make_float2(zk.x + zm.x, zk.y - zm.y);
    const float2 d = make_float2(zk.x - zm.x, zk.y + zm.y);
    const float2 c = cmul(w, d);
    st_out<true>(&o[k], make_float2(a.x + c.y, a.y - c.x));
    if (k != h - k) st_out<true>(&o[h - k], make_float2(a.x - c.y, -(a.y + c.x)));
}

// Real forward: the last pass of the half-size transform (span ncols = h/R) writing the
// packed spectrum (h+1 values per signal) instead of Z.  XTW_PASS as in fft_pass_kernel.
template <int LOGR, bool XTW_PASS>
__global__ __launch_bounds__(PassShape<LOGR>::WG) void fft_pass_pack_kernel(
    const float2* __restrict__ z, float2* __restrict__ out, const float2* __restrict__ pyr,
    const float2* __restrict__ fine, const float2* __restrict__ xtw, int log_xtw, int logn,
    long long tiles_per_signal, long long total_tiles)
{
    using SH = PassShape<LOGR>;
    using PL = Plan<LOGR, SH::LOGP>;
    constexpr int R = SH::R, T = SH::T, CB = SH::CB, WG = SH::WG;
    constexpr int R0 = 1 << PL::rlog(0), B0 = PL::P / R0;
    static_assert(!XTW_PASS || LOGR + kLogXtwCols <= kLogMaxPyramid, "transposed twiddles cover radix <= 2^8");
    __shared__ float lds[CB * SH::RSTRIDE];
    const long long tile = xcd_tile(blockIdx.x, total_tiles);
    const long long g = tile / tiles_per_signal;
    const long long q = tile - g * tiles_per_signal;
    const long long h = 1ll << (logn - 1);
    const int logns = logn - 1 - LOGR;
    const long long ncols = 1ll << logns;
    const float2* zs = z + g * h;
    float2* os = out + g * (h + 1);
    const int tid = threadIdx.x;
    const int cA = tid & (CB - 1), tA = tid >> SH::LOGCB;
    const int cB = tid / T, tB = tid - cB * T;
    const long long col = mirror_col(q, cA, CB, ncols);

    float2 v[PL::P];
    PassTwiddles<PL, false, XTW_PASS ? 2 : 0> ptw;
    ptw.fetch(tA, col, logns, pyr, fine, xtw, log_xtw);
#pragma unroll
    for (int b = 0; b < B0; ++b)
#pragma unroll
        for (int rr = 0; rr < R0; ++rr) v[b * R0 + rr] = ld_in<true>(&zs[col + (long long)in_index<PL, 0>(tA + b * T, rr) * ncols]);
    ptw.apply(v);
    float* reA = lds + cA * SH::RSTRIDE;
    float* imA = reA + SH::SPAN;
    float* reB = lds + cB * SH::RSTRIDE;
    float* imB = reB + SH::SPAN;
    stage_compute<PL, 0, false>(v, tA, pyr);
    pass_stages<PL, 1, false, false, SH>(v, tA, reA, imA, tB, reB, imB, pyr);
    // Z of column c, element e (point c + e*ncols), into the column's LDS row
    constexpr int SL = PL::NSTAGES - 1;
    constexpr int RL = 1 << PL::rlog(SL);
    constexpr int BL = PL::P / RL;
    float* row = lds + cA * SH::RSTRIDE;
#pragma unroll
    for (int b = 0; b < BL; ++b)
#pragma unroll
        for (int rr = 0; rr < RL; ++rr) {
            const int e = out_index<PL, SL>(tA + b * T, rr);
            row[e] = v[b * RL + rr].x;
            row[R + e] = v[b * RL + rr].y;
        }
    __syncthreads();
    // pairs of the low columns: (column l, element r) with its mirror; consecutive threads
    // take consecutive columns of one element
    constexpr int HALF = CB / 2;
    static_assert((HALF * R) % WG == 0 && WG % HALF == 0, "pairs divide evenly over the workgroup");
    // a thread's pairs all come from one low column (p mod HALF = tid mod HALF)
    const int l = tid % HALF;
    const long long lc = q * HALF + l;
    const float2 wl = twiddle_at(pyr, fine, logn, lc);                    // W_2h^lc
    const float2 wm = twiddle_at(pyr, fine, logn, lc ? ncols - lc : 0);   // W_2h^(ncols-lc)
    const float* lrow = lds + l * SH::RSTRIDE;
    const float* mrow = lds + (l + HALF) * SH::RSTRIDE;
#pragma unroll
    for (int i = 0; i < HALF * R / WG; ++i) {
        const int r = (tid + i * WG) / HALF;
        const float2 zk = make_float2(lrow[r], lrow[R + r]);
        const long long k = lc + (long long)r * ncols;
        if (lc == 0) {
            // column 0 pairs with itself: (r, R-r), each pair once
            if (r > R / 2) continue;
            const int rm = (R - r) & (R - 1);
            pack_out(os, k, h, zk, make_float2(lrow[rm], lrow[R + rm]), pyr[2 * R + r]);
        } else {
            const float2 zm = make_float2(mrow[R - 1 - r], mrow[2 * R - 1 - r]);
            const long long km = h - k;
            if (k <= km)
                pack_out(os, k, h, zk, zm, cmul(wl, pyr[2 * R + r]));
            else
                pack_out(os, km, h, zm, zk, cmul(wm, pyr[2 * R + (R - 1 - r)]));
        }
    }
    if (q == 0) {
        // column ncols/2 (the slot of column 0's mirror) pairs with itself: (r, R-1-r)
        const float* srow = lds + HALF * SH::RSTRIDE;
        const float2 wh = twiddle_at(pyr, fine, logn, ncols / 2);
        for (int r = tid; r < R / 2; r += WG) {
            const long long k = ncols / 2 + (long long)r * ncols;
            pack_out(os, k, h, make_float2(srow[r], srow[R + r]), make_float2(srow[R - 1 - r], srow[2 * R - 1 - r]),
                     cmul(wh, pyr[2 * R + r]));
        }
    }
}

// ---- host side ----------------------------------------------------------------------

using PassFn = void (*)(const float2*, float2*, const float2*, const float2*, const float2*, int, int, int, long long,
                        long long);

struct PassCfg {
    PassFn fn;
    int wg;
    int cb;
};

template <int LOGR, bool INV, bool FIRST, bool XTW_PASS, bool LAST>
constexpr PassCfg pass_cfg()
{
    // radices above 2^8 have no transposed-twiddle variant
    constexpr bool X = XTW_PASS && LOGR + kLogXtwCols <= kLogMaxPyramid;
    return PassCfg{fft_pass_kernel<LOGR, INV, FIRST, X, LAST>, PassShape<LOGR>::WG, PassShape<LOGR>::CB};
}

template <bool INV, bool FIRST, bool XTW_PASS, bool LAST, int... L>
constexpr void fill_pass(PassCfg* tab, std::integer_sequence<int, L...>)
{
    ((tab[L + kLogMinPass] = pass_cfg<L + kLogMinPass, INV, FIRST, XTW_PASS, LAST>()), ...);
}

// Pass kernels by [inverse][variant][log2 radix]; variant 0 = first pass, 1 = later pass,
// 2 = later pass with twiddles from the transposed copy, 3 and 4 = the same two as the
// last pass (a first pass is never the last).
struct PassTables {
    PassCfg t[2][5][kLogMaxPass + 1];
    PassTables()
    {
        constexpr auto L = std::make_integer_sequence<int, kLogMaxPass - kLogMinPass + 1>{};
        fill_pass<false, true, false, false>(t[0][0], L);
        fill_pass<false, false, false, false>(t[0][1], L);
        fill_pass<false, false, true, false>(t[0][2], L);
        fill_pass<false, false, false, true>(t[0][3], L);
        fill_pass<false, false, true, true>(t[0][4], L);
        fill_pass<true, true, false, false>(t[1][0], L);
        fill_pass<true, false, false, false>(t[1][1], L);
        fill_pass<true, false, true, false>(t[1][2], L);
        fill_pass<true, false, false, true>(t[1][3], L);
        fill_pass<true, false, true, true>(t[1][4], L);
    }
};

static const PassTables& pass_tables()
{
    static const PassTables tables;
    return tables;
}

// The fused edge passes of the real kinds, per log2 radix.
using UnpackFn = void (*)(const float2*, float2*, const float2*, const float2*, int, long long, long long);
using PackFn = void (*)(const float2*, float2*, const float2*, const float2*, const float2*, int, int, long long,
                        long long);

struct EdgePassCfg {
    UnpackFn unpack;
    PackFn pack;       // twiddles from the pyramid / fine tables
    PackFn pack_xtw;   // twiddles from the transposed copy
    int wg;
    int cb;
};

template <int LOGR>
constexpr EdgePassCfg edge_cfg()
{
    constexpr bool X = LOGR + kLogXtwCols <= kLogMaxPyramid;
    return EdgePassCfg{fft_pass_unpack_kernel<LOGR>, fft_pass_pack_kernel<LOGR, false>, fft_pass_pack_kernel<LOGR, X>,
                       PassShape<LOGR>::WG, PassShape<LOGR>::CB};
}

template <int... L>
constexpr void fill_edge(EdgePassCfg* tab, std::integer_sequence<int, L...>)
{
    ((tab[L + kLogMinPass] = edge_cfg<L + kLogMinPass>()), ...);
}

struct EdgePassTables {
    EdgePassCfg t[kLogMaxPass + 1];
    EdgePassTables() { fill_edge(t, std::make_integer_sequence<int, kLogMaxPass - kLogMinPass + 1>{}); }
};

static const EdgePassTables& edge_tables()
{
    static const EdgePassTables tables;
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

// The transposed twiddles cover a later pass of radix 2^logr and span 2^logns when the span
// is their column count and the pass level 2^(logns+logr) is within their table.
static bool xtw_covers(const DeviceTables& tb, int logr, int logns)
{
    return tb.xtw && logns == kLogXtwCols && logns + logr <= tb.log_xtw && logr + kLogXtwCols <= kLogMaxPyramid;
}

static hipError_t launch_pass(int logr, bool inv, bool first, bool last, const float2* in, float2* out,
                              const DeviceTables& tb, int logc, int logns, long long count, hipStream_t s)
{
    const bool xtw = !first && xtw_covers(tb, logr, logns);
    const PassCfg& c = pass_tables().t[inv][first ? 0 : (xtw ? 2 : 1) + (last ? 2 : 0)][logr];
    const long long tiles = ((1ll << logc) >> logr) / c.cb;
    const long long total = tiles * count;
    if (total > 0x7fffffffll) return hipErrorInvalidValue;
    hipLaunchKernelGGL(c.fn, dim3((unsigned)total), dim3(c.wg), 0, s, in, out, tb.pyramid, tb.fine, tb.xtw, tb.log_xtw,
                       logc, logns, tiles, total);
    return hipGetLastError();
}

// Complex transform of `count` signals of 2^logc points through `np` passes, from `in`
// to `out`, using `tmp` (2^logc points per signal) for the intermediate.  Real kinds
// (edge != kComplexForward/Inverse) fuse their unpack into the first pass (kRealInverse:
// `in` holds Y) or their pack into the last (kRealForward: `out` receives the packed
// spectrum, 2^logc + 1 values per signal).
static hipError_t run_passes(Kind edge, bool inv, int logc, long long count, const float2* in, float2* out,
                             float2* tmp, const DeviceTables& tb, hipStream_t s)
{
    int logr[3];
    const int np = plan_passes(logc, logr);
    if (!np) return hipErrorInvalidValue;
    const int logn_real = logc + 1;   // log2 of the real length, for the edge passes
    // alternate buffers so that the last pass lands in `out` (an intermediate may pass
    // through `out` too: it holds at least 2^logc points per signal)
    const float2* src = in;
    int logns = 0;
    for (int p = 0; p < np; ++p) {
        float2* dst = ((np - 1 - p) % 2 == 0) ? out : tmp;
        hipError_t e;
        const bool unpack = edge == kRealInverse && p == 0, pack = edge == kRealForward && p == np - 1;
        if (unpack || pack) {
            const EdgePassCfg& c = edge_tables().t[logr[p]];
            // a mirror tile needs at least one tile of columns (always so above 2^16 points)
            if (((1ll << logc) >> logr[p]) < c.cb) return hipErrorInvalidValue;
            const long long tiles = ((1ll << logc) >> logr[p]) / c.cb;
            const long long total = tiles * count;
            if (total > 0x7fffffffll) return hipErrorInvalidValue;
            if (unpack)
                hipLaunchKernelGGL(c.unpack, dim3((unsigned)total), dim3(c.wg), 0, s, src, dst, tb.pyramid, tb.fine,
                                   logn_real, tiles, total);
            else
                hipLaunchKernelGGL(xtw_covers(tb, logr[p], logns) ? c.pack_xtw : c.pack, dim3((unsigned)total), dim3(c.wg), 0, s, src, dst, tb.pyramid, tb.fine,
                                   tb.xtw, tb.log_xtw, logn_real, tiles, total);
            e = hipGetLastError();
        } else {
            e = launch_pass(logr[p], inv, p == 0, p == np - 1, src, dst, tb, logc, logns, count, s);
        }
        if (e != hipSuccess) return e;
        src = dst;
        logns += logr[p];
    }
    return hipSuccess;
}

size_t multipass_scratch_bytes(Kind kind, int log2n, int count)
{
    const bool real = (kind == kRealForward || kind == kRealInverse);
    const int logc = real ? log2n - 1 : log2n;
    return sizeof(float2) * ((size_t)1 << logc) * (size_t)count;
}

hipError_t launch_multipass(Kind kind, int log2n, int count, const void* in, void* out, const DeviceTables& tb,
                            hipStream_t s, void* scratch, size_t scratch_bytes)
{
    const bool real = (kind == kRealForward || kind == kRealInverse);
    const int logc = real ? log2n - 1 : log2n;
    int logr[3];
    if (!plan_passes(logc, logr)) return hipErrorInvalidValue;
    if (log2n > tb.log_max) return hipErrorInvalidValue;
    const size_t tmp_bytes = multipass_scratch_bytes(kind, log2n, count);
    const bool inv = kind == kComplexInverse || kind == kRealInverse;
    if (scratch) {
        if (scratch_bytes < tmp_bytes) return hipErrorInvalidValue;
        return run_passes(kind, inv, logc, count, (const float2*)in, (float2*)out, (float2*)scratch, tb, s);
    }
    float2* tmp = nullptr;
    hipError_t e = tb.pool ? hipMallocFromPoolAsync((void**)&tmp, tmp_bytes, tb.pool, s)
                           : hipMallocAsync((void**)&tmp, tmp_bytes, s);
    if (e != hipSuccess) return e;
    e = run_passes(kind, inv, logc, count, (const float2*)in, (float2*)out, tmp, tb, s);
    hipError_t e2 = hipFreeAsync(tmp, s);
    return e != hipSuccess ? e : e2;
}

// ---- the fixed scratch of the no-allocation entry point ---------------------------------
//
// CkFftTransformBatchNoAlloc takes its multi-pass intermediates from this array instead of
// allocating them.  Every GPU the library runs on holds its own copy, at its own address,
// from the code object's load until the process ends.  In host code `scratch` names a
// host-side stand-in, never device memory: kernels get the address hipGetSymbolAddress
// returned for the device (g_scratch_ptr).  The intermediate is written in full by one pass
// before the next reads it, so nothing needs clearing between calls; calls that overlap on
// one GPU share the array and race.
static __device__ __align__(256) unsigned char scratch[CKFFT_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxScratchDevices] = {};
}  // namespace

hipError_t noalloc_scratch_setup(int device)
{
    if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return e;
    }
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

void* noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch_ptr[device];
}

}  // namespace ckfft_gpu
