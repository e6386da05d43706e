// fft_device.h -- device-side complex arithmetic and small DFT butterflies (gfx950).
//
// Sign convention: forward transforms use exp(-2 pi i k n / N), inverse exp(+...),
// unnormalised, as in the reference's complex transform
// (src/ckfft/fft_default.cpp:204-256 states the radix-4 form of the same DFT).
// Twiddles come from the device table (FP64-generated, rounded to FP32); the only
// constants here are the radix-8/16 internal rotations, written to FP32 precision.
#pragma once
#include <hip/hip_runtime.h>

namespace ckfft_gpu {

// ---- a transform's own input and output ------------------------------------------------
// A transform reads each input value once and writes each output value once.  With NT
// these accesses carry the non-temporal hint, so that streaming a batch through the
// caches does not displace what is reused (twiddle tables, the intermediate of the next
// pass) and the written results leave without competing with later reads.  Measured from
// HBM (buffer sets beyond the 256 MB Infinity Cache): complex 2^24 170 -> 136 us (reads of
// each pass), real 65536 x 256 forward + inverse step 88 -> 77 us.
typedef float NtF2 __attribute__((ext_vector_type(2)));
typedef float NtF4 __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ __forceinline__ float2 ld_in(const float2* p)
{
    if constexpr (NT) {
        const NtF2 u = __builtin_nontemporal_load(reinterpret_cast<const NtF2*>(p));
        return make_float2(u.x, u.y);
    } else {
        return *p;
    }
}

template <bool NT>
__device__ __forceinline__ void st_out(float2* p, float2 v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(NtF2{v.x, v.y}, reinterpret_cast<NtF2*>(p));
    else
        *p = v;
}

template <bool NT>
__device__ __forceinline__ void st_out4(float4* p, float4 v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(NtF4{v.x, v.y, v.z, v.w}, reinterpret_cast<NtF4*>(p));
    else
        *p = v;
}

// Cache-policy operand of the raw buffer stores: 2 is the non-temporal bit on gfx950.
template <bool NT>
constexpr int kStoreAux = NT ? 2 : 0;

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// a*b with one explicit fused multiply-add per component.  All HIP code is compiled with
// -ffp-contract=off, so these are the only fused operations and every kernel that runs
// the same transform rounds the same way (results do not depend on the kernel).
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// Forward twiddles are stored; the inverse transform uses their conjugates.
template <bool INV>
__device__ __forceinline__ float2 dirw(float2 w) { return INV ? make_float2(w.x, -w.y) : w; }

// a * (-i) for forward, a * (+i) for inverse: the W_4^1 rotation.
template <bool INV>
__device__ __forceinline__ float2 rot4(float2 a)
{
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// a * W_8^1 (forward: exp(-i pi/4)).
template <bool INV>
__device__ __forceinline__ float2 rot8(float2 a)
{
    const float h = 0.70710678118654752f;
    return INV ? make_float2((a.x - a.y) * h, (a.x + a.y) * h)
               : make_float2((a.x + a.y) * h, (a.y - a.x) * h);
}

// a * W_8^3 (forward: exp(-3 i pi/4)).
template <bool INV>
__device__ __forceinline__ float2 rot8_3(float2 a)
{
    const float h = 0.70710678118654752f;
    return INV ? make_float2(-(a.x + a.y) * h, (a.x - a.y) * h)
               : make_float2((a.y - a.x) * h, -(a.x + a.y) * h);
}

// a * W_16^m for the exponents the 4x4 radix-16 needs (m = n2*k1 in {1,2,3,4,6,9}).
template <bool INV, int M>
__device__ __forceinline__ float2 rot16(float2 a)
{
    const float c = 0.92387953251128674f;   // cos(pi/8)
    const float s = 0.38268343236508977f;   // sin(pi/8)
    if constexpr (M == 0) return a;
    else if constexpr (M == 2) return rot8<INV>(a);
    else if constexpr (M == 4) return rot4<INV>(a);
    else if constexpr (M == 6) return rot8_3<INV>(a);
    else {
        // W_16^M = (cos t, -sin t), t = 2 pi M / 16 for forward; conjugate for inverse
        constexpr int m = M & 15;
        float wc, ws;
        if constexpr (m == 1) { wc = c;  ws = s; }
        else if constexpr (m == 3) { wc = s;  ws = c; }
        else if constexpr (m == 9) { wc = -c; ws = -s; }
        else { static_assert(m == 1 || m == 3 || m == 9, "unsupported W16 exponent"); }
        const float wi = INV ? ws : -ws;
        return make_float2(a.x * wc - a.y * wi, a.x * wi + a.y * wc);
    }
}

// exp(-2 pi i q / 64) for q in [0, 64), each component the FP32 rounding of the exact
// value (cos/sin of 2 pi q / 64 evaluated in double precision).  Indexed with
// compile-time constants, so the compiler folds them into immediates.
__device__ constexpr float kW64Re[64] = {1.0f, 0.9951847195625305f, 0.9807852506637573f, 0.9569403529167175f, 0.9238795042037964f, 0.8819212913513184f, 0.8314695954322815f, 0.7730104327201843f, 0.7071067690849304f, 0.6343932747840881f, 0.5555702447891235f, 0.4713967442512512f, 0.3826834261417389f, 0.290284663438797f, 0.19509032368659973f, 0.0980171412229538f, 0.0f, -0.0980171412229538f, -0.19509032368659973f, -0.290284663438797f, -0.3826834261417389f, -0.4713967442512512f, -0.5555702447891235f, -0.6343932747840881f, -0.7071067690849304f, -0.7730104327201843f, -0.8314695954322815f, -0.8819212913513184f, -0.9238795042037964f, -0.9569403529167175f, -0.9807852506637573f, -0.9951847195625305f, -1.0f, -0.9951847195625305f, -0.9807852506637573f, -0.9569403529167175f, -0.9238795042037964f, -0.8819212913513184f, -0.8314695954322815f, -0.7730104327201843f, -0.7071067690849304f, -0.6343932747840881f, -0.5555702447891235f, -0.4713967442512512f, -0.3826834261417389f, -0.290284663438797f, -0.19509032368659973f, -0.0980171412229538f, 0.0f, 0.0980171412229538f, 0.19509032368659973f, 0.290284663438797f, 0.3826834261417389f, 0.4713967442512512f, 0.5555702447891235f, 0.6343932747840881f, 0.7071067690849304f, 0.7730104327201843f, 0.8314695954322815f, 0.8819212913513184f, 0.9238795042037964f, 0.9569403529167175f, 0.9807852506637573f, 0.9951847195625305f};
__device__ constexpr float kW64Im[64] = {0.0f, -0.0980171412229538f, -0.19509032368659973f, -0.290284663438797f, -0.3826834261417389f, -0.4713967442512512f, -0.5555702447891235f, -0.6343932747840881f, -0.7071067690849304f, -0.7730104327201843f, -0.8314695954322815f, -0.8819212913513184f, -0.9238795042037964f, -0.9569403529167175f, -0.9807852506637573f, -0.9951847195625305f, -1.0f, -0.9951847195625305f, -0.9807852506637573f, -0.9569403529167175f, -0.9238795042037964f, -0.8819212913513184f, -0.8314695954322815f, -0.7730104327201843f, -0.7071067690849304f, -0.6343932747840881f, -0.5555702447891235f, -0.4713967442512512f, -0.3826834261417389f, -0.290284663438797f, -0.19509032368659973f, -0.0980171412229538f, 0.0f, 0.0980171412229538f, 0.19509032368659973f, 0.290284663438797f, 0.3826834261417389f, 0.4713967442512512f, 0.5555702447891235f, 0.6343932747840881f, 0.7071067690849304f, 0.7730104327201843f, 0.8314695954322815f, 0.8819212913513184f, 0.9238795042037964f, 0.9569403529167175f, 0.9807852506637573f, 0.9951847195625305f, 1.0f, 0.9951847195625305f, 0.9807852506637573f, 0.9569403529167175f, 0.9238795042037964f, 0.8819212913513184f, 0.8314695954322815f, 0.7730104327201843f, 0.7071067690849304f, 0.6343932747840881f, 0.5555702447891235f, 0.4713967442512512f, 0.3826834261417389f, 0.290284663438797f, 0.19509032368659973f, 0.0980171412229538f};

// exp(-2 pi i q / 64) for the forward direction, its conjugate for the inverse.
template <bool INV = false>
__device__ __forceinline__ float2 w64(int q)
{
    return make_float2(kW64Re[q & 63], INV ? -kW64Im[q & 63] : kW64Im[q & 63]);
}

// exp(-2 pi i q / 128), q in [0, 128), rounded as kW64 (entry 2q equals kW64's entry q).
__device__ constexpr float kW128Re[128] = {1.0f, 0.9987954497337341f, 0.9951847195625305f, 0.9891765117645264f, 0.9807852506637573f, 0.9700312614440918f, 0.9569403529167175f, 0.9415440559387207f, 0.9238795042037964f, 0.903989315032959f, 0.8819212913513184f, 0.8577286005020142f, 0.8314695954322815f, 0.803207516670227f, 0.7730104327201843f, 0.7409511208534241f, 0.7071067690849304f, 0.6715589761734009f, 0.6343932747840881f, 0.5956993103027344f, 0.5555702447891235f, 0.5141027569770813f, 0.4713967442512512f, 0.4275550842285156f, 0.3826834261417389f, 0.3368898630142212f, 0.290284663438797f, 0.24298018217086792f, 0.19509032368659973f, 0.1467304676771164f, 0.0980171412229538f, 0.049067676067352295f, 0.0f, -0.049067676067352295f, -0.0980171412229538f, -0.1467304676771164f, -0.19509032368659973f, -0.24298018217086792f, -0.290284663438797f, -0.3368898630142212f, -0.3826834261417389f, -0.4275550842285156f, -0.4713967442512512f, -0.5141027569770813f, -0.5555702447891235f, -0.5956993103027344f, -0.6343932747840881f, -0.6715589761734009f, -0.7071067690849304f, -0.7409511208534241f, -0.7730104327201843f, -0.803207516670227f, -0.8314695954322815f, -0.8577286005020142f, -0.8819212913513184f, -0.903989315032959f, -0.9238795042037964f, -0.9415440559387207f, -0.9569403529167175f, -0.9700312614440918f, -0.9807852506637573f, -0.9891765117645264f, -0.9951847195625305f, -0.9987954497337341f, -1.0f, -0.9987954497337341f, -0.9951847195625305f, -0.9891765117645264f, -0.9807852506637573f, -0.9700312614440918f, -0.9569403529167175f, -0.9415440559387207f, -0.9238795042037964f, -0.903989315032959f, -0.8819212913513184f, -0.8577286005020142f, -0.8314695954322815f, -0.803207516670227f, -0.7730104327201843f, -0.7409511208534241f, -0.7071067690849304f, -0.6715589761734009f, -0.6343932747840881f, -0.5956993103027344f, -0.5555702447891235f, -0.5141027569770813f, -0.4713967442512512f, -0.4275550842285156f, -0.3826834261417389f, -0.3368898630142212f, -0.290284663438797f, -0.24298018217086792f, -0.19509032368659973f, -0.1467304676771164f, -0.0980171412229538f, -0.049067676067352295f, 0.0f, 0.049067676067352295f, 0.0980171412229538f, 0.1467304676771164f, 0.19509032368659973f, 0.24298018217086792f, 0.290284663438797f, 0.3368898630142212f, 0.3826834261417389f, 0.4275550842285156f, 0.4713967442512512f, 0.5141027569770813f, 0.5555702447891235f, 0.5956993103027344f, 0.6343932747840881f, 0.6715589761734009f, 0.7071067690849304f, 0.7409511208534241f, 0.7730104327201843f, 0.803207516670227f, 0.8314695954322815f, 0.8577286005020142f, 0.8819212913513184f, 0.903989315032959f, 0.9238795042037964f, 0.9415440559387207f, 0.9569403529167175f, 0.9700312614440918f, 0.9807852506637573f, 0.9891765117645264f, 0.9951847195625305f, 0.9987954497337341f};
__device__ constexpr float kW128Im[128] = {0.0f, -0.049067676067352295f, -0.0980171412229538f, -0.1467304676771164f, -0.19509032368659973f, -0.24298018217086792f, -0.290284663438797f, -0.3368898630142212f, -0.3826834261417389f, -0.4275550842285156f, -0.4713967442512512f, -0.5141027569770813f, -0.5555702447891235f, -0.5956993103027344f, -0.6343932747840881f, -0.6715589761734009f, -0.7071067690849304f, -0.7409511208534241f, -0.7730104327201843f, -0.803207516670227f, -0.8314695954322815f, -0.8577286005020142f, -0.8819212913513184f, -0.903989315032959f, -0.9238795042037964f, -0.9415440559387207f, -0.9569403529167175f, -0.9700312614440918f, -0.9807852506637573f, -0.9891765117645264f, -0.9951847195625305f, -0.9987954497337341f, -1.0f, -0.9987954497337341f, -0.9951847195625305f, -0.9891765117645264f, -0.9807852506637573f, -0.9700312614440918f, -0.9569403529167175f, -0.9415440559387207f, -0.9238795042037964f, -0.903989315032959f, -0.8819212913513184f, -0.8577286005020142f, -0.8314695954322815f, -0.803207516670227f, -0.7730104327201843f, -0.7409511208534241f, -0.7071067690849304f, -0.6715589761734009f, -0.6343932747840881f, -0.5956993103027344f, -0.5555702447891235f, -0.5141027569770813f, -0.4713967442512512f, -0.4275550842285156f, -0.3826834261417389f, -0.3368898630142212f, -0.290284663438797f, -0.24298018217086792f, -0.19509032368659973f, -0.1467304676771164f, -0.0980171412229538f, -0.049067676067352295f, 0.0f, 0.049067676067352295f, 0.0980171412229538f, 0.1467304676771164f, 0.19509032368659973f, 0.24298018217086792f, 0.290284663438797f, 0.3368898630142212f, 0.3826834261417389f, 0.4275550842285156f, 0.4713967442512512f, 0.5141027569770813f, 0.5555702447891235f, 0.5956993103027344f, 0.6343932747840881f, 0.6715589761734009f, 0.7071067690849304f, 0.7409511208534241f, 0.7730104327201843f, 0.803207516670227f, 0.8314695954322815f, 0.8577286005020142f, 0.8819212913513184f, 0.903989315032959f, 0.9238795042037964f, 0.9415440559387207f, 0.9569403529167175f, 0.9700312614440918f, 0.9807852506637573f, 0.9891765117645264f, 0.9951847195625305f, 0.9987954497337341f, 1.0f, 0.9987954497337341f, 0.9951847195625305f, 0.9891765117645264f, 0.9807852506637573f, 0.9700312614440918f, 0.9569403529167175f, 0.9415440559387207f, 0.9238795042037964f, 0.903989315032959f, 0.8819212913513184f, 0.8577286005020142f, 0.8314695954322815f, 0.803207516670227f, 0.7730104327201843f, 0.7409511208534241f, 0.7071067690849304f, 0.6715589761734009f, 0.6343932747840881f, 0.5956993103027344f, 0.5555702447891235f, 0.5141027569770813f, 0.4713967442512512f, 0.4275550842285156f, 0.3826834261417389f, 0.3368898630142212f, 0.290284663438797f, 0.24298018217086792f, 0.19509032368659973f, 0.1467304676771164f, 0.0980171412229538f, 0.049067676067352295f};

template <bool INV = false>
__device__ __forceinline__ float2 w128(int q)
{
    return make_float2(kW128Re[q & 127], INV ? -kW128Im[q & 127] : kW128Im[q & 127]);
}

// ---- in-register DFTs on a[0..R-1] (natural order in, natural order out) -------

template <bool INV>
__device__ __forceinline__ void dft2(float2& a0, float2& a1)
{
    float2 t = a0;
    a0 = cadd(t, a1);
    a1 = csub(t, a1);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3)
{
    float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
    float2 t2 = cadd(a1, a3), t3 = rot4<INV>(csub(a1, a3));
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

template <bool INV>
__device__ __forceinline__ void dft8(float2* a)
{
    // even/odd split: E = DFT4(a0,a2,a4,a6), O = DFT4(a1,a3,a5,a7)
    dft4<INV>(a[0], a[2], a[4], a[6]);
    dft4<INV>(a[1], a[3], a[5], a[7]);
    float2 o1 = rot8<INV>(a[3]), o2 = rot4<INV>(a[5]), o3 = rot8_3<INV>(a[7]);
    float2 e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6], o0 = a[1];
    a[0] = cadd(e0, o0); a[4] = csub(e0, o0);
    a[1] = cadd(e1, o1); a[5] = csub(e1, o1);
    a[2] = cadd(e2, o2); a[6] = csub(e2, o2);
    a[3] = cadd(e3, o3); a[7] = csub(e3, o3);
}

template <bool INV>
__device__ __forceinline__ void dft16(float2* a)
{
    // n = n2 + 4 n1, k = k1 + 4 k2: column DFT4 over n1, twiddle W16^(n2 k1),
    // row DFT4 over n2.
    dft4<INV>(a[0], a[4], a[8], a[12]);
    dft4<INV>(a[1], a[5], a[9], a[13]);
    dft4<INV>(a[2], a[6], a[10], a[14]);
    dft4<INV>(a[3], a[7], a[11], a[15]);
    // b[n2][k1] sits in a[n2 + 4 k1]
    a[5] = rot16<INV, 1>(a[5]);
    a[9] = rot16<INV, 2>(a[9]);
    a[13] = rot16<INV, 3>(a[13]);
    a[6] = rot16<INV, 2>(a[6]);
    a[10] = rot16<INV, 4>(a[10]);
    a[14] = rot16<INV, 6>(a[14]);
    a[7] = rot16<INV, 3>(a[7]);
    a[11] = rot16<INV, 6>(a[11]);
    a[15] = rot16<INV, 9>(a[15]);
    // row DFT4 over n2 for each k1: inputs a[4 k1 + n2], outputs X[k1 + 4 k2]
    dft4<INV>(a[0], a[1], a[2], a[3]);
    dft4<INV>(a[4], a[5], a[6], a[7]);
    dft4<INV>(a[8], a[9], a[10], a[11]);
    dft4<INV>(a[12], a[13], a[14], a[15]);
    // now a[4 k1 + k2] = X[k1 + 4 k2]: transpose the 4x4 index to natural order
    float2 t;
    t = a[1]; a[1] = a[4]; a[4] = t;
    t = a[2]; a[2] = a[8]; a[8] = t;
    t = a[3]; a[3] = a[12]; a[12] = t;
    t = a[6]; a[6] = a[9]; a[9] = t;
    t = a[7]; a[7] = a[13]; a[13] = t;
    t = a[11]; a[11] = a[14]; a[14] = t;
}

// a * W_32^M (forward: exp(-2 pi i M / 32)), M in [0, 16): the rotations of the radix-2
// combine of dft32.  The exponents that are multiples of 2 are W_16 rotations.
template <bool INV, int M>
__device__ __forceinline__ float2 rot32(float2 a)
{
    if constexpr (M % 2 == 1)
        return cmul(a, w128<INV>(4 * M));
    else if constexpr (M / 2 == 5)              // W16^5 = W16^4 W16^1
        return rot4<INV>(rot16<INV, 1>(a));
    else if constexpr (M / 2 == 7)              // W16^7 = W16^4 W16^3
        return rot4<INV>(rot16<INV, 3>(a));
    else
        return rot16<INV, M / 2>(a);
}

// 32-point DFT as a radix-2 step over two 16-point DFTs: X[k] = E[k] + W32^k O[k],
// X[k+16] = E[k] - W32^k O[k], E/O the DFTs of the even/odd inputs.
template <bool INV>
__device__ __forceinline__ void dft32(float2* a)
{
    float2 e[16], o[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        e[i] = a[2 * i];
        o[i] = a[2 * i + 1];
    }
    dft16<INV>(e);
    dft16<INV>(o);
    o[1] = rot32<INV, 1>(o[1]);
    o[2] = rot32<INV, 2>(o[2]);
    o[3] = rot32<INV, 3>(o[3]);
    o[4] = rot32<INV, 4>(o[4]);
    o[5] = rot32<INV, 5>(o[5]);
    o[6] = rot32<INV, 6>(o[6]);
    o[7] = rot32<INV, 7>(o[7]);
    o[8] = rot4<INV>(o[8]);
    o[9] = rot32<INV, 9>(o[9]);
    o[10] = rot32<INV, 10>(o[10]);
    o[11] = rot32<INV, 11>(o[11]);
    o[12] = rot32<INV, 12>(o[12]);
    o[13] = rot32<INV, 13>(o[13]);
    o[14] = rot32<INV, 14>(o[14]);
    o[15] = rot32<INV, 15>(o[15]);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        a[k] = cadd(e[k], o[k]);
        a[k + 16] = csub(e[k], o[k]);
    }
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2* a)
{
    if constexpr (R == 1) {
    } else if constexpr (R == 2) {
        dft2<INV>(a[0], a[1]);
    } else if constexpr (R == 4) {
        dft4<INV>(a[0], a[1], a[2], a[3]);
    } else if constexpr (R == 8) {
        dft8<INV>(a);
    } else if constexpr (R == 16) {
        dft16<INV>(a);
    } else {
        static_assert(R == 32, "radix must be 1, 2, 4, 8, 16 or 32");
        dft32<INV>(a);
    }
}

}  // namespace ckfft_gpu
