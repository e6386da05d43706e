/*
 * ckfft_oracle.h -- CPU restatement of ckfft's scalar ("default") transform path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the GPU product is
 * compared against; it is never linked into, loaded by or called from the product
 * library (ckfft_amd/libckfft_amd.so).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may use it.
 *
 * Arithmetic: IEEE FP32, separate multiply and add (built with -ffp-contract=off),
 * same operation order as the reference, so that results are bit-identical with the
 * reference built from its own sources on the same libm.
 *
 * Functions and the reference code each one restates (paths relative to the
 * reference checkout):
 *   ocf_create          src/ckfft/context.cpp:24-114   (table layout, twiddles 90-105)
 *   ocf_complex         src/ckfft/fft.cpp:13-46, src/ckfft/fft_default.cpp:12-266
 *   ocf_real_forward    src/ckfft/fft_real.cpp:13-60, src/ckfft/fft_real_default.cpp:13-63
 *   ocf_real_inverse    src/ckfft/fft_real.cpp:62-105, src/ckfft/fft_real_default.cpp:65-114
 *   ocf_check_*         src/ckfft/ckfft.cpp:14-114 (argument validation, 0/1 returns)
 *
 * Pinned by (tests/test_oracle.py): bit-exact agreement with golden vectors produced
 * by the genuine reference (tests/golden/), bit-exact agreement with oracle/_ref when
 * it is built, an FP64 brute-force DFT for N <= 1024, closed forms for N <= 4, and the
 * round-trip identities (complex: n*x, real: 2n*x).
 */
#ifndef CKFFT_ORACLE_H
#define CKFFT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float re, im; } ocf_cplx;

typedef struct {
    int max_n;          /* the context's maxCount */
    ocf_cplx* fwd;      /* fwd[i] = (cos t, sin t), t = (-2*pi_f*i)/maxCount in float */
    ocf_cplx* inv;      /* inv[i] = (cos t, -sin t) */
} ocf_ctx;

/* dirs: 1 forward table, 2 inverse table, 3 both.  Returns NULL on bad arguments. */
ocf_ctx* ocf_create(int max_n, int dirs);
void     ocf_destroy(ocf_ctx* c);

/* Twiddle table as the reference's context builds it (for table-level checks). */
void ocf_twiddles(int max_n, ocf_cplx* fwd, ocf_cplx* inv);

/* All transforms return 1 on success, 0 on invalid arguments (same rules as the
 * reference API).  out != in is required.  Outputs are unnormalised.
 * ocf_real_forward writes exactly n/2+1 values (the reference also writes out[3]
 * at n == 4; the oracle does not). */
int ocf_complex(const ocf_ctx* c, int n, const ocf_cplx* in, ocf_cplx* out, int inverse);
int ocf_real_forward(const ocf_ctx* c, int n, const float* in, ocf_cplx* out);
int ocf_real_inverse(const ocf_ctx* c, int n, const ocf_cplx* in, float* out, ocf_cplx* tmp);

/* Batched drivers used by the CPU baseline: `count` independent transforms laid out
 * back to back, split over `threads` POSIX threads that share the context. */
int ocf_complex_batch(const ocf_ctx* c, int n, int count, const ocf_cplx* in,
                      ocf_cplx* out, int inverse, int threads);
int ocf_real_forward_batch(const ocf_ctx* c, int n, int count, const float* in,
                           ocf_cplx* out, int threads);
int ocf_real_inverse_batch(const ocf_ctx* c, int n, int count, const ocf_cplx* in,
                           float* out, int threads);

#ifdef __cplusplus
}
#endif
#endif
