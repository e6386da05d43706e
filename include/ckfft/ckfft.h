/*
 * ckfft.h -- drop-in C API of the MI355X transform library.
 *
 * Same six entry points, types, enum values, argument meanings and 0/1/NULL error
 * conventions as the reference's public header, so existing callers recompile or
 * relink unchanged.  Each declaration names the reference declaration it replaces
 * (inc/ckfft/ckfft.h in the reference checkout).
 *
 * Transforms are 1-D, power-of-two sizes, FP32, out-of-place and unnormalised:
 *   complex forward  X[k] = sum_n x[n] exp(-2 pi i k n / N)
 *   complex inverse  uses exp(+2 pi i k n / N); forward then inverse gives N*x
 *   real forward     N real samples -> N/2+1 complex values, equal to 2*DFT(x)[0..N/2]
 *   real inverse     N/2+1 complex values -> N real samples; forward then inverse
 *                    gives 2N*x
 *
 * Buffers may be host memory (as with the reference) or device memory from
 * hipMalloc/torch; the library tells them apart per call.  Every call is synchronous:
 * results are in `output` when it returns.  The work runs on the GPU.
 *
 * Thread safety: a context holds no per-call state and may be used from several
 * threads at once.
 */
#ifndef CKFFT_CKFFT_H
#define CKFFT_CKFFT_H

#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Interleaved complex sample.  Replaces inc/ckfft/ckfft.h:10-15. */
typedef struct
{
    float real;
    float imag;
}
CkFftComplex;

/* Opaque context.  Replaces inc/ckfft/ckfft.h:18. */
typedef struct _CkFftContext CkFftContext;

/* Which transform directions a context supports.  Replaces inc/ckfft/ckfft.h:21-27. */
typedef enum
{
    kCkFftDirection_Forward = (1 << 0),
    kCkFftDirection_Inverse = (1 << 1),
    kCkFftDirection_Both    = (kCkFftDirection_Forward | kCkFftDirection_Inverse)
}
CkFftDirection;

/* Create a context for transforms of size up to nMax (a power of two).
 * Replaces inc/ckfft/ckfft.h:59 (CkFftInit).
 *
 * buf/bufSize: optional caller memory for the host part of the context.  If bufSize
 * is given and buf is NULL or *bufSize is too small, the required byte count is
 * written to *bufSize and NULL is returned.  buf without bufSize is an error.
 * The device-resident twiddle tables are always allocated by the library (device
 * memory cannot live in a caller's host buffer) and are released by CkFftShutdown.
 *
 * Returns NULL if nMax <= 0 or not a power of two, if direction is not one of the
 * three values above, on a query/too-small buffer, or if no GPU is usable. */
CkFftContext* CkFftInit(int nMax, CkFftDirection direction, void* buf, size_t* bufSize);

/* Real forward transform: n floats in, n/2+1 complex out.
 * Replaces inc/ckfft/ckfft.h:85 (CkFftRealForward).
 * Returns 0 if the context is NULL or lacks the forward direction, n is not a power
 * of two or exceeds nMax, a pointer is NULL, or input == output; otherwise 1. */
int CkFftRealForward(CkFftContext* context, int n, const float* input, CkFftComplex* output);

/* Real inverse transform: n/2+1 complex in, n floats out.  tmpBuf must be non-NULL
 * (n/2+1 elements, as in the reference); the GPU path does not write to it.
 * Replaces inc/ckfft/ckfft.h:108 (CkFftRealInverse). */
int CkFftRealInverse(CkFftContext* context, int n, const CkFftComplex* input, float* output, CkFftComplex* tmpBuf);

/* Complex forward transform, n in, n out.  Replaces inc/ckfft/ckfft.h:129. */
int CkFftComplexForward(CkFftContext* context, int n, const CkFftComplex* input, CkFftComplex* output);

/* Complex inverse transform, n in, n out.  Replaces inc/ckfft/ckfft.h:150. */
int CkFftComplexInverse(CkFftContext* context, int n, const CkFftComplex* input, CkFftComplex* output);

/* Destroy a context; NULL is ignored.  Frees the host part only if the library
 * allocated it, and always frees the device tables.  Replaces inc/ckfft/ckfft.h:158. */
void CkFftShutdown(CkFftContext* context);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* CKFFT_CKFFT_H */
