/*
 * ckfft_gpu.h -- batched, stream-ordered extensions of the drop-in API.
 *
 * The reference has one transform per call on host memory and no counterpart for
 * these entry points.  They are for callers that keep signals resident in device
 * memory and want one launch per batch instead of one synchronous call per signal.
 * Arithmetic and results are exactly those of the single-signal calls in ckfft.h:
 * the same kernels compute every signal, so a signal's output does not depend on
 * the batch it is in.
 *
 * Layout: `count` signals back to back.
 *   complex:       in/out stride n complex values
 *   real forward:  in stride n floats,         out stride n/2+1 complex values
 *   real inverse:  in stride n/2+1 complex,    out stride n floats
 * Pointers must be device memory.  `stream` is a hipStream_t (NULL = the null
 * stream); the call only enqueues work and returns, so the caller synchronises.
 * Return 1 when the work was enqueued, 0 on invalid arguments (same rules as the
 * single-signal calls, plus count >= 0 and device pointers).
 */
#ifndef CKFFT_CKFFT_GPU_H
#define CKFFT_CKFFT_GPU_H

#include "ckfft/ckfft.h"

#ifdef __cplusplus
extern "C" {
#endif

int CkFftComplexForwardBatch(CkFftContext* context, int n, int count,
                             const CkFftComplex* input, CkFftComplex* output, void* stream);
int CkFftComplexInverseBatch(CkFftContext* context, int n, int count,
                             const CkFftComplex* input, CkFftComplex* output, void* stream);
int CkFftRealForwardBatch(CkFftContext* context, int n, int count,
                          const float* input, CkFftComplex* output, void* stream);
int CkFftRealInverseBatch(CkFftContext* context, int n, int count,
                          const CkFftComplex* input, float* output, void* stream);

/* Copy the device twiddle table of level L (L values exp(-2 pi i k / L), k < L,
 * generated in FP64 on the GPU and rounded to FP32) to host memory `out`.
 * L must be a power of two no larger than the context's table.  Returns 1/0. */
int CkFftGetTwiddles(CkFftContext* context, int L, CkFftComplex* out);

/* Description of the last GPU-side failure on the calling thread (a HIP error and
 * the step that returned it), or "no error".  Valid until the thread's next call. */
const char* CkFftGpuLastError(void);

/* Library version and build information as a static string. */
const char* CkFftGpuVersion(void);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* CKFFT_CKFFT_GPU_H */
