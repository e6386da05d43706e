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

/* ---- a batched call that allocates nothing ------------------------------------------
 *
 * No counterpart in the reference.  For callers that cannot allocate device memory: code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.  It is
 * the batched call above for any kind (kCkFftKind_*, below), with the same arguments,
 * validation and results (bit for bit), except that the sizes above one workgroup
 * (complex n > 2^15, real n > 2^16) take their intermediate from a fixed device array the
 * library holds on every GPU it runs on, instead of allocating it.  That array is
 * CKFFT_NOALLOC_SCRATCH_BYTES (64 MiB) per GPU, held until the process ends; a transform
 * whose intermediate (8 n bytes per complex signal, 4 n per real signal, times count)
 * exceeds it is rejected (0, and CkFftGpuLastError says why).  The call allocates nothing
 * on any path.  Calls that overlap on one GPU share the array and race: serialise them
 * (one stream, or events).  Returns 1 when the work was enqueued, 0 otherwise. */
#define CKFFT_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int CkFftTransformBatchNoAlloc(CkFftContext* context, int kind, int n, int count, const void* input, void* output,
                               void* stream);

/* ---- uneven batches in persistent launches --------------------------------------------
 *
 * A batch plan takes any mix of transforms -- each with its own kind, size and device
 * buffers -- and runs them in persistent kernels (workgroups pull work items from an
 * on-device queue, largest first): one launch for the classes that need a whole CU's
 * LDS, one with two workgroups per CU for the rest.  Complex 2^16 transforms run as two
 * table-driven pass launches over all of them, both directions (every workgroup of the
 * GPU on one transform's tiles at a time), and transforms the persistent kernels do not take
 * (complex n <= 1, real n <= 4, any n above 2^16) run as ordinary launches; all on the same
 * stream after the persistent launches.  CkFftBatchInfo reports how many transforms each
 * route takes.
 * Results are bit-identical to the single-signal calls for every transform.
 * Executions of one plan must not overlap (they share its queue).  The reference has
 * no counterpart for these entry points. */

enum {
    kCkFftKind_ComplexForward = 0,
    kCkFftKind_ComplexInverse = 1,
    kCkFftKind_RealForward = 2,
    kCkFftKind_RealInverse = 3
};

typedef struct {
    int kind;             /* kCkFftKind_* */
    int n;                /* transform size, a power of two <= the context's nMax */
    const void* input;    /* device memory, layout of the single-signal call */
    void* output;         /* device memory */
} CkFftBatchDesc;

typedef struct {
    int transforms;             /* descriptors in the plan */
    int persistent_transforms;  /* handled by the persistent launches */
    int work_items;             /* queue entries */
    int workgroups;             /* persistent workgroups, both persistent launches */
    int extra_launches;         /* transforms run as ordinary launches */
    long long algorithmic_bytes;/* one read of every input, one write of every output */
    int pass_transforms;        /* complex 2^16 run by the table-driven pass launches */
} CkFftBatchInfo;

typedef struct _CkFftBatchPlan CkFftBatchPlan;

/* Flag: capture the plan's launches (the persistent kernels and the table-driven passes)
 * in a hipGraph once and replay it.  With or without it, an execution runs the table-driven
 * passes, the two-per-CU persistent launch and the one-per-CU persistent launch as three
 * branches side by side, on streams the plan owns, forked from and joined back into the
 * caller's stream. */
#define kCkFftBatch_Graph 1

/* Validate the descriptors (same rules as the single-signal calls, plus device
 * pointers) and build the queue.  NULL on invalid arguments or allocation failure
 * (CkFftBatchLastError tells which). */
CkFftBatchPlan* CkFftBatchPlanCreate(CkFftContext* context, const CkFftBatchDesc* descs, int count, int flags);
/* Enqueue one execution on `stream` (a hipStream_t; NULL = null stream): work after it on
 * `stream` sees the outputs.  Executions of one plan must not overlap.  1/0. */
int CkFftBatchPlanExecute(CkFftBatchPlan* plan, void* stream);
int CkFftBatchPlanGetInfo(const CkFftBatchPlan* plan, CkFftBatchInfo* info);
void CkFftBatchPlanDestroy(CkFftBatchPlan* plan);
const char* CkFftBatchLastError(void);

/* Choose each class's layout in batch plans on this GPU.  A class is (kind, n); its layouts
 * are the points per thread (16 or 32, for 32 <= n <= 8192 complex points -- which sets the
 * waves per transform and the transforms per workgroup), the persistent kernel that runs it
 * (one workgroup per CU with a 136 KiB LDS tile, or two per CU with half that, for the
 * classes that fit; 16384-point transforms then exchange one LDS plane at a time) and, for
 * complex n = 65536, table-driven passes over all such transforms or an ordinary two-pass
 * launch each.  The radix plan is fixed per size: it sets the arithmetic, and layouts differ
 * only in speed, never in results.  CkFftAutotune times every layout of every class (complex
 * 32 .. 65536 points, real 64 .. 65536 samples) in a plan of that class alone and keeps the
 * fastest in the context (used by later plans); if cachePath is not NULL it writes the
 * choices and times as JSON. */
int CkFftAutotune(CkFftContext* context, const char* cachePath);
/* Choose the layouts inside a given mix: starting from the faster of the context's choices
 * and the defaults, for each class of descs (largest share of the bytes first, two sweeps)
 * time the whole plan (graph replays) with each of the class's other layouts against the
 * current layouts, in alternating rounds, and keep the best that is more than 1% faster.
 * The result is timed the same way against the layouts on entry and kept only if faster.
 * usBefore / usAfter (may be NULL) receive the plan's microseconds per execution with the
 * layouts on entry and with those kept (never more).  1/0; on 0 the entry layouts stay. */
int CkFftAutotuneMix(CkFftContext* context, const CkFftBatchDesc* descs, int count, const char* cachePath,
                     double* usBefore, double* usAfter);
/* Load choices written by CkFftAutotune or CkFftAutotuneMix.  1 if any entry was applied. */
int CkFftLoadTuning(CkFftContext* context, const char* cachePath);

/* ---- serving loop ------------------------------------------------------------------
 *
 * A server admits transforms as they arrive, from any thread, and runs them with
 * continuous batching: a dispatcher thread takes every admitted request into the next
 * free slot (up to maxInFlight batches on the device at once, each one persistent
 * launch on the slot's own stream; transforms above one workgroup's size -- complex
 * n >= 2^16, real n >= 2^17 -- run as multi-pass launches on that stream), so requests
 * arriving while others run join the next batch and their work overlaps on the device.
 * Admission is first-come-first-served against an HBM budget: a request is admitted
 * while the bytes of admitted, unfinished requests (one read of the input, one write of
 * the output) plus the server's own resident buffers stay within hbmBudget; others wait
 * (back-pressure, not failure).  Each slot holds a resident intermediate for the
 * multi-pass sizes (8n bytes complex, 4n real, up to 64 MiB: the context's nMax at most);
 * a request that needs more has its intermediate charged with it.
 * Twiddles, descriptor buffers, intermediates and streams stay resident.
 * A request's result is bit-identical to the single-signal call on the same input,
 * whatever it was batched with.  Buffers are device memory owned by the caller and
 * must stay valid until the request completes.  The reference has no counterpart. */

typedef struct _CkFftServer CkFftServer;

typedef struct {
    long long completed;          /* requests finished since the last reset */
    long long failed;
    double p50_us, p99_us, max_us, mean_us;  /* latency, arrival to completion seen, over the
                                              last 2^20 completions since the last reset */
    double requests_per_s;        /* completed / (last completion - first arrival) */
    double GB_per_s;              /* algorithmic bytes over the same span */
    long long peak_inflight_bytes;/* largest admitted-unfinished byte count */
    long long budget_bytes;       /* hbmBudget minus the resident bytes */
    long long resident_bytes;     /* the server's own device buffers */
    long long batches;            /* batches launched */
    double mean_batch;            /* requests per batch */
    double max_client_us;         /* the slowest request (max_us): arrival to its submission */
    double max_wait_us;           /* its submission to its batch's launch */
    double max_device_us;         /* its batch's launch to completion seen */
} CkFftServerStats;

/* NULL on invalid arguments, a budget below the server's resident buffers, or
 * allocation failure (CkFftServerLastError tells which).  maxInFlight in [1, 64]. */
CkFftServer* CkFftServerCreate(CkFftContext* context, size_t hbmBudget, int maxInFlight);
/* Submit one transform (kind, n and buffers as in CkFftBatchDesc).  Returns a ticket
 * >= 0, or -1 if the request is invalid or alone exceeds the budget.  The server tracks
 * the last 2^18 tickets; a submission waits while the ticket 2^18 before it is still
 * unfinished. */
long long CkFftServerSubmit(CkFftServer* server, int kind, int n, const void* input, void* output);
/* Submit descs[i] at arrivalUs[i] microseconds after the call (non-decreasing times),
 * from the calling thread, writing tickets[i] (may be NULL).  Latency is measured from
 * the scheduled arrival.  1 if every request was accepted. */
int CkFftServerReplay(CkFftServer* server, const CkFftBatchDesc* descs, const double* arrivalUs, int count,
                      long long* tickets);
/* 1 done, 0 pending, -1 failed or unknown ticket.  A ticket more than 2^18 submissions
 * old has finished: it reports -1 if it failed, else 1 (failed tickets are remembered). */
int CkFftServerQuery(CkFftServer* server, long long ticket);
/* Block until the ticket completes; 1 on success. */
int CkFftServerWait(CkFftServer* server, long long ticket);
/* Block until every submitted request completes; 1 if none failed. */
int CkFftServerDrain(CkFftServer* server);
int CkFftServerGetStats(CkFftServer* server, CkFftServerStats* stats);
void CkFftServerResetStats(CkFftServer* server);
/* Finishes outstanding requests, then frees the server. */
void CkFftServerDestroy(CkFftServer* server);
const char* CkFftServerLastError(void);

/* ---- precision mode ------------------------------------------------------------------
 *
 * kCkFftPrecision_Fp32 (default): every stage in FP32 on the vector units.
 * kCkFftPrecision_SplitBf16x3 (opt-in): complex transforms with 16 <= n <= 65536 run
 * their radix-16 stages as 32x32 DFT-matrix products on the matrix cores, each FP32
 * operand split exactly into three BF16 parts, six partial products accumulated in FP32
 * (error at the FP32 level; results differ in the last bits from the FP32 mode).  Real
 * transforms and other sizes keep the FP32 kernels.  The mode applies to every later call
 * on the context: single-signal, batched, batch plans created afterwards (those
 * transforms then run as ordinary launches) and servers.  The reference has no
 * counterpart.  Returns 1, or 0 for an invalid context or mode. */
enum {
    kCkFftPrecision_Fp32 = 0,
    kCkFftPrecision_SplitBf16x3 = 1
};
int CkFftSetPrecision(CkFftContext* context, int precision);
/* The context's mode, or -1 for an invalid context. */
int CkFftGetPrecision(const CkFftContext* context);

/* ---- device memory of a context ------------------------------------------------------
 *
 * CkFftInit's query mode (buf == NULL, bufSize != NULL) reports the host part of a context.
 * Its device part -- the twiddle tables, generated in FP64 on the GPU at init -- is one
 * block whose size CkFftDeviceTableBytes reports (0 for invalid arguments).  CkFftInitArena
 * is CkFftInit with that block placed in caller-provided device memory (`deviceArena`, at
 * least CkFftDeviceTableBytes bytes, 256-byte aligned); the arena stays the caller's and
 * CkFftShutdown does not free it (it may be reused once the context is shut down).  Query
 * mode works as in CkFftInit.  Host-pointer calls, batch plans and servers still allocate
 * their own buffers.  NULL on invalid arguments or a too small / misaligned / non-device
 * arena (CkFftGpuLastError tells which).  The reference has no counterpart (its context
 * holds host tables only: src/ckfft/context.cpp:26-67). */
size_t CkFftDeviceTableBytes(int nMax, CkFftDirection direction);
CkFftContext* CkFftInitArena(int nMax, CkFftDirection direction, void* buf, size_t* bufSize, void* deviceArena,
                             size_t deviceArenaBytes);

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
