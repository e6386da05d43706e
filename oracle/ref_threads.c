/* ref_threads.c -- a POSIX-thread driver for the genuine reference library (test and
 * benchmark infrastructure only; nothing here is part of the product).
 *
 * The CPU baseline times the reference's own transform path on the host's cores.  The
 * reference is single-threaded per call and allows one context to be used from several
 * threads at once (inc/ckfft/ckfft.h:39-41), so a batch is split over threads that share
 * one context.  Driving the calls from C keeps the per-call overhead of a Python thread
 * loop (and its interpreter lock) out of the timing, which matters for the short
 * transforms (a 4096-point call takes tens of microseconds).
 *
 * Built by oracle/Makefile (`make ref`) into _ref/, linked against _ref/libckfft_ref.so
 * (the reference built from its own sources).  Only the reference's six C entry points
 * are called (their prototypes below restate inc/ckfft/ckfft.h:59-158).
 */
#include <pthread.h>
#include <stddef.h>
#include <stdlib.h>

typedef struct { float real, imag; } rt_complex;
typedef struct _CkFftContext rt_context;

int CkFftRealForward(rt_context* c, int n, const float* in, rt_complex* out);
int CkFftRealInverse(rt_context* c, int n, const rt_complex* in, float* out, rt_complex* tmp);
int CkFftComplexForward(rt_context* c, int n, const rt_complex* in, rt_complex* out);
int CkFftComplexInverse(rt_context* c, int n, const rt_complex* in, rt_complex* out);

typedef struct {
    rt_context* ctx;
    int count, passes, threads, index, max_n, failed;
    const int* kinds;           /* 0 complex forward, 1 complex inverse, 2 real forward, 3 real inverse */
    const int* ns;
    const void* const* ins;
    void* const* outs;
} rt_job;

static void* rt_worker(void* arg)
{
    rt_job* j = (rt_job*)arg;
    rt_complex* tmp = (rt_complex*)malloc(sizeof(rt_complex) * ((size_t)j->max_n / 2 + 1));
    if (!tmp) {
        j->failed = 1;
        return NULL;
    }
    for (int p = 0; p < j->passes; ++p)
        for (int i = j->index; i < j->count; i += j->threads) {
            const int n = j->ns[i];
            int ok;
            switch (j->kinds[i]) {
            case 0: ok = CkFftComplexForward(j->ctx, n, (const rt_complex*)j->ins[i], (rt_complex*)j->outs[i]); break;
            case 1: ok = CkFftComplexInverse(j->ctx, n, (const rt_complex*)j->ins[i], (rt_complex*)j->outs[i]); break;
            case 2: ok = CkFftRealForward(j->ctx, n, (const float*)j->ins[i], (rt_complex*)j->outs[i]); break;
            default: ok = CkFftRealInverse(j->ctx, n, (const rt_complex*)j->ins[i], (float*)j->outs[i], tmp); break;
            }
            if (!ok) j->failed = 1;
        }
    free(tmp);
    return NULL;
}

/* `passes` passes over `count` transforms (kinds[i], ns[i], ins[i] -> outs[i]), transform i
 * run by thread i % threads, all threads sharing `ctx`.  Returns 1, or 0 if a call was
 * refused or a thread could not start. */
int refthr_run(rt_context* ctx, int count, const int* kinds, const int* ns, const void* const* ins,
               void* const* outs, int threads, int passes)
{
    if (!ctx || count < 0 || threads < 1 || passes < 0) return 0;
    int max_n = 1;
    for (int i = 0; i < count; ++i)
        if (ns[i] > max_n) max_n = ns[i];
    rt_job* jobs = (rt_job*)calloc((size_t)threads, sizeof(rt_job));
    pthread_t* tid = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
    int ok = jobs && tid;
    int started = 0;
    for (int t = 0; ok && t < threads; ++t) {
        jobs[t] = (rt_job){ctx, count, passes, threads, t, max_n, 0, kinds, ns, ins, outs};
        if (threads == 1) {
            rt_worker(&jobs[0]);
        } else if (pthread_create(&tid[t], NULL, rt_worker, &jobs[t]) == 0) {
            ++started;
        } else {
            ok = 0;
        }
    }
    for (int t = 0; t < started; ++t) pthread_join(tid[t], NULL);
    for (int t = 0; ok && t < threads; ++t)
        if (jobs[t].failed) ok = 0;
    free(jobs);
    free(tid);
    return ok;
}
