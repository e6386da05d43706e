/*
 * ckfft_oracle.c -- CPU restatement of ckfft's scalar transform path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See ckfft_oracle.h for
 * the list of reference functions restated here and for what pins each of them.
 *
 * Build: gcc -O2 -ffp-contract=off (no FMA contraction, no -ffast-math), so every
 * complex multiply is (a.re*b.re - a.im*b.im, a.im*b.re + a.re*b.im) rounded after
 * each product and each sum, as in src/ckfft/math_util.h:29-33.
 */
#define _GNU_SOURCE
#include "ckfft_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---- FP32 complex helpers: src/ckfft/math_util.h:17-33 ---------------------- */

static inline ocf_cplx cx(float re, float im) { ocf_cplx z; z.re = re; z.im = im; return z; }
static inline ocf_cplx cx_add(ocf_cplx a, ocf_cplx b) { return cx(a.re + b.re, a.im + b.im); }
static inline ocf_cplx cx_sub(ocf_cplx a, ocf_cplx b) { return cx(a.re - b.re, a.im - b.im); }
/* a*b with the reference's term order: re = ar*br - ai*bi, im = ai*br + ar*bi */
static inline ocf_cplx cx_mul(ocf_cplx a, ocf_cplx b)
{
    return cx(a.re * b.re - a.im * b.im, a.im * b.re + a.re * b.im);
}

/* src/ckfft/math_util.h:12-15: the test is on the unsigned value.  The oracle also
 * rejects n <= 0, which the reference would let through for n == INT_MIN. */
static int pow2_positive(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* ---- twiddles: src/ckfft/context.cpp:90-105 -------------------------------- */

void ocf_twiddles(int max_n, ocf_cplx* fwd, ocf_cplx* inv)
{
    const float two_pi_neg = -2.0f * (float)M_PI;   /* exact: scaling by 2 */
    for (int k = 0; k < max_n; ++k) {
        /* angle formed in float exactly as the reference does: ((-2*pi_f)*k)/maxCount */
        float ang = (two_pi_neg * (float)k) / (float)max_n;
        float c = cosf(ang);
        float s = sinf(ang);
        if (fwd) fwd[k] = cx(c, s);
        if (inv) inv[k] = cx(c, -s);
    }
}

ocf_ctx* ocf_create(int max_n, int dirs)
{
    if (!pow2_positive(max_n)) return NULL;
    if (dirs != 1 && dirs != 2 && dirs != 3) return NULL;
    ocf_ctx* c = (ocf_ctx*)calloc(1, sizeof(ocf_ctx));
    if (!c) return NULL;
    c->max_n = max_n;
    if (dirs & 1) c->fwd = (ocf_cplx*)malloc(sizeof(ocf_cplx) * (size_t)max_n);
    if (dirs & 2) c->inv = (ocf_cplx*)malloc(sizeof(ocf_cplx) * (size_t)max_n);
    if (((dirs & 1) && !c->fwd) || ((dirs & 2) && !c->inv)) { ocf_destroy(c); return NULL; }
    ocf_twiddles(max_n, c->fwd, c->inv);
    return c;
}

void ocf_destroy(ocf_ctx* c)
{
    if (!c) return;
    free(c->fwd);
    free(c->inv);
    free(c);
}

/* ---- radix-4 DIT recursion: src/ckfft/fft_default.cpp:12-266 ---------------- */

/* Combine F0 with three quarter-results that are already multiplied by their
 * twiddles.  Forward uses -i for the odd outputs, inverse +i
 * (src/ckfft/fft_default.cpp:226-256). */
static inline void radix4_combine(ocf_cplx f0, ocf_cplx g1, ocf_cplx g2, ocf_cplx g3,
                                  int inverse, ocf_cplx* y0, ocf_cplx* y1,
                                  ocf_cplx* y2, ocf_cplx* y3)
{
    ocf_cplx s02 = cx_add(f0, g2), d02 = cx_sub(f0, g2);
    ocf_cplx s13 = cx_add(g1, g3), d13 = cx_sub(g1, g3);
    *y0 = cx_add(s02, s13);
    *y2 = cx_sub(s02, s13);
    if (inverse) {
        *y1 = cx(d02.re - d13.im, d02.im + d13.re);
        *y3 = cx(d02.re + d13.im, d02.im - d13.re);
    } else {
        *y1 = cx(d02.re + d13.im, d02.im - d13.re);
        *y3 = cx(d02.re - d13.im, d02.im + d13.re);
    }
}

/* m-point transform of x[0], x[s], ..., x[(m-1)s] into y[0..m-1].  Twiddle for the
 * combine at local size L is w[k * (s*ws)] with ws = maxCount/n_top. */
static void dit4(const ocf_cplx* x, ocf_cplx* y, int m, int inverse, int s,
                 const ocf_cplx* w, int ws)
{
    if (m == 4) {
        /* leaf for sizes that are powers of 4 (fft_default.cpp:22-67) */
        radix4_combine(x[0], x[s], x[2 * s], x[3 * s], inverse, &y[0], &y[1], &y[2], &y[3]);
        return;
    }
    if (m == 8) {
        /* leaf for odd powers of two (fft_default.cpp:68-167): four 2-point
         * transforms, then one radix-4 combine without twiddles and one with. */
        ocf_cplx p[8];
        for (int j = 0; j < 4; ++j) {
            ocf_cplx a = x[j * s], b = x[(j + 4) * s];
            p[2 * j] = cx_add(a, b);
            p[2 * j + 1] = cx_sub(a, b);
        }
        radix4_combine(p[0], p[2], p[4], p[6], inverse, &y[0], &y[2], &y[4], &y[6]);
        const int t = s * ws;
        radix4_combine(p[1], cx_mul(p[3], w[t]), cx_mul(p[5], w[2 * t]), cx_mul(p[7], w[3 * t]),
                       inverse, &y[1], &y[3], &y[5], &y[7]);
        return;
    }
    /* general level (fft_default.cpp:168-266) */
    const int q = m / 4;
    for (int j = 0; j < 4; ++j)
        dit4(x + j * s, y + j * q, q, inverse, 4 * s, w, ws);
    const int t = s * ws;
    for (int i = 0; i < q; ++i) {
        /* the i == 0 column is multiplied by w[0] too, as in the reference */
        ocf_cplx g1 = cx_mul(y[q + i], w[i * t]);
        ocf_cplx g2 = cx_mul(y[2 * q + i], w[2 * i * t]);
        ocf_cplx g3 = cx_mul(y[3 * q + i], w[3 * i * t]);
        radix4_combine(y[i], g1, g2, g3, inverse, &y[i], &y[q + i], &y[2 * q + i], &y[3 * q + i]);
    }
}

/* src/ckfft/fft.cpp:13-46 */
static void complex_core(const ocf_ctx* c, int n, const ocf_cplx* in, ocf_cplx* out, int inverse)
{
    if (n == 1) {
        out[0] = in[0];
    } else if (n == 2) {
        out[0] = cx_add(in[0], in[1]);
        out[1] = cx_sub(in[0], in[1]);
    } else {
        dit4(in, out, n, inverse, 1, inverse ? c->inv : c->fwd, c->max_n / n);
    }
}

/* src/ckfft/fft_real.cpp:13-60 and src/ckfft/fft_real_default.cpp:13-63 */
static void real_forward_core(const ocf_ctx* c, int n, const float* x, ocf_cplx* y)
{
    if (n == 1) {
        y[0] = cx(x[0] * 2.0f, 0.0f);
        return;
    }
    if (n == 2) {
        y[0] = cx((x[0] + x[1]) * 2.0f, 0.0f);
        y[1] = cx((x[0] - x[1]) * 2.0f, 0.0f);
        return;
    }
    if (n == 4) {
        float s02 = (x[0] + x[2]) * 2.0f, d02 = (x[0] - x[2]) * 2.0f;
        float s13 = (x[1] + x[3]) * 2.0f, d13 = (x[1] - x[3]) * 2.0f;
        y[0] = cx(s02 + s13, 0.0f);
        y[1] = cx(d02, -d13);
        y[2] = cx(s02 - s13, 0.0f);
        return;   /* the reference also stores y[3] = (d02, d13) past the n/2+1 outputs */
    }
    const int h = n / 2;
    dit4((const ocf_cplx*)x, y, h, 0, 1, c->fwd, c->max_n / h);
    y[h] = y[0];
    const int ws = c->max_n / n;
    for (int i = 0; i < n / 4; ++i) {
        ocf_cplx z0 = y[i], z1 = y[h - i];
        ocf_cplx sum = cx(z0.re + z1.re, z0.im - z1.im);
        ocf_cplx dif = cx(z0.re - z1.re, z0.im + z1.im);
        ocf_cplx a = c->fwd[i * ws];
        y[i] = cx_sub(sum, cx_mul(cx(-a.im, a.re), dif));
        dif.re = -dif.re;
        sum.im = -sum.im;
        ocf_cplx b = c->fwd[(h - i) * ws];
        y[h - i] = cx_sub(sum, cx_mul(cx(-b.im, b.re), dif));
    }
    ocf_cplx mid = y[n / 4];
    y[n / 4] = cx(mid.re * 2.0f, -mid.im * 2.0f);
}

/* src/ckfft/fft_real.cpp:62-105 and src/ckfft/fft_real_default.cpp:65-114 */
static void real_inverse_core(const ocf_ctx* c, int n, const ocf_cplx* y, float* x, ocf_cplx* tmp)
{
    if (n == 1) {
        x[0] = y[0].re;
        return;
    }
    if (n == 2) {
        x[0] = y[0].re + y[1].re;
        x[1] = y[0].re - y[1].re;
        return;
    }
    if (n == 4) {
        float s02 = y[0].re + y[2].re;
        float s13 = 2.0f * y[1].re;
        ocf_cplx d02 = cx_sub(y[0], y[2]);
        float d13i = 2.0f * y[1].im;
        x[0] = s02 + s13;
        x[1] = d02.re - d13i;
        x[2] = s02 - s13;
        x[3] = d02.re + d13i;
        return;
    }
    const int h = n / 2;
    const int ws = c->max_n / n;
    for (int i = 0; i < n / 4; ++i) {
        ocf_cplx z0 = y[i], z1 = y[h - i];
        ocf_cplx sum = cx(z0.re + z1.re, z0.im - z1.im);
        ocf_cplx dif = cx(z0.re - z1.re, z0.im + z1.im);
        ocf_cplx a = c->inv[i * ws];
        tmp[i] = cx_add(sum, cx_mul(cx(-a.im, a.re), dif));
        dif.re = -dif.re;
        sum.im = -sum.im;
        ocf_cplx b = c->inv[(h - i) * ws];
        tmp[h - i] = cx_add(sum, cx_mul(cx(-b.im, b.re), dif));
    }
    tmp[n / 4] = cx(y[n / 4].re * 2.0f, -y[n / 4].im * 2.0f);
    dit4(tmp, (ocf_cplx*)x, h, 1, 1, c->inv, c->max_n / h);
}

/* ---- validated entry points: src/ckfft/ckfft.cpp:36-114 --------------------- */

static int size_ok(const ocf_ctx* c, int n) { return pow2_positive(n) && n <= c->max_n; }

int ocf_complex(const ocf_ctx* c, int n, const ocf_cplx* in, ocf_cplx* out, int inverse)
{
    if (!c || !(inverse ? c->inv : c->fwd)) return 0;
    if (!size_ok(c, n)) return 0;
    if (!in || !out || (const void*)in == (const void*)out) return 0;
    complex_core(c, n, in, out, inverse);
    return 1;
}

int ocf_real_forward(const ocf_ctx* c, int n, const float* in, ocf_cplx* out)
{
    if (!c || !c->fwd) return 0;
    if (!size_ok(c, n)) return 0;
    if (!in || !out || (const void*)in == (const void*)out) return 0;
    real_forward_core(c, n, in, out);
    return 1;
}

int ocf_real_inverse(const ocf_ctx* c, int n, const ocf_cplx* in, float* out, ocf_cplx* tmp)
{
    if (!tmp) return 0;
    if (!c || !c->inv) return 0;
    if (!size_ok(c, n)) return 0;
    if (!in || !out || (const void*)in == (const void*)out) return 0;
    real_inverse_core(c, n, in, out, tmp);
    return 1;
}

/* ---- batched drivers for CPU timing ----------------------------------------- */

typedef struct {
    const ocf_ctx* c;
    int n, kind, lo, hi;        /* kind: 0 cfwd, 1 cinv, 2 rfwd, 3 rinv */
    const void* in;
    void* out;
} ocf_job;

static void* ocf_worker(void* arg)
{
    ocf_job* j = (ocf_job*)arg;
    const int n = j->n;
    const size_t half = (size_t)n / 2 + 1;
    ocf_cplx* tmp = NULL;
    if (j->kind == 3) tmp = (ocf_cplx*)malloc(sizeof(ocf_cplx) * half);
    for (int b = j->lo; b < j->hi; ++b) {
        switch (j->kind) {
        case 0:
        case 1:
            complex_core(j->c, n, (const ocf_cplx*)j->in + (size_t)b * n,
                         (ocf_cplx*)j->out + (size_t)b * n, j->kind == 1);
            break;
        case 2:
            real_forward_core(j->c, n, (const float*)j->in + (size_t)b * n,
                              (ocf_cplx*)j->out + (size_t)b * half);
            break;
        default:
            real_inverse_core(j->c, n, (const ocf_cplx*)j->in + (size_t)b * half,
                              (float*)j->out + (size_t)b * n, tmp);
            break;
        }
    }
    free(tmp);
    return NULL;
}

static int run_batch(const ocf_ctx* c, int n, int count, int kind, const void* in, void* out,
                     int threads)
{
    if (!c || !size_ok(c, n) || count < 0 || !in || !out || in == out) return 0;
    if (!((kind == 0 || kind == 2) ? c->fwd : c->inv)) return 0;
    if (threads < 1) threads = 1;
    if (threads > count) threads = count > 0 ? count : 1;
    ocf_job* jobs = (ocf_job*)calloc((size_t)threads, sizeof(ocf_job));
    pthread_t* tid = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
    if (!jobs || !tid) { free(jobs); free(tid); return 0; }
    for (int t = 0; t < threads; ++t) {
        jobs[t].c = c; jobs[t].n = n; jobs[t].kind = kind; jobs[t].in = in; jobs[t].out = out;
        jobs[t].lo = (int)((long long)count * t / threads);
        jobs[t].hi = (int)((long long)count * (t + 1) / threads);
    }
    if (threads == 1) {
        ocf_worker(&jobs[0]);
    } else {
        for (int t = 0; t < threads; ++t) pthread_create(&tid[t], NULL, ocf_worker, &jobs[t]);
        for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
    }
    free(jobs);
    free(tid);
    return 1;
}

int ocf_complex_batch(const ocf_ctx* c, int n, int count, const ocf_cplx* in, ocf_cplx* out,
                      int inverse, int threads)
{
    return run_batch(c, n, count, inverse ? 1 : 0, in, out, threads);
}

int ocf_real_forward_batch(const ocf_ctx* c, int n, int count, const float* in, ocf_cplx* out,
                           int threads)
{
    return run_batch(c, n, count, 2, in, out, threads);
}

int ocf_real_inverse_batch(const ocf_ctx* c, int n, int count, const ocf_cplx* in, float* out,
                           int threads)
{
    return run_batch(c, n, count, 3, in, out, threads);
}
