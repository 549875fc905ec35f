/*
 * oracle.c -- CPU restatement of glsCV's FFT / phase-correlation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path and
 * the timed CPU baseline; nothing in the product library links or calls it.
 *
 * Contents (each function cites the reference lines it restates):
 *   orc_fft2d_f64            FP64 2D DFT (Stockham radix-2, exact twiddles) with
 *                            the gls::fft flag semantics.  The numerical ground
 *                            truth for rel-RMS checks.
 *   orc_fft2d_r4_f32         FP32 restatement of fft_dit_Stockham_radix4_type0
 *                            (UnitTest_glsCV/fft_stockhum.cpp:771-1033) with the
 *                            flag handling of glsCV/glsFft.cpp:586-664.  This is
 *                            "the reference's CPU path" used as the CPU baseline.
 *   orc_fft2d_r4_f32_mt      the same, rows/columns split over pthreads (the
 *                            reference's commented-out omp loops,
 *                            fft_stockhum.cpp:461-462,501-502).
 *   orc_poc_f32              glsCV/glsBasicOperation.cpp:195-206 (POC), with an
 *                            epsilon guard (0/0 -> 0 instead of NaN).
 *   orc_mul_spectrums_f32    glsCV/glsBasicOperation.cpp:320-327 (binary shader
 *                            cases 6 and 7, :182-193: a*b and a*conj(b)).
 *   orc_min_max_loc_f32      glsCV/glsMinMaxLoc.cpp:65-118, 149-199 (both passes:
 *                            per-row scans, then the scan over the rows).
 *   orc_argmax_first_f32     glsCV/glsMinMaxLoc.cpp:65-118 (first maximum in
 *                            row-major order; NaN never wins).
 *   orc_weighted_centroid    glsCV/glsPhaseCorrelate.cpp:60-137 (5x5 box,
 *                            clamped at the borders, FP64 sums).
 *   orc_phase_correlate      glsCV/glsPhaseCorrelate.cpp:142-231 (the chain).
 *   orc_mag_spectrums_f32,   glsCV/glsBasicOperation.cpp:337-351 (scalar shader
 *   orc_log_mag_spectrums_f32  cases 9 and 13, :121-135).
 *   orc_normalize_f32        glsCV/glsNormalize.cpp:52-75 with glsNorm.cpp norms.
 *
 * Fixes relative to fft_stockhum.cpp (all behaviour-preserving for the
 * transforms the reference supports):
 *   - the butterfly loops' `n = n++` (never advances under GCC/Clang) -> n++;
 *   - the ping-pong bank is kept per transform and per stage (the reference
 *     toggles a shared bank per row, which is wrong for odd stage counts);
 *   - per-element printing removed;
 *   - 1D INVERSE conjugates both input and output (the reference skips one of
 *     the two conjugations for 1-row / 1-column inputs);
 *   - length-2 and length-1 dimensions are handled (the reference requires >=4);
 *   - rows and columns may have different power-of-two lengths.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <pthread.h>

#include "oracle.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Planted-fault switch for the tests that prove the checks have teeth:
 * 0 = none, 1 = radix-4 twiddle index off by one, 2 = FP64 column pass with the
 * wrong sign, 3 = SHIFT applied to rows only, 4 = POC without the conjugate,
 * 5 = logMagSpectrums without the offset, 6 = NORM_MINMAX with alpha and beta
 * swapped, 7 = mulSpectrums with a conjugated result, 8 = minMaxLoc keeping the
 * last maximum of a row instead of the first. */
static int g_fault = 0;
void orc_plant_fault(int which) { g_fault = which; }

static int is_pow2(int x) { return x > 0 && (x & (x - 1)) == 0; }
static int ilog2(int x) { int q = 0; while ((1 << q) < x) q++; return q; }

/* ------------------------------------------------------------------------- */
/* FP64 reference                                                            */
/* ------------------------------------------------------------------------- */

/* In-place 1D DFT of n complex doubles at element stride `es` (in complex
 * units).  Stockham radix-2 autosort, twiddle w = exp(sign*2*pi*i*k/L)
 * evaluated directly in double for every stage (no recurrences). */
static void fft1d_f64(double *data, int n, long es, int sign, double *t0, double *t1)
{
    if (n == 1) return;
    for (int i = 0; i < n; i++) {
        t0[2 * i] = data[2 * i * es];
        t0[2 * i + 1] = data[2 * i * es + 1];
    }
    double *x = t0, *y = t1;
    /* Stockham: sub-transform length Ls doubles each stage. */
    for (int Ls = 1; Ls < n; Ls <<= 1) {
        int half = n / 2;
        for (int j = 0; j < half; j++) {
            int k = j % Ls;
            double ang = sign * 2.0 * M_PI * (double)k / (double)(2 * Ls);
            double wr = cos(ang), wi = sin(ang);
            double ar = x[2 * j], ai = x[2 * j + 1];
            double br0 = x[2 * (j + half)], bi0 = x[2 * (j + half) + 1];
            double br = br0 * wr - bi0 * wi, bi = br0 * wi + bi0 * wr;
            int o = (j - k) * 2 + k;
            y[2 * o] = ar + br;
            y[2 * o + 1] = ai + bi;
            y[2 * (o + Ls)] = ar - br;
            y[2 * (o + Ls) + 1] = ai - bi;
        }
        double *t = x; x = y; y = t;
    }
    for (int i = 0; i < n; i++) {
        data[2 * i * es] = x[2 * i];
        data[2 * i * es + 1] = x[2 * i + 1];
    }
}

int orc_fft2d_f64(const double *in, double *out, int rows, int cols, int flags)
{
    if (!is_pow2(rows) || !is_pow2(cols)) return -1;
    size_t total = (size_t)rows * cols;
    double *buf = (double *)malloc(total * 2 * sizeof(double));
    int mx = rows > cols ? rows : cols;
    double *t0 = (double *)malloc((size_t)mx * 2 * sizeof(double));
    double *t1 = (double *)malloc((size_t)mx * 2 * sizeof(double));
    if (!buf || !t0 || !t1) { free(buf); free(t0); free(t1); return -2; }
    memcpy(buf, in, total * 2 * sizeof(double));
    int sign = (flags & ORC_FFT_INVERSE) ? +1 : -1;
    for (int r = 0; r < rows; r++) fft1d_f64(buf + (size_t)r * cols * 2, cols, 1, sign, t0, t1);
    for (int c = 0; c < cols; c++) fft1d_f64(buf + (size_t)c * 2, rows, cols, g_fault == 2 ? -sign : sign, t0, t1);
    double s = (flags & ORC_FFT_SCALE) ? 1.0 / (double)total : 1.0;
    int hr = rows / 2, hc = cols / 2;
    for (int r = 0; r < rows; r++) {
        for (int c = 0; c < cols; c++) {
            int orow = r, ocol = c;
            if (flags & ORC_FFT_SHIFT) { orow = (r + hr) % rows; ocol = g_fault == 3 ? c : (c + hc) % cols; }
            size_t si = ((size_t)r * cols + c) * 2, di = ((size_t)orow * cols + ocol) * 2;
            out[di] = buf[si] * s;
            out[di + 1] = buf[si + 1] * s;
        }
    }
    free(buf); free(t0); free(t1);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* FP32 restatement of fft_dit_Stockham_radix4_type0                          */
/* ------------------------------------------------------------------------- */

typedef struct { float re, im; } cf;

static inline cf cmul(cf a, cf b)   /* mul() of fft_stockhum.cpp / glsFft.cpp:226-231 */
{
    cf r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}
static inline cf cadd(cf a, cf b) { cf r = { a.re + b.re, a.im + b.im }; return r; }
static inline cf csub(cf a, cf b) { cf r = { a.re - b.re, a.im - b.im }; return r; }
static inline cf cscl(cf a, cf s) { cf r = { a.re * s.re, a.im * s.im }; return r; }

/* insertZeroBits: fft_stockhum.cpp:256-266 / glsFft.cpp:220-225 */
static inline unsigned insert_zero_bits(unsigned src, int idx, int num)
{
    unsigned ret = src << num;
    ret &= ~((1u << (idx + num)) - 1u);
    ret |= src & ((1u << idx) - 1u);
    return ret;
}

/* Twiddle table w[n] = exp(-2*pi*i*n/N) with the angle rounded to float first
 * (fft_stockhum.cpp:805-810, glsFft.cpp:545-550). */
static void r4_twiddles(cf *w, int N)
{
    for (int n = 0; n < N / 2; n++) {
        float jw = (float)(-2 * M_PI * n / N);
        w[n].re = cosf(jw);
        w[n].im = sinf(jw);
    }
}

/* One 1D transform of length N held contiguously in x (clobbered; y scratch).
 * Returns the buffer that holds the result.  Stage schedule of
 * fft_stockhum.cpp:822-911: a twiddle-free radix-2 pre-pass when log2 N is odd
 * (:822-861), then radix-4 DIT stages p = 0,2,..; q = Q-2,..,0 (:864-911).
 * xscl is applied to the inputs of the first radix-4 stage (glsFft.cpp:600),
 * yscl to the outputs of the last one (glsFft.cpp:601,646-647). */
static cf *r4_transform(cf *x, cf *y, int N, const cf *w, cf xscl, cf yscl)
{
    if (N == 1) {
        x[0] = cscl(cscl(x[0], xscl), yscl);
        return x;
    }
    if (N == 2) {  /* below the reference's N>=4 limit: a single radix-2 butterfly */
        cf a = cscl(x[0], xscl), b = cscl(x[1], xscl);
        y[0] = cscl(cadd(a, b), yscl);
        y[1] = cscl(csub(a, b), yscl);
        return y;
    }
    int Q = ilog2(N);
    if (Q % 2) {
        for (int n = 0; n < N / 4; n++) {
            cf x0 = x[n], x1 = x[n + N / 4], x2 = x[n + N / 2], x3 = x[n + 3 * (N / 4)];
            y[n] = cadd(x0, x2);
            y[n + N / 4] = cadd(x1, x3);
            y[n + N / 2] = csub(x0, x2);
            y[n + 3 * (N / 4)] = csub(x1, x3);
        }
        Q -= 1;
        cf *t = x; x = y; y = t;
    }
    const cf j = { 0.0f, 1.0f };
    for (int p = 0, q = Q - 2; q >= 0; p += 2, q -= 2) {
        cf xs = { 1.0f, 1.0f }, ys = { 1.0f, 1.0f };
        if (p == 0) xs = xscl;
        if (q == 0) ys = yscl;
        for (int n = 0; n < N / 4; n++) {
            int iw = (n >> q) << q;
            int ix = (int)insert_zero_bits((unsigned)n, q, 2);
            const cf w1 = w[g_fault == 1 && iw + 1 < N / 2 ? iw + 1 : iw];
            const cf w2 = cmul(w1, w1);
            const cf w3 = cmul(w1, w2);
            const cf a = cscl(x[ix], xs);
            const cf b = cmul(cscl(x[ix + (1 << q)], xs), w1);
            const cf c = cmul(cscl(x[ix + (1 << q) * 2], xs), w2);
            const cf d = cmul(cscl(x[ix + (1 << q) * 3], xs), w3);
            const cf apc = cadd(a, c);
            const cf amc = csub(a, c);
            const cf bpd = cadd(b, d);
            const cf jbmd = cmul(j, csub(b, d));
            y[n] = cscl(cadd(apc, bpd), ys);
            y[n + N / 4] = cscl(csub(amc, jbmd), ys);
            y[n + N / 2] = cscl(csub(apc, bpd), ys);
            y[n + 3 * (N / 4)] = cscl(cadd(amc, jbmd), ys);
        }
        cf *t = x; x = y; y = t;
    }
    return x;
}

typedef struct {
    cf *data;
    int rows, cols;
    int dim;          /* 0: rows (length cols), 1: columns (length rows) */
    int first, last;  /* transform index range [first,last) */
    cf xscl, yscl;
    const cf *w;
} r4_job;

static void *r4_worker(void *arg)
{
    r4_job *jb = (r4_job *)arg;
    int N = jb->dim == 0 ? jb->cols : jb->rows;
    cf *t0 = (cf *)malloc((size_t)N * sizeof(cf));
    cf *t1 = (cf *)malloc((size_t)N * sizeof(cf));
    if (!t0 || !t1) { free(t0); free(t1); return (void *)1; }
    for (int b = jb->first; b < jb->last; b++) {
        long es = jb->dim == 0 ? 1 : jb->cols;
        cf *base = jb->dim == 0 ? jb->data + (size_t)b * jb->cols : jb->data + b;
        for (int i = 0; i < N; i++) t0[i] = base[i * es];
        cf *r = r4_transform(t0, t1, N, jb->w, jb->xscl, jb->yscl);
        for (int i = 0; i < N; i++) base[i * es] = r[i];
    }
    free(t0); free(t1);
    return NULL;
}

static int r4_dim(cf *data, int rows, int cols, int dim, cf xscl, cf yscl, int nthreads)
{
    int N = dim == 0 ? cols : rows;
    int count = dim == 0 ? rows : cols;
    cf *w = (cf *)malloc((size_t)(N / 2 + 1) * sizeof(cf));
    if (!w) return -2;
    r4_twiddles(w, N);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > count) nthreads = count;
    r4_job *jobs = (r4_job *)calloc((size_t)nthreads, sizeof(r4_job));
    pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    int rc = 0;
    for (int t = 0; t < nthreads; t++) {
        jobs[t].data = data; jobs[t].rows = rows; jobs[t].cols = cols; jobs[t].dim = dim;
        jobs[t].first = (int)((long)count * t / nthreads);
        jobs[t].last = (int)((long)count * (t + 1) / nthreads);
        jobs[t].xscl = xscl; jobs[t].yscl = yscl; jobs[t].w = w;
    }
    if (nthreads == 1) {
        if (r4_worker(&jobs[0])) rc = -2;
    } else {
        for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, r4_worker, &jobs[t]);
        for (int t = 0; t < nthreads; t++) { void *res; pthread_join(th[t], &res); if (res) rc = -2; }
    }
    free(jobs); free(th); free(w);
    return rc;
}

int orc_fft2d_r4_f32_mt(const float *in, float *out, int rows, int cols, int flags, int nthreads)
{
    if (!is_pow2(rows) || !is_pow2(cols)) return -1;
    size_t total = (size_t)rows * cols;
    cf *data = (cf *)malloc(total * sizeof(cf));
    if (!data) return -2;
    memcpy(data, in, total * sizeof(cf));
    const int inv = (flags & ORC_FFT_INVERSE) != 0, scl = (flags & ORC_FFT_SCALE) != 0;
    const int do_rows = cols != 1 || rows == 1, do_cols = rows != 1;
    /* glsFft.cpp:600-601: rows conjugate their input (INVERSE) and scale 1/N;
     * glsFft.cpp:646-647: columns conjugate their output (INVERSE) and scale. */
    cf one = { 1.0f, 1.0f };
    int rc = 0;
    if (do_rows) {
        cf xs = one, ys = one;
        if (inv) xs.im = -1.0f;
        if (scl) { ys.re = ys.im = 1.0f / (float)cols; }
        if (inv && !do_cols) ys.im = -ys.im;   /* 1D fix: conjugate the output too */
        rc = r4_dim(data, rows, cols, 0, xs, ys, nthreads);
    }
    if (!rc && do_cols) {
        cf xs = one, ys = one;
        if (inv && !do_rows) xs.im = -1.0f;    /* 1D fix: conjugate the input too */
        if (scl) { ys.re = ys.im = 1.0f / (float)rows; }
        if (inv) ys.im = -ys.im;
        rc = r4_dim(data, rows, cols, 1, xs, ys, nthreads);
    }
    if (rc) { free(data); return rc; }
    /* SHIFT: the tile permutation {2,3,0,1} of glsFft.cpp:656-664 is a
     * circular shift by half the length in each transformed dimension. */
    cf *o = (cf *)out;
    int hr = rows / 2, hc = cols / 2;
    for (int r = 0; r < rows; r++) {
        for (int c = 0; c < cols; c++) {
            int orow = r, ocol = c;
            if (flags & ORC_FFT_SHIFT) { orow = (r + hr) % rows; ocol = (c + hc) % cols; }
            o[(size_t)orow * cols + ocol] = data[(size_t)r * cols + c];
        }
    }
    free(data);
    return 0;
}

int orc_fft2d_r4_f32(const float *in, float *out, int rows, int cols, int flags)
{
    return orc_fft2d_r4_f32_mt(in, out, rows, cols, flags, 1);
}

/* ------------------------------------------------------------------------- */
/* Phase correlation chain                                                    */
/* ------------------------------------------------------------------------- */

/* glsBasicOperation.cpp:195-206: C = A*conj(B) / |A*conj(B)|.  The reference
 * divides 0/0 into NaN; here a magnitude below ORC_POC_EPS gives 0. */
void orc_poc_f32(const float *a, const float *b, float *c, long n)
{
    for (long i = 0; i < n; i++) {
        float ar = a[2 * i], ai = a[2 * i + 1], br = b[2 * i], bi = b[2 * i + 1];
        float re = ar * br + ai * bi;
        float im = -ar * bi + ai * br;
        if (g_fault == 4) { re = ar * br - ai * bi; im = ar * bi + ai * br; }
        float mag = sqrtf(re * re + im * im);
        if (mag > ORC_POC_EPS) { c[2 * i] = re / mag; c[2 * i + 1] = im / mag; }
        else { c[2 * i] = 0.0f; c[2 * i + 1] = 0.0f; }
    }
}

/* glsBasicOperation.cpp:320-327 with the binary shader's cases 6 and 7
 * (:182-193): c = a*b, or a*conj(b) when conj != 0, on interleaved (re, im). */
void orc_mul_spectrums_f32(const float *a, const float *b, float *c, long n, int conj)
{
    for (long i = 0; i < n; i++) {
        float ar = a[2 * i], ai = a[2 * i + 1], br = b[2 * i], bi = b[2 * i + 1];
        float re, im;
        if (conj) { re = ar * br + ai * bi; im = -ar * bi + ai * br; }
        else { re = ar * br - ai * bi; im = ar * bi + ai * br; }
        if (g_fault == 7) im = -im;
        c[2 * i] = re;
        c[2 * i + 1] = im;
    }
}

/* glsMinMaxLoc.cpp:65-118 (shader) and :149-199 (driver), both passes as the
 * reference runs them.  Pass 0: every row scans channel 0 of its pixels (cn
 * interleaved channels) for its first strict minimum and maximum, starting
 * from FLT_MAX / -FLT_MAX at location 0.  Pass 1: the row results are scanned
 * the same way; the location is (x of the winning row, that row). */
int orc_min_max_loc_f32(const float *v, int rows, int cols, int cn, float *minv, float *maxv, int *minx,
                        int *miny, int *maxx, int *maxy)
{
    float *rmin = (float *)malloc(sizeof(float) * (size_t)rows);
    float *rmax = (float *)malloc(sizeof(float) * (size_t)rows);
    int *rminl = (int *)malloc(sizeof(int) * (size_t)rows);
    int *rmaxl = (int *)malloc(sizeof(int) * (size_t)rows);
    if (!rmin || !rmax || !rminl || !rmaxl) { free(rmin); free(rmax); free(rminl); free(rmaxl); return -2; }
    for (int y = 0; y < rows; y++) {
        float mn = FLT_MAX, mx = -FLT_MAX;
        int mnl = 0, mxl = 0;
        for (int x = 0; x < cols; x++) {
            const float d = v[((size_t)y * cols + x) * cn];
            if (d < mn) { mn = d; mnl = x; }
            if (g_fault == 8 ? d >= mx : d > mx) { mx = d; mxl = x; }
        }
        rmin[y] = mn; rmax[y] = mx; rminl[y] = mnl; rmaxl[y] = mxl;
    }
    float mn = FLT_MAX, mx = -FLT_MAX;
    int mny = 0, mxy = 0;
    for (int y = 0; y < rows; y++) {
        if (rmin[y] < mn) { mn = rmin[y]; mny = y; }
        if (rmax[y] > mx) { mx = rmax[y]; mxy = y; }
    }
    *minv = mn; *maxv = mx;
    *minx = rminl[mny]; *miny = mny;
    *maxx = rmaxl[mxy]; *maxy = mxy;
    free(rmin); free(rmax); free(rminl); free(rmaxl);
    return 0;
}

/* glsMinMaxLoc.cpp:65-118: a strict '>' scan per row, then across rows; the
 * result is the first maximum in row-major order.  NaN compares false and is
 * never selected; an all-NaN input reports (0,0). */
void orc_argmax_first_f32(const float *v, int rows, int cols, int *px, int *py, float *pmax)
{
    float best = -FLT_MAX;
    int bx = 0, by = 0;
    for (int y = 0; y < rows; y++) {
        for (int x = 0; x < cols; x++) {
            float d = v[(size_t)y * cols + x];
            if (d > best) { best = d; bx = x; by = y; }
        }
    }
    *px = bx; *py = by;
    if (pmax) *pmax = best;
}

/* glsPhaseCorrelate.cpp:60-137 */
void orc_weighted_centroid(const float *v, int rows, int cols, int peak_x, int peak_y,
                           int box_w, int box_h, double *cx, double *cy, double *response)
{
    int minr = peak_y - (box_h >> 1), maxr = peak_y + (box_h >> 1);
    int minc = peak_x - (box_w >> 1), maxc = peak_x + (box_w >> 1);
    if (minr < 0) minr = 0;
    if (minc < 0) minc = 0;
    if (maxr > rows - 1) maxr = rows - 1;
    if (maxc > cols - 1) maxc = cols - 1;
    double sx = 0.0, sy = 0.0, sum = 0.0;
    for (int y = minr; y <= maxr; y++) {
        const float *row = v + (size_t)y * cols;
        for (int x = minc; x <= maxc; x++) {
            sx += (double)x * row[x];
            sy += (double)y * row[x];
            sum += (double)row[x];
        }
    }
    if (response) *response = sum;
    sum += DBL_EPSILON;
    *cx = sx / sum;
    *cy = sy / sum;
}

/* glsPhaseCorrelate.cpp:142-231.  img1/img2/win are real rows x cols planes;
 * win may be NULL.  Writes the integer peak of the shifted correlation
 * surface, the sub-pixel shift (center - centroid) and the response / (M*N).
 * surface_out (optional, rows*cols floats) receives the real correlation
 * surface that the argmax and centroid ran on. */
int orc_phase_correlate(const float *img1, const float *img2, const float *win, int rows, int cols,
                        int nthreads, orc_poc_result *res, float *surface_out)
{
    if (!is_pow2(rows) || !is_pow2(cols)) return -1;
    size_t n = (size_t)rows * cols;
    float *a = (float *)calloc(n * 2, sizeof(float));
    float *b = (float *)calloc(n * 2, sizeof(float));
    float *c = (float *)malloc(n * 2 * sizeof(float));
    float *s = (float *)malloc(n * sizeof(float));
    int rc = -2;
    if (!a || !b || !c || !s) goto done;
    /* window multiply (:174-185) and real->complex with imag = 0 */
    for (size_t i = 0; i < n; i++) {
        float w = win ? win[i] : 1.0f;
        a[2 * i] = win ? w * img1[i] : img1[i];
        b[2 * i] = win ? w * img2[i] : img2[i];
    }
    if ((rc = orc_fft2d_r4_f32_mt(a, a, rows, cols, 0, nthreads))) goto done;
    if ((rc = orc_fft2d_r4_f32_mt(b, b, rows, cols, 0, nthreads))) goto done;
    orc_poc_f32(a, b, c, (long)n);
    if ((rc = orc_fft2d_r4_f32_mt(c, c, rows, cols, ORC_FFT_INVERSE | ORC_FFT_SHIFT, nthreads))) goto done;
    for (size_t i = 0; i < n; i++) s[i] = c[2 * i];   /* split: keep the real plane */
    orc_argmax_first_f32(s, rows, cols, &res->peak_x, &res->peak_y, NULL);
    double cx, cy, resp;
    orc_weighted_centroid(s, rows, cols, res->peak_x, res->peak_y, 5, 5, &cx, &cy, &resp);
    res->response = resp / ((double)rows * (double)cols);
    res->shift_x = (double)cols / 2.0 - cx;
    res->shift_y = (double)rows / 2.0 - cy;
    if (surface_out) memcpy(surface_out, s, n * sizeof(float));
    rc = 0;
done:
    free(a); free(b); free(c); free(s);
    return rc;
}

/* ---- spectrum display chain ------------------------------------------------
 * glsBasicOperation.cpp:337-343 magSpectrums: the scalar shader's case 9
 * (:121-123), dst = sqrt(re*re + im*im) on the RG32F texture, one R32F plane
 * out.  c holds n interleaved (re, im) pairs. */
void orc_mag_spectrums_f32(const float *c, float *m, long n)
{
    for (long i = 0; i < n; i++) {
        const float re = c[2 * i], im = c[2 * i + 1];
        const float rr = re * re, ii = im * im;
        m[i] = sqrtf(rr + ii);
    }
}

/* glsBasicOperation.cpp:345-351 logMagSpectrums: case 13 (:133-135),
 * dst = log(abs(sqrt(re*re + im*im) + offset)). */
void orc_log_mag_spectrums_f32(const float *c, float *m, long n, float offset)
{
    for (long i = 0; i < n; i++) {
        const float re = c[2 * i], im = c[2 * i + 1];
        const float rr = re * re, ii = im * im;
        const float a = sqrtf(rr + ii) + (g_fault == 5 ? 0.0f : offset);
        m[i] = logf(fabsf(a));
    }
}

/* glsNormalize.cpp:52-75.  NORM_MINMAX: min and max over the plane (as
 * doubles, gls::minMaxLoc), scl = (float)(beta - alpha) / (float)(max - min),
 * dst = scl*src + (float)(alpha - scl*min), two float operations per element
 * (gls::multiply then gls::add).  A constant plane (max == min) maps to alpha
 * (the reference divides by zero there).  NORM_L1 / NORM_L2: glsNorm.cpp
 * norm = sum |x| / sqrt(sum x^2) (summed here in FP64), scl = (float)alpha /
 * (float)norm, dst = scl*src; beta is unused.  NaN elements are skipped by the
 * min/max (fmin/fmax). */
int orc_normalize_f32(const float *src, float *dst, long n, double alpha, double beta, int norm_type)
{
    if (n <= 0) return -1;
    if (norm_type == ORC_NORM_MINMAX) {
        float mn = INFINITY, mx = -INFINITY;
        for (long i = 0; i < n; i++) { mn = fminf(mn, src[i]); mx = fmaxf(mx, src[i]); }
        const double minv = mn, maxv = mx;
        float scl = (maxv > minv) ? (float)(beta - alpha) / (float)(maxv - minv) : 0.0f;
        if (g_fault == 6) scl = (float)(alpha - beta) / (float)(maxv - minv);
        const float off = (float)(alpha - (double)scl * minv);
        for (long i = 0; i < n; i++) {
            const float t = scl * src[i];
            dst[i] = t + off;
        }
        return 0;
    }
    if (norm_type == ORC_NORM_L1 || norm_type == ORC_NORM_L2) {
        double acc = 0.0;
        for (long i = 0; i < n; i++) acc += norm_type == ORC_NORM_L1 ? fabs((double)src[i]) : (double)src[i] * src[i];
        const double norm = norm_type == ORC_NORM_L2 ? sqrt(acc) : acc;
        const float scl = norm > 0.0 ? (float)alpha / (float)norm : 0.0f;
        for (long i = 0; i < n; i++) dst[i] = scl * src[i];
        return 0;
    }
    return -2;
}
