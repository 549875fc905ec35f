/*
 * oracle.h -- CPU restatement of glsCV's FFT / phase-correlation path.
 * TEST INFRASTRUCTURE ONLY: used by tests and by the CPU-baseline leg of the
 * benchmark as the checker; never linked into the product library.
 */
#ifndef GLSX_ORACLE_H
#define GLSX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* gls::fft flag values (glsCV/glsFft.h:43-47) */
#define ORC_FFT_INVERSE 1
#define ORC_FFT_SCALE 2
#define ORC_FFT_SHIFT 64

/* magnitude below which the POC quotient is defined as 0 */
#define ORC_POC_EPS 1e-30f

typedef struct {
    int peak_x, peak_y;        /* integer argmax of the shifted correlation surface */
    double shift_x, shift_y;   /* (cols/2, rows/2) - weighted centroid */
    double response;           /* 5x5 sum / (rows*cols) */
} orc_poc_result;

void orc_plant_fault(int which);   /* test hook: see oracle.c */
int orc_fft2d_f64(const double *in, double *out, int rows, int cols, int flags);
int orc_fft2d_r4_f32(const float *in, float *out, int rows, int cols, int flags);
int orc_fft2d_r4_f32_mt(const float *in, float *out, int rows, int cols, int flags, int nthreads);
void orc_poc_f32(const float *a, const float *b, float *c, long n);
void orc_mul_spectrums_f32(const float *a, const float *b, float *c, long n, int conj);
int orc_min_max_loc_f32(const float *v, int rows, int cols, int cn, float *minv, float *maxv, int *minx,
                        int *miny, int *maxx, int *maxy);
void orc_argmax_first_f32(const float *v, int rows, int cols, int *px, int *py, float *pmax);
void orc_weighted_centroid(const float *v, int rows, int cols, int peak_x, int peak_y,
                           int box_w, int box_h, double *cx, double *cy, double *response);
int orc_phase_correlate(const float *img1, const float *img2, const float *win, int rows, int cols,
                        int nthreads, orc_poc_result *res, float *surface_out);

/* spectrum display chain (glsCam.cpp:196-202) */
#define ORC_NORM_L1 2
#define ORC_NORM_L2 4
#define ORC_NORM_MINMAX 32
void orc_mag_spectrums_f32(const float *c, float *m, long n);
void orc_log_mag_spectrums_f32(const float *c, float *m, long n, float offset);
int orc_normalize_f32(const float *src, float *dst, long n, double alpha, double beta, int norm_type);

#ifdef __cplusplus
}
#endif
#endif
