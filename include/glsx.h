/*
 * glsx.h -- C ABI of the MI355X (gfx950) implementation of glsCV's 2D FFT and
 * phase-correlation path.
 *
 * Plain C types only: opaque handles, raw device/host pointers, sizes and int
 * status codes.  The C++ API in glsMat.h / glsFft.h / glsPhaseCorrelate.h
 * (namespace gls, same signatures as the reference) is a thin layer over
 * these entry points; any other host language binds them directly.
 *
 * Status: every function returns GLSX_OK (0) or a negative GLSX_ERR_* code;
 * glsx_last_error() returns the message of the calling thread's last failure.
 * The reference throws cv::Exception from GLS_Assert (glsCV/glsMacro.h:40);
 * the C++ layer turns a non-zero status back into a thrown gls::Exception.
 *
 * Streams: `stream` is a hipStream_t passed as void*; NULL means the default
 * stream.  All work is stream-ordered; functions that return host results
 * (download, phase correlation into a host struct) synchronise that stream.
 */
#ifndef GLSX_H
#define GLSX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------- */
#define GLSX_OK 0
#define GLSX_ERR_ARG (-1)         /* bad argument: shape, type, null pointer  */
#define GLSX_ERR_UNSUPPORTED (-2) /* valid but outside the supported contract */
#define GLSX_ERR_HIP (-3)         /* HIP runtime failure                       */
#define GLSX_ERR_NOMEM (-4)       /* device or host allocation failed          */
#define GLSX_ERR_STATE (-5)       /* not initialised / wrong state             */
#define GLSX_ERR_BUDGET (-6)      /* request exceeds an HBM budget             */

/* ---- gls::fft flags: glsCV/glsFft.h:43-47 (same numeric values) ------------- */
#define GLSX_FFT_INVERSE 1 /* DFT_INVERSE: unnormalised +i transform        */
#define GLSX_FFT_SCALE 2   /* DFT_SCALE: multiply by 1/(rows*cols)          */
#define GLSX_FFT_SHIFT 64  /* fftshift of the output (quadrant swap)         */
/* Extension, no counterpart in glsFft.h: run the radix-16 stages as 16x16
 * DFT-matrix products on the matrix cores (3-term BF16 hi/lo split, FP32
 * accumulation).  Honoured per transform length only after that length has
 * passed the FP64 check of glsx_fft_split_mfma_check(); otherwise, and for
 * lengths below 16, the FP32 kernels run.  Output bits differ from the FP32
 * kernels' (same error bound). */
#define GLSX_FFT_SPLIT_MFMA 0x10000
/* Relative RMS error vs FP64 a length must hold to use GLSX_FFT_SPLIT_MFMA:
 * 1e-5 / sqrt(2) per dimension, so that a 2D transform stays within 1e-5. */
#define GLSX_SPLIT_MFMA_TOL 7.0710678118654752e-6

/* ---- cv type codes (OpenCV encoding, glsCV/glsMat.h type()) ---------------- */
#define GLSX_CV_8UC1 0
#define GLSX_CV_32SC1 4
#define GLSX_CV_32FC1 5
#define GLSX_CV_32FC2 13

/* largest supported transform length per dimension */
#define GLSX_MAX_FFT_LEN 16384

/* ---- context (glsCV/glsCV.h:67-68 glsCvInit / glsCvTerminate) -------------- */
int glsx_init(int device);
/* Waits for every initialised device, frees the library's tables, scratch
 * and cached GlsMat blocks.  Plans and servers use those tables: while any is
 * alive, glsx_shutdown changes nothing and returns GLSX_ERR_STATE -- destroy
 * them first.  GlsMat buffers may outlive it (their memory is their own). */
int glsx_shutdown(void);
const char *glsx_last_error(void);

typedef struct {
    char name[64];
    char arch[32];
    int compute_units;
    size_t lds_per_block;
    size_t global_mem_bytes;
} glsx_device_info;
int glsx_get_device_info(glsx_device_info *out);

/* ---- GlsMat storage (glsCV/glsMat.h:45-214, glsMat.cpp:67-637) --------------
 * A refcounted device buffer: rows x cols elements of cv type `type`, row
 * major, pitch = cols * elem_size.  create -> refs = 1; the release that drops
 * the count to 0 gives the device memory back (GlsMat::~GlsMat,
 * glsMat.cpp:121-125), stream-ordered: the block goes to a per-device cache
 * and is handed to a later glsx_buf_create of the same size only after the
 * work that used it has finished (an event recorded at the release), so the
 * release neither blocks nor synchronises other streams.  hipFree, which
 * waits for the whole device, runs only when the cache exceeds 4 GiB per
 * device, when an allocation fails, and at glsx_shutdown.
 *   glsx_buf_release: the library knows the stream of every call that took
 *     the handle (it records an event there after each such call); the block
 *     is reusable after the last of them, and the stream may have been
 *     destroyed by then.  If the buffer was used on two different streams,
 *     or its device pointer was handed out (glsx_buf_info with dev_ptr), the
 *     library cannot see all of its uses and this release waits for the
 *     device first.
 *   glsx_buf_release_async: the caller states that every use of the buffer
 *     is ordered before the work enqueued so far on `stream` (which must
 *     exist); the block is reusable once `stream` passes this point.  Never
 *     waits. */
typedef struct glsx_buf_s *glsx_buf;
int glsx_buf_create(int rows, int cols, int type, glsx_buf *out);
int glsx_buf_retain(glsx_buf buf);
int glsx_buf_release(glsx_buf buf);
int glsx_buf_release_async(glsx_buf buf, void *stream);
int glsx_buf_info(glsx_buf buf, int *rows, int *cols, int *type, void **dev_ptr, size_t *bytes);
/* GlsMat::upload (glsMat.cpp:374-419); host_pitch 0 = tightly packed */
int glsx_buf_upload(glsx_buf buf, const void *host, size_t host_pitch, void *stream);
/* GlsMat::download (glsMat.cpp:493-552); synchronises `stream` */
int glsx_buf_download(glsx_buf buf, void *host, size_t host_pitch, void *stream);
/* GlsMat::copyTo / clone (glsMat.cpp:629-637): dst must match size and type */
int glsx_buf_copy(glsx_buf src, glsx_buf dst, void *stream);

/* Asynchronous transfers (the reference's disabled PBO upload/download paths,
 * glsMat.cpp:55-56, 311-552).  Pinned host memory from glsx_host_alloc makes
 * the copies DMA transfers that overlap kernels on other streams; the _async
 * copies are stream-ordered and return at once, so `host` must stay valid
 * (and, for uploads, unchanged) until the stream has passed the copy --
 * glsx_stream_sync or an event.  With pageable memory the runtime stages
 * the copy and may block. */
int glsx_host_alloc(size_t bytes, void **out);
int glsx_host_free(void *ptr);
int glsx_buf_upload_async(glsx_buf buf, const void *host, size_t host_pitch, void *stream);
int glsx_buf_download_async(glsx_buf buf, void *host, size_t host_pitch, void *stream);
int glsx_stream_sync(void *stream);

/* ---- FFT: gls::fft(const GlsMat& src, GlsMat& dst, int flag), glsFft.h:55 ----
 * src: CV_32FC2 (interleaved complex) or CV_32FC1 (real, imag taken as 0);
 * dst: CV_32FC2 of the same size (in place, src == dst, is allowed).
 * rows, cols: powers of two, 1..GLSX_MAX_FFT_LEN (1 x N and N x 1 are 1D).
 * flags: any combination of GLSX_FFT_INVERSE | GLSX_FFT_SCALE | GLSX_FFT_SHIFT
 * (| GLSX_FFT_SPLIT_MFMA).
 * Some shapes use scratch the library keeps per stream (allocated on the first
 * such call, kept until glsx_shutdown): complex transforms of 8192 rows and
 * 2048..8192 columns (one array of the transform's size, rows*cols*8 bytes)
 * and the shifted real-input transform of 4096 x 4096 (the same size). */
int glsx_fft2d(glsx_buf src, glsx_buf dst, int flags, void *stream);
/* Same on raw device pointers (for callers that own their memory). */
int glsx_fft2d_dev(const void *src, int src_type, void *dst, int rows, int cols, int flags, void *stream);

/* Batched 1D transforms on raw device pointers: `count` transforms of length
 * `len` (power of two, 1..GLSX_MAX_FFT_LEN); element i of transform b is at
 * index b*batch_stride + i*elem_stride (in elements of src / dst).  This is
 * one sweep of the 2D FFT: rows are (elem_stride 1, batch_stride cols),
 * columns are (elem_stride cols, batch_stride 1).  Flags as for glsx_fft2d,
 * applied to this dimension only (INVERSE conjugates input and output,
 * SCALE is 1/len, SHIFT rotates the output by len/2).  The reference runs
 * these as its per-dimension stage loops (glsFft.cpp:562-652). */
int glsx_fft1d_dev(const void *src, int src_type, void *dst, int len, long long count, long long elem_stride,
                   long long batch_stride, int flags, void *stream);

/* FP64 check of the split-precision MFMA passes for transform length `len`
 * (no counterpart in the reference: it gates GLSX_FFT_SPLIT_MFMA).  The first
 * call per device and length runs a batch of pseudo-random transforms through
 * the MFMA pass and through a direct FP64 DFT on the device and keeps the
 * relative RMS error; later calls return the kept result.  *rel_rms: the
 * error; *enabled: 1 when it is <= GLSX_SPLIT_MFMA_TOL (the flag is then
 * honoured for this length), else 0.  Synchronous: call it before capturing a
 * stream that uses the flag (an unchecked length is checked on first use). */
int glsx_fft_split_mfma_check(int len, double *rel_rms, int *enabled);

/* ---- phase correlation: gls::phaseCorrelate, glsPhaseCorrelate.h:58 --------
 * src1, src2: CV_32FC1 rows x cols (powers of two, rows 2..16384, cols
 * 2..8192); window: CV_32FC1 of the same size or NULL.  Both images go through
 * one complex FFT (img1 + i img2) whose two spectra are separated by symmetry.
 * Result: integer peak of the fft-shifted correlation
 * surface (first maximum in row-major order), the sub-pixel shift
 * (cols/2, rows/2) - 5x5 weighted centroid, and the 5x5 response / (rows*cols)
 * (glsPhaseCorrelate.cpp:60-137, 216-230). */
typedef struct {
    int peak_x, peak_y;
    double shift_x, shift_y;
    double response;
} glsx_poc_result;
int glsx_phase_correlate(glsx_buf src1, glsx_buf src2, glsx_buf window, glsx_poc_result *out, void *stream);
/* Raw device pointers; `out` is a host struct, the call synchronises `stream`.
 * The chain's workspace (12 bytes per pixel) is scratch the library keeps per
 * stream: calls on different streams run concurrently, calls on one stream
 * from several host threads are serialised. */
int glsx_phase_correlate_dev(const float *src1, const float *src2, const float *window, int rows, int cols,
                             glsx_poc_result *out, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate: code
 * under HIP graph capture, serving loops and harnesses that forbid hipMalloc.
 * The chain of glsx_phase_correlate_dev with no device or host allocation
 * and no synchronisation: its workspace is a static device array of
 * GLSX_NOALLOC_WORKSPACE_BYTES that each device holds for the life of the
 * process (up to 2048 x 2048 images; larger ones are rejected with
 * GLSX_ERR_UNSUPPORTED), and the result is written to `dev_out`, device
 * memory of the caller's, in stream order.  The device must have been set up
 * with glsx_init (GLSX_ERR_STATE otherwise).  Every call on one device shares
 * that device's array: calls that overlap on one device -- on two streams, or
 * from two host threads -- race and give wrong results; order them (one
 * stream, one thread, or events). */
#define GLSX_NOALLOC_WORKSPACE_BYTES ((size_t)64 << 20)
int glsx_phase_correlate_dev_noalloc(const float *src1, const float *src2, const float *window, int rows, int cols,
                                     glsx_poc_result *dev_out, void *stream);

/* ---- the chain's operations one at a time (glsPhaseCorrelate.cpp:174-210) ----
 * glsx_phase_correlate runs these fused into the FFT passes; here each is its
 * own launch, for callers that compose the chain or use the operations
 * elsewhere.  Buffers must match in size; in place (dst == a source) is
 * allowed for multiply, the spectrum products and a one-plane merge.  The
 * _dev forms reject (GLSX_ERR_ARG) a destination that overlaps a source
 * without being it, and split / merge planes that overlap the image.
 *
 * gls::multiply(src0, src1, dst) (glsBasicOperation.cpp:265-272, shader case
 * 2): dst = src0 * src1 per float, CV_32FC1 or CV_32FC2 (all three the same
 * type).  _dev: n floats. */
int glsx_multiply(glsx_buf src0, glsx_buf src1, glsx_buf dst, void *stream);
int glsx_multiply_dev(const float *src0, const float *src1, float *dst, long long n, void *stream);
/* gls::mulSpectrums(src0, src1, dst, conj) (glsBasicOperation.cpp:320-327,
 * shader cases 6-7): dst = src0 * src1, or src0 * conj(src1) with conj != 0;
 * gls::mulSpectrumsPhaseOnly (:329-335, case 8): dst = c / |c| with
 * c = src0 * conj(src1), computed as the shader does (IEEE square root and
 * division), except that |c| <= 1e-30 (or NaN) gives 0 where the shader
 * divides 0/0 into NaN.  CV_32FC2 interleaved (re, im).  _dev: n complex
 * values, 8-byte aligned pointers, mode one of: */
#define GLSX_SPECTRUM_MUL 0
#define GLSX_SPECTRUM_MUL_CONJ 1
#define GLSX_SPECTRUM_PHASE_ONLY 2
int glsx_mul_spectrums(glsx_buf src0, glsx_buf src1, glsx_buf dst, int conj, void *stream);
int glsx_mul_spectrums_phase_only(glsx_buf src0, glsx_buf src1, glsx_buf dst, void *stream);
int glsx_mul_spectrums_dev(const void *src0, const void *src1, void *dst, long long n, int mode, void *stream);
/* gls::split (glsSplit.cpp:192-210): one CV_32FC1 plane per channel of a
 * CV_32FC1 / CV_32FC2 source (n_planes = its channel count, planes created by
 * the caller); gls::merge (glsMerge.cpp:193-211): 1 or 2 CV_32FC1 planes into
 * a CV_32FC1 / CV_32FC2 destination.  _dev: n pixels of two channels;
 * split_dev's plane1 may be NULL (channel 0 only, as phaseCorrelate keeps the
 * real plane). */
int glsx_split(glsx_buf src, const glsx_buf *planes, int n_planes, void *stream);
int glsx_merge(const glsx_buf *planes, int n_planes, glsx_buf dst, void *stream);
int glsx_split_dev(const float *src, float *plane0, float *plane1, long long n, void *stream);
int glsx_merge_dev(const float *plane0, const float *plane1, float *dst, long long n, void *stream);
/* gls::minMaxLoc (glsMinMaxLoc.cpp:149-199, shader :65-118) of channel 0 of a
 * CV_32FC1 / CV_32FC2 image: the first minimum and maximum in row-major order
 * (strict comparisons from FLT_MAX / -FLT_MAX: NaN is never chosen, and an
 * image with no value below FLT_MAX (above -FLT_MAX) reports FLT_MAX
 * (-FLT_MAX) at (0, 0)), locations as (x, y).  No mask (the reference asserts
 * an empty one).  glsx_min_max_loc synchronises `stream` and writes a host
 * struct, as the reference downloads its 1x1 results; the _dev form writes
 * `dev_out` (device memory) in stream order and does not synchronise (its
 * partials use scratch the library keeps per stream). */
typedef struct {
    float min_val, max_val;
    int min_x, min_y, max_x, max_y;
} glsx_minmax_result;
int glsx_min_max_loc(glsx_buf src, glsx_minmax_result *out, void *stream);
int glsx_min_max_loc_dev(const void *src, int src_type, int rows, int cols, glsx_minmax_result *dev_out,
                         void *stream);

/* ---- spectrum display chain (glsCam/glsCam.cpp:196-202) ---------------------
 * gls::magSpectrums (glsBasicOperation.cpp:337-343): dst = sqrt(re^2 + im^2);
 * gls::logMagSpectrums (:345-351): dst = log(abs(sqrt(re^2 + im^2) + offset)).
 * src CV_32FC2, dst CV_32FC1 of the same size. */
int glsx_mag_spectrums(glsx_buf src, glsx_buf dst, void *stream);
int glsx_log_mag_spectrums(glsx_buf src, glsx_buf dst, float offset, void *stream);
/* Raw device pointers: n complex values in, n floats out; log_mag selects
 * logMagSpectrums. */
int glsx_spectrum_dev(const void *src, void *dst, long long n, int log_mag, float offset, void *stream);

/* gls::normalize (glsNormalize.h:47, glsNormalize.cpp:52-75) on CV_32FC1
 * (src == dst allowed).  GLSX_NORM_MINMAX maps min -> alpha and max -> beta
 * (scl = (float)(beta - alpha) / (float)(max - min), dst = scl*src +
 * (float)(alpha - scl*min); a constant plane maps to alpha); GLSX_NORM_L1 /
 * GLSX_NORM_L2 scale to norm alpha (norms summed in FP64; beta unused). */
#define GLSX_NORM_L1 2      /* cv::NORM_L1 */
#define GLSX_NORM_L2 4      /* cv::NORM_L2 */
#define GLSX_NORM_MINMAX 32 /* cv::NORM_MINMAX */
int glsx_normalize(glsx_buf src, glsx_buf dst, double alpha, double beta, int norm_type, void *stream);
int glsx_normalize_dev(const float *src, float *dst, long long n, double alpha, double beta, int norm_type,
                       void *stream);

/* Extension (no single reference call; it fuses gls::fft with
 * gls::magSpectrums / logMagSpectrums): dst (CV_32FC1, rows x cols) =
 * |FFT2(src)|, or log(|FFT2(src)| + offset) when log_mag != 0, with the
 * flags of glsx_fft2d.  The magnitude is taken in the column pass's store, so
 * the complex spectrum is written and read once (through a scratch buffer
 * the library keeps per stream, allocated or grown on the first call with a
 * larger size; glsx_normalize keeps its partials the same way).  Make that
 * first call before capturing the stream into a graph.  The buffers stay
 * allocated until glsx_shutdown; calls on one stream from several host
 * threads are serialised while they enqueue (they share that stream's
 * buffer). */
int glsx_fft2d_mag(glsx_buf src, glsx_buf dst, int flags, int log_mag, float offset, void *stream);
int glsx_fft2d_mag_dev(const void *src, int src_type, float *dst, int rows, int cols, int flags, int log_mag,
                       float offset, void *stream);

/* Extension: glsCam's whole FFT view (glsCam.cpp:199-202) -- dst (CV_32FC1) =
 * normalize(log(|FFT2(src)| + offset), alpha, beta, NORM_MINMAX) -- in three
 * launches: the row pass, the column pass storing the log magnitude with
 * per-workgroup min/max partials, and the affine map.  Same values as
 * glsx_fft2d_mag followed by glsx_normalize. */
int glsx_spectrum_view(glsx_buf src, glsx_buf dst, int flags, float offset, double alpha, double beta, void *stream);
int glsx_spectrum_view_dev(const void *src, int src_type, float *dst, int rows, int cols, int flags, float offset,
                           double alpha, double beta, void *stream);

/* ---- plans: many FFTs / phase correlations in one persistent launch ---------
 * No counterpart in the reference, which issues every radix stage of every
 * FFT as its own draw (glsFft.cpp:562-652) and runs phaseCorrelate as a fixed
 * host-driven call sequence (glsPhaseCorrelate.cpp:178-219).  A plan cuts all
 * row and column sweeps of its transforms into tiles that one persistent
 * kernel executes in dependency order; shapes may differ per entry. */
typedef struct {
    const void *src; /* device pointer: rows x cols CV_32FC2 (or CV_32FC1)     */
    void *dst;       /* device pointer: rows x cols CV_32FC2, may equal src     */
    int rows, cols;  /* powers of two in [1, 4096]                              */
    int src_type;    /* GLSX_CV_32FC2 or GLSX_CV_32FC1                          */
    int flags;       /* GLSX_FFT_*                                              */
} glsx_fft_desc;

typedef struct {
    const float *src1, *src2; /* device pointers, rows x cols CV_32FC1           */
    const float *window;      /* device pointer or NULL                          */
    int rows, cols;           /* powers of two in [2, 4096]                      */
} glsx_poc_desc;
/* A phase-correlation batch runs the chain of glsx_phase_correlate per
 * descriptor (packed spectrum of the pair, row-pair cross-power pass, column
 * pass into the surface, finalize).  The row-pair tiles need cols/8 threads:
 * a configured workgroup below that is rejected. */

typedef struct glsx_plan_s *glsx_plan;

#define GLSX_PLAN_GRAPH 1    /* capture the launch in a hipGraph; execute replays it */
#define GLSX_PLAN_AUTOTUNE 2 /* time candidate launch shapes (result cached in the
                                tuning database: $GLSX_TUNE_DB or glsx_tune.db
                                beside libglsx.so)                                 */

/* Launch shape: workgroup size (256, 512 or 1024), the widest column tile
 * (columns per tile, 1..64) and the execution mode: 0 = one launch per
 * dependency level with one workgroup per tile (kernel boundaries order the
 * levels), 1 = one persistent launch whose workgroups take tiles from a
 * counter and wait on per-pass completion counters, 2 = one launch per
 * dependency level and kernel: every pass runs on the kernel, tiles and
 * workgroup size the plain one-pass-per-launch path would use for it (the
 * other two fields are then unused).  Every mode computes the same bits.
 * NULL = defaults (FFT batches (256, 16, 0), phase-correlation batches
 * (512, 16, 2)) or, with GLSX_PLAN_AUTOTUNE, the tuned shape. */
typedef struct {
    int threads, col_wcap, persistent;
} glsx_plan_config;

typedef struct {
    int threads, col_wcap, persistent, grid, n_passes, n_items, n_levels;
    size_t lds_bytes;
    int uses_graph;
    int tuned; /* 0 no, 1 measured now, 2 from the database */
    double tuned_us;
} glsx_plan_info;

int glsx_plan_fft_batch(const glsx_fft_desc *descs, int count, const glsx_plan_config *config, int options,
                        glsx_plan *out);
/* Workspace a phase-correlation batch needs (spectra, surfaces, partials,
 * results); 0 on invalid descriptors. */
size_t glsx_poc_batch_workspace_bytes(const glsx_poc_desc *descs, int count);
int glsx_plan_poc_batch(const glsx_poc_desc *descs, int count, void *workspace, size_t workspace_bytes,
                        const glsx_plan_config *config, int options, glsx_plan *out);
/* Enqueues the plan on `stream` (asynchronous).  One execution of a plan at a
 * time: two executions in flight on different streams write the same outputs
 * and, in mode 2, share the plan's side streams. */
int glsx_plan_execute(glsx_plan plan, void *stream);
/* copies the batch's results (count entries, descriptor order); syncs `stream` */
int glsx_plan_poc_results(glsx_plan plan, glsx_poc_result *host_out, void *stream);
/* syncs the device; GLSX_ERR_STATE if a bounded dependency wait expired */
int glsx_plan_check(glsx_plan plan);
int glsx_plan_get_info(glsx_plan plan, glsx_plan_info *info);
int glsx_plan_destroy(glsx_plan plan);

/* ---- phase-correlation server: continuous batching under an HBM budget -----
 * No counterpart in the reference (one synchronous gls::phaseCorrelate call at
 * a time, glsPhaseCorrelate.cpp:142-236).  Submitted image pairs are copied to
 * pinned staging; a scheduler thread admits pending requests into the next
 * batch whenever a stream slot is free and the budget has room, and runs each
 * batch as one plan (execution mode: glsx_server_config.persistent).  Results
 * are bit-identical to a request run alone.  A request larger than the whole
 * budget completes with GLSX_ERR_BUDGET. */
typedef struct glsx_server_s *glsx_server;

typedef struct {
    size_t hbm_budget_bytes; /* device memory for images, spectra and tables    */
    int max_batch;           /* most requests per launch                        */
    int max_inflight;        /* concurrent batches (streams), 1..8              */
    int threads;             /* workgroup size of the batch kernels (0 = 512)    */
    int persistent;          /* execution mode of a batch, as glsx_plan_config's:
                                0 per level, 1 one persistent launch, 2 per
                                level and kernel (the plain passes' kernels) */
} glsx_server_config;

typedef struct {
    long long id;
    int status;              /* GLSX_OK or the error of this request            */
    int batch_size;          /* requests in the launch that served it           */
    glsx_poc_result result;
    double latency_s;        /* submit -> completion seen by the scheduler      */
} glsx_server_completion;

typedef struct {
    long long submitted, completed, rejected, batches, pending;
    double mean_batch;
    size_t budget_bytes, peak_bytes;
    int inflight, max_inflight_seen;
} glsx_server_stats;

int glsx_server_create(const glsx_server_config *config, glsx_server *out);
/* img1, img2 (and window or NULL): host pointers, rows x cols floats; copied
 * before the call returns. */
int glsx_server_submit(glsx_server server, const float *img1, const float *img2, const float *window, int rows,
                       int cols, long long *id_out);
/* The same for images already in pinned host memory (glsx_host_alloc): no
 * host copy; the uploads read the caller's buffers, which must stay valid and
 * unchanged until poll returns the request's completion.  Other memory is
 * rejected with GLSX_ERR_ARG. */
int glsx_server_submit_pinned(glsx_server server, const float *img1, const float *img2, const float *window,
                              int rows, int cols, long long *id_out);
/* Up to `max` completions; waits up to timeout_ms when none is ready. */
int glsx_server_poll(glsx_server server, glsx_server_completion *out, int max, int *n_out, int timeout_ms);
int glsx_server_get_stats(glsx_server server, glsx_server_stats *stats);
/* Finishes every submitted request, then frees the server. */
int glsx_server_destroy(glsx_server server);

#ifdef __cplusplus
}
#endif
#endif /* GLSX_H */
