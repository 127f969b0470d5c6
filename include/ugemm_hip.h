/*
 * ugemm_hip.h -- C ABI of libugemm_hip.so, an MI355X (gfx950) GEMM backend
 * for the ugemm C API.
 *
 * Semantics (all entry points):
 *   C <- alpha * op(A) * op(B) + beta * C
 *   major 'R' (row-major) or 'C' (column-major); transa/transb 'N' or 'T'
 *   (any value other than 'N'/'n' means transpose, as in gemm_cpu.h:327-360).
 *   Row-major 'N' A is M x K (lda >= K); 'T' A is stored K x M (lda >= M).
 *   Row-major 'N' B is K x N (ldb >= N); 'T' B is stored N x K (ldb >= K).
 *   C is M x N (ldc >= N).  Column-major swaps the roles of rows and columns
 *   exactly as ugemm.h:259-285 (dgemm_cpu) indexes them.
 *   beta == 0 never reads C (BLAS rule; gemm_cpu.h:119-123, sgemm_ocl2.h:86-88).
 *   alpha == 0 or K == 0 scales C by beta with the correct major indexing.
 *
 * Errors: functions returning int return UGEMM_OK (0) or a negative
 * ugemm_status_t; the void uut-style entry points record the status, which
 * ugemm_hip_last_error() returns (per calling thread).
 *
 * Streams are passed as `void *` (a hipStream_t, or NULL for the null stream)
 * so that this header needs no HIP include.
 */
#ifndef UGEMM_HIP_H
#define UGEMM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    UGEMM_OK = 0,
    UGEMM_ERR_ARG = -1,         /* invalid argument (bad ld, NULL pointer, bad flag) */
    UGEMM_ERR_HIP = -2,         /* a HIP runtime call failed */
    UGEMM_ERR_NOMEM = -3,       /* device or host allocation failed */
    UGEMM_ERR_NODEV = -4,       /* no usable gfx950 device */
    UGEMM_ERR_UNSUPPORTED = -5, /* mode/layout combination not provided */
    UGEMM_ERR_TUNE_FILE = -6,   /* tuning cache could not be read or written */
    UGEMM_ERR_SCRATCH = -7      /* ugemm_hip_sgemm_noalloc_dev: workspace beyond its static scratch */
} ugemm_status_t;

/* Arithmetic forms of the SGEMM path. */
typedef enum {
    UGEMM_MODE_FP32 = 0,      /* plain FP32: vector FMA, FP32 accumulate */
    UGEMM_MODE_BF16X3 = 1,    /* 2-way BF16 split, hi*hi + hi*lo + lo*hi on MFMA, FP32 accumulate */
    UGEMM_MODE_BF16X1 = 2,    /* BF16 round-to-nearest operands, one MFMA product (opt-in, low precision) */
    UGEMM_MODE_FP8_BLOCK = 3, /* FP8-e4m3 operands with one FP32 scale per 128x128 block (opt-in) */
    UGEMM_MODE_BF16X6 = 4,    /* 3-way BF16 split, six MFMA products: FP32-level error at large K */
    UGEMM_MODE_FP16X3 = 5,    /* 2-way FP16 split with a power-of-two scale per row of op(A) and column
                                 of op(B), hi*hi + hi*lo + lo*hi on MFMA: FP32-level error, 3 products */
    UGEMM_MODE_FP32_MFMA = 6, /* FP32 operands and accumulation on the FP32 matrix cores: FP32 error class,
                                 no operand preparation (no counterpart in the reference) */
    UGEMM_MODE_COUNT = 7
} ugemm_mode_t;

/* ------------------------------------------------------------------ */
/* 1. uut-compatible entry points: host pointers, synchronous.         */
/*    Replace the function pointer handed to test_sgemm/test_dgemm     */
/*    (check_sgemm.c:87-103, check_dgemm.c:86-102); same argument list  */
/*    as sgemm_cpu/dgemm_cpu (ugemm.h:162-176, ugemm.h:287-301).       */
/*    sgemm_hip runs UGEMM_MODE_FP32 unless UGEMM_HIP_MODE is set      */
/*    (fp32|bf16x3|bf16x1|fp8|bf16x6|fp16x3|fp32_mfma).                */
/* ------------------------------------------------------------------ */
void sgemm_hip(char major, char transa, char transb, int M, int N, int K,
               float alpha, const float *A, int lda, const float *B, int ldb,
               float beta, float *C, int ldc);
void dgemm_hip(char major, char transa, char transb, int M, int N, int K,
               double alpha, const double *A, int lda, const double *B, int ldb,
               double beta, double *C, int ldc);

/* Host-pointer SGEMM with an explicit arithmetic mode. */
int ugemm_hip_sgemm(char major, char transa, char transb, int M, int N, int K,
                    float alpha, const float *A, int lda, const float *B, int ldb,
                    float beta, float *C, int ldc, int mode);

/* ------------------------------------------------------------------ */
/* 2. GPU-backend trio, as sgemm_ocl_init / sgemm_ocl / sgemm_ocl_finish */
/*    (sgemm_ocl2.h:158-224, selected by sgemm_test.c:29-33).          */
/* ------------------------------------------------------------------ */
/* Select the device (-1: $UGEMM_HIP_DEVICE or 0) and reserve max_bytes of
 * device staging memory for the host-pointer entry points (0: grow on demand).
 * Calling any entry point without init initialises device 0 lazily. */
int ugemm_hip_init(int device, size_t max_bytes);
void ugemm_hip_finish(void);
/* Packed row-major SGEMM (lda = K or M, ldb = N or K, ldc = N by transposition),
 * the argument list of sgemm_ocl (sgemm_ocl2.h:166). */
void sgemm_hip_packed(char ta, char tb, int m, int n, int k, float alpha,
                      float *a, float *b, float beta, float *c);

/* ------------------------------------------------------------------ */
/* 3. Device-pointer, stream-ordered entry points (no host sync).      */
/* ------------------------------------------------------------------ */
int ugemm_hip_sgemm_dev(char major, char transa, char transb, int M, int N, int K,
                        float alpha, const float *dA, int lda, const float *dB, int ldb,
                        float beta, float *dC, int ldc, int mode, void *stream);
int ugemm_hip_dgemm_dev(char major, char transa, char transb, int M, int N, int K,
                        double alpha, const double *dA, int lda, const double *dB, int ldb,
                        double beta, double *dC, int ldc, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): ugemm_hip_sgemm_dev with every workspace (split-form planes,
 * FP16X3 row maxima, FP8 planes and block exponents, split-K partial
 * products) taken from one static device array of
 * UGEMM_HIP_NOALLOC_SCRATCH_BYTES per GPU, held from library load to process
 * exit.  It calls no hipMalloc/hipFree.  A call whose workspace would exceed
 * the array returns UGEMM_ERR_SCRATCH before it launches anything.  The
 * array keeps the last call's contents (what must start at zero is cleared
 * by the call), and every call on one GPU shares it: calls that overlap on
 * one GPU (other streams, other host threads) race and must be serialised
 * by the caller.  The plain FP32 and FP32-MFMA forms need no workspace
 * unless split along K, so their size is not limited by the array. */
#define UGEMM_HIP_NOALLOC_SCRATCH_BYTES (64u << 20)
int ugemm_hip_sgemm_noalloc_dev(char major, char transa, char transb, int M, int N, int K,
                                float alpha, const float *dA, int lda, const float *dB, int ldb,
                                float beta, float *dC, int ldc, int mode, void *stream);

/* Bias + activation epilogue, applied after alpha/beta in this order:
 *   C[m][n] += bias[m] (UGEMM_BIAS_ROWS) or bias[n] (UGEMM_BIAS_COLS);
 *   C = C > 0 ? C : slope * C (UGEMM_ACT_LEAKY_RELU)
 * -- the post-GEMM loop of gl_convolution_LReLU (sgemm_gl1.h:210-217), fused
 * into the GEMM's store.  m, n index the caller's C (either major).  Modes:
 * FP32 and the split forms (FP8 returns UGEMM_ERR_UNSUPPORTED).  d_bias may
 * be NULL (activation only). */
enum { UGEMM_BIAS_ROWS = 0, UGEMM_BIAS_COLS = 1 };
enum { UGEMM_ACT_NONE = 0, UGEMM_ACT_LEAKY_RELU = 1 };
int ugemm_hip_sgemm_ex_dev(char major, char transa, char transb, int M, int N, int K,
                           float alpha, const float *dA, int lda, const float *dB, int ldb,
                           float beta, float *dC, int ldc, int mode,
                           const float *d_bias, int bias_dim, int act, float slope, void *stream);

/* Convolution as one implicit GEMM (the column matrix of im2col,
 * sgemm_gl1.h:166-190, is gathered inside the GEMM's tile loader and never
 * stored), with the bias + LeakyReLU epilogue:
 *   outputs[c][p] = act(sum_q weights[c][q] * col[q][p] + bias[c])
 *   inputs [ich][h][w], weights [ch][ich*k*k], outputs [ch][hcol*wcol],
 *   hcol = (h + 2*pad - k)/stride + 1, wcol likewise.
 * convolution_LReLU_hip has the argument list of gl_convolution_LReLU
 * (sgemm_gl1.h:191-218) and ocl_convolution_LReLU (sgemm_ocl1.h:302-337):
 * host pointers, synchronous, slope 0.1; status in ugemm_hip_last_error().
 * ugemm_hip_conv2d_dev: device pointers, stream-ordered, act/slope/bias
 * selectable (d_bias may be NULL); mode UGEMM_MODE_FP32 or
 * UGEMM_MODE_FP32_MFMA (im2col gathered in the tile loader) or a split form
 * (BF16X1/X3/X6, FP16X3: the im2col operand is gathered into its planes);
 * FP8 is refused (UNSUPPORTED). */
void convolution_LReLU_hip(float *inputs, int ich, int w, int h, float *weights, int k, int pad, int stride,
                           float *outputs, int ch, float *bias);
int ugemm_hip_conv2d_dev(const float *d_in, int ich, int w, int h, const float *d_weights, int k, int pad,
                         int stride, float *d_out, int ch, const float *d_bias, int act, float slope, int mode,
                         void *stream);

/* ------------------------------------------------------------------ */
/* 4. Grouped (batched, uneven) SGEMM: one persistent launch.          */
/*    Every problem is row-major; pointers are device pointers.        */
/* ------------------------------------------------------------------ */
typedef struct {
    char transa, transb;
    int M, N, K;
    float alpha, beta;
    const float *A; int lda;
    const float *B; int ldb;
    float *C; int ldc;
} ugemm_problem_t;

typedef struct ugemm_grouped_plan ugemm_grouped_plan_t;

/* Build a plan: validates the problems, builds the tile table and the
 * per-XCD tile queues, uploads them, and (capture != 0) records the launch
 * in a hipGraph so that ugemm_hip_grouped_run replays one graph launch.
 * mode: UGEMM_MODE_FP32, UGEMM_MODE_FP32_MFMA (the same persistent kernel
 * with FP32 matrix-core tiles), or a split form (UGEMM_MODE_BF16X1/X3/X6,
 * UGEMM_MODE_FP16X3: the plan owns every problem's prepared planes; a run is
 * a preparation launch and a core launch, FP16X3 adding a row-maximum pass);
 * FP8 returns UGEMM_ERR_UNSUPPORTED. */
int ugemm_hip_grouped_plan_create(ugemm_grouped_plan_t **plan, int count,
                                  const ugemm_problem_t *problems, int mode,
                                  int capture, void *stream);
int ugemm_hip_grouped_run(ugemm_grouped_plan_t *plan, void *stream);
void ugemm_hip_grouped_plan_destroy(ugemm_grouped_plan_t *plan);
/* A plan of convolutions (each the operation of ugemm_hip_conv2d_dev:
 * implicit im2col, bias, activation), run by the same persistent kernel in
 * one launch -- the layers of a network, or the images of a batch.  The
 * returned plan is run and destroyed with the functions above.  Buffers are
 * device pointers and must stay valid while the plan is used. */
typedef struct {
    const float *in; int ich, w, h;              /* inputs  [ich][h][w] */
    const float *weights; int k, pad, stride;    /* weights [ch][ich*k*k] */
    float *out; int ch;                          /* outputs [ch][hcol*wcol] */
    const float *bias; int act; float slope;     /* bias [ch] or NULL; UGEMM_ACT_* */
} ugemm_conv_problem_t;
int ugemm_hip_conv_grouped_plan_create(ugemm_grouped_plan_t **plan, int count,
                                       const ugemm_conv_problem_t *convs, int capture, void *stream);
/* The same with an arithmetic mode: UGEMM_MODE_FP32 (as above) or
 * UGEMM_MODE_FP32_MFMA (the tiles on the FP32 matrix cores); any other mode
 * is UGEMM_ERR_UNSUPPORTED. */
int ugemm_hip_conv_grouped_plan_create_mode(ugemm_grouped_plan_t **plan, int count,
                                            const ugemm_conv_problem_t *convs, int mode, int capture,
                                            void *stream);
/* One-shot convenience: plan, run (no capture), destroy. */
int ugemm_hip_sgemm_grouped(int count, const ugemm_problem_t *problems, int mode,
                            void *stream);
/* Tile variants of the grouped kernel (no counterpart in the reference):
 * 0-3 one tile shape for every problem, 4 a tile shape per problem (128 x
 * 128 or 64 x 128 by the padding of its rows).  ugemm_hip_tune_grouped
 * times them; ugemm_hip_force_variant(UGEMM_MODE_FP32, v) with v < the
 * count pins one for the plans created next on the thread. */
int ugemm_hip_grouped_variant_count(void);
const char *ugemm_hip_grouped_variant_name(int variant);
/* The same per plan mode: UGEMM_MODE_FP32 (the variants above),
 * UGEMM_MODE_FP32_MFMA (the same five tile choices on the FP32 matrix
 * cores) and the split forms (their split-core variants, as
 * ugemm_hip_variant_name, then three grouped-only variants, names ending
 * in "_split_b", that split B in the core from its FP32 source: plans whose
 * every B has one layout; UGEMM_ERR_UNSUPPORTED otherwise).  Count:
 * UGEMM_ERR_ARG for a mode grouped plans do not provide; name: NULL out of
 * range. */
int ugemm_hip_grouped_mode_variant_count(int mode);
const char *ugemm_hip_grouped_mode_variant_name(int mode, int variant);

/* ------------------------------------------------------------------ */
/* 5. BLAS-1/2 helpers (HBM-bound kernels).                           */
/*    saxpy/daxpy: y[n*incy] += alpha * x[n*incx], n < N               */
/*      (saxpy_cpu / daxpy_cpu, ugemm.h:76-101).                       */
/*    sgemv/dgemv, the element mapping of sgemv_cpu / dgemv_cpu        */
/*    (ugemm.h:103-160), m < M, n < N:                                  */
/*      'N': y[m*incy] = alpha*sum_n A[m + n*lda]*x[n*incx] + beta*y   */
/*      other: y[m*incy] = alpha*sum_n A[n + m*lda]*x[n*incx] + beta*y */
/*    x is read with incx (the reference reads it with incy,            */
/*    ugemm.h:116,147 -- identical when incx == incy).  Increments      */
/*    must be >= 1; beta == 0 does not read y.                          */
/*    The *_hip functions have the reference's argument lists (host     */
/*    pointers, synchronous, status in ugemm_hip_last_error()); the     */
/*    _dev functions take device pointers and a stream.                 */
/* ------------------------------------------------------------------ */
void saxpy_hip(int N, float alpha, const float *x, int incx, float *y, int incy);
void daxpy_hip(int N, double alpha, const double *x, int incx, double *y, int incy);
/* Three-operand axpy with the argument lists of saxpy_avx / daxpy_avx
 * (ugemm.h:41-74): z[i] = a * x[i] + y[i] with one rounding (the FMA branch
 * the reference's -march=native build takes), i < N, z stored with
 * non-temporal stores.  Unlike the reference, which works in whole vectors
 * of 8 (4) and so also writes z up to the next multiple and needs 32-byte
 * aligned arrays, exactly N elements are written, at any alignment; z may
 * be y.  Host pointers, synchronous, status in ugemm_hip_last_error(); the
 * _dev versions take device pointers and a stream. */
void saxpy_avx_hip(int N, float a, float *x, float *y, float *z);
void daxpy_avx_hip(int N, double a, double *x, double *y, double *z);
int ugemm_hip_saxpy3_dev(int N, float a, const float *dx, const float *dy, float *dz, void *stream);
int ugemm_hip_daxpy3_dev(int N, double a, const double *dx, const double *dy, double *dz, void *stream);
void sgemv_hip(char trans, int M, int N, float alpha, const float *A, int lda, const float *x, int incx,
               float beta, float *y, int incy);
void dgemv_hip(char trans, int M, int N, double alpha, const double *A, int lda, const double *x, int incx,
               double beta, double *y, int incy);
int ugemm_hip_saxpy_dev(int N, float alpha, const float *dx, int incx, float *dy, int incy, void *stream);
int ugemm_hip_daxpy_dev(int N, double alpha, const double *dx, int incx, double *dy, int incy, void *stream);
int ugemm_hip_sgemv_dev(char trans, int M, int N, float alpha, const float *dA, int lda, const float *dx, int incx,
                        float beta, float *dy, int incy, void *stream);
int ugemm_hip_dgemv_dev(char trans, int M, int N, double alpha, const double *dA, int lda, const double *dx,
                        int incx, double beta, double *dy, int incy, void *stream);

/* ------------------------------------------------------------------ */
/* 6. Status and tuning.                                              */
/* ------------------------------------------------------------------ */
int ugemm_hip_last_error(void);
const char *ugemm_hip_last_error_message(void);
const char *ugemm_hip_status_string(int status);
/* Time every compiled kernel variant on an (M, N, K, layout, mode) problem
 * -- and split-K counts when 128x128 tiles would leave CUs idle -- and
 * remember the fastest for that shape bucket. Allocates its own buffers.
 * The tuning replaces nothing in the reference (it has no tuner for this
 * path); the cache file format is documented in tune.hip. */
int ugemm_hip_tune_sgemm(char transa, char transb, int M, int N, int K, int mode);
/* Time every tile variant of a grouped plan (each run a few times) and
 * rebuild the plan with the fastest; returns its index.  It first waits for
 * all work on the device (the caller's uploads of A, B and C), runs on a
 * private stream, and restores every problem's C (or output) afterwards, so
 * C is left as the caller left it; it briefly allocates a copy of all C. */
int ugemm_hip_tune_grouped(ugemm_grouped_plan_t *plan);
/* Tuning cache file.  ugemm_hip_tune_load merges a file written by
 * ugemm_hip_tune_save (entries of another architecture or CU count are
 * skipped) and returns the number of entries taken.  When the environment
 * variable UGEMM_HIP_TUNE_CACHE names a file, it is loaded at initialisation
 * (ugemm_hip_init or the first call), so that uut callers of sgemm_hip pick
 * up a saved table; a missing or unreadable file is ignored there. */
int ugemm_hip_tune_load(const char *path);
int ugemm_hip_tune_save(const char *path);
/* Variant chosen for a problem (index into the mode's kernel table); -1 = none. */
int ugemm_hip_tune_lookup(char transa, char transb, int M, int N, int K, int mode);
/* Force a variant for the next calls on this thread (-1 restores tuning/heuristics). */
int ugemm_hip_force_variant(int mode, int variant);
/* Force the split-K count (1..64; 1 = no split) of the next calls on this
 * thread, mode -1 for DGEMM; 0 restores tuning/heuristics. A split call
 * writes alpha = 1 partial products of K chunks to a workspace and adds
 * them in a fixed order (bitwise reproducible). No counterpart in the
 * reference; FP8 and grouped plans do not split. */
int ugemm_hip_force_splits(int mode, int splits);
int ugemm_hip_variant_count(int mode);
const char *ugemm_hip_variant_name(int mode, int variant);
const char *ugemm_hip_version(void);

/* Launch timing for benchmarks (no counterpart in the reference): while
 * enabled on the calling thread, the GEMM core kernel of every call (not the
 * operand-preparation kernels) is bracketed by HIP events on its own stream.
 * ugemm_hip_timing_collect waits for them, stores the summed milliseconds and
 * returns how many launches were timed (then starts a new collection). */
int ugemm_hip_timing_enable(int on);
int ugemm_hip_timing_collect(double *total_ms);

/* Sustained matrix-core rate of the device (diagnostic, no counterpart in
 * the reference): every SIMD issues nothing but independent
 * v_mfma_f32_16x16x32 instructions (f16 when fp16 != 0, else bf16) for
 * `rounds` rounds of 16, on operands taken from d_operands, 1024 16-byte
 * fragments (8192 16-bit values) in device memory -- planes of real data,
 * since the rate under the power limit depends on the bits multiplied.
 * *tflops receives the rate; it synchronizes the device.  It bounds what
 * a GEMM core built on these MFMAs can reach on this device; the published
 * peak assumes the top clock. */
int ugemm_hip_mfma_ceiling(const void *d_operands, int fp16, int rounds, double *tflops);

#ifdef __cplusplus
}
#endif

#endif /* UGEMM_HIP_H */
