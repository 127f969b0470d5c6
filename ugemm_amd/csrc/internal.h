// Internal declarations shared by the libugemm_hip.so translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/ugemm_hip.h"

namespace ugemm {

// Optional epilogue after C = alpha * AB + beta * C, in this order (the
// post-GEMM loop of gl_convolution_LReLU, sgemm_gl1.h:210-217):
//   C[m, n] += bias[rows ? m : n];  C[m, n] = C > 0 ? C : slope * C  (act == 1)
struct Epilogue {
    const float *bias = nullptr;
    int bias_rows = 1;     // bias indexed by the row m (else by the column n)
    int act = 0;           // 0: none, 1: LeakyReLU
    float slope = 0.1f;
    __host__ __device__ bool active() const { return bias != nullptr || act != 0; }
};

__host__ __device__ inline float apply_epilogue(const Epilogue &e, float v, int m, int n) {
    if (e.bias) v += e.bias[e.bias_rows ? m : n];
    if (e.act == 1) v = v > 0.0f ? v : v * e.slope;
    return v;
}

// Implicit im2col operand (the column matrix of im2col, sgemm_gl1.h:166-190,
// never materialised): element (k, n) of the K x N matrix, K = ich*kh*kw,
// N = hcol*wcol, is in[c*H*W + (ho*stride - pad + i)*W + (wo*stride - pad + j)]
// for k = (c*kh + i)*kw + j, n = ho*wcol + wo, and 0 outside the image.
// ktab[k] = {c*H*W + i*W + j, (i << 16) | j}, built on the device per call.
struct ConvDesc {
    const float *in;
    const int2 *ktab;
    int H, W, wcol, stride, pad;
};

// A problem of a grouped split-form plan: its prepared planes (blocked
// layout, NP planes of Mp x Kp and Np x Kp) and C; and one C tile of it.
// k rows per block of the row-maximum pass over a row-contiguous operand
// (blocks cover 256 columns x this many k).
constexpr int kAmaxRcontSpan = 128;

struct GroupedSplitProblem {
    const void *Ap, *Bp;
    long a_plane, b_plane;
    int M, N, Kp, Mp, Np;
    float alpha, beta;
    float *C;
    long ldc;
    const unsigned *amax_a, *amax_b;   // FP16X3: per-row maxima (Mp, Np entries)
    // Variants that split B in the core (no B planes): B's FP32 source, its
    // leading dimension, the true K, and whether vector loads are allowed.
    const float *Bf;
    long ldb;
    int K, vec_b;
};
struct GroupedSplitTile {
    int problem, m0, n0, pad;
};
// One operand of a grouped split-form plan to prepare: its FP32 source and
// its planes, and the first block of the job in the preparation grid.
// A block takes kPrepSpan consecutive k tiles of one 64-row stripe: 64 x 128
// tiles of a k-contiguous job, 64 x 64 tiles of a row-contiguous one (nbx
// blocks along the rows either way).  kcont == kPrepRowScaled (FP16X3,
// k-contiguous, k_pad <= 8192): a block takes two whole rows, computes their
// maxima (written to amax) and splits them, so the operand needs no
// row-maximum pass.
constexpr int kPrepSpan = 4;
constexpr int kPrepRowScaled = 2;
struct PrepJob {
    const float *X;
    long ld;
    void *out;
    long plane_stride;
    int rows, K, rows_pad, k_pad;
    int kcont, vec;
    int block0, nbx;
    const unsigned *amax;   // FP16X3: the operand's row maxima (row scales), else null
};
// FP16X3 grouped plans: the row-maximum jobs (out = maxima), the job of
// each of their blocks, and the maxima buffer of all problems, zeroed before
// the pass.
struct GroupedScales {
    unsigned *amax;
    long count;
    const PrepJob *jobs;
    const int *block_job;
    long blocks;
};

// One row-major GEMM after major/transpose normalisation:
//   C[m, n] = alpha * sum_k opA(m, k) * opB(k, n) + beta * C[m, n]
//   opA(m, k) = ta ? A[k * lda + m] : A[m * lda + k]
//   opB(k, n) = tb ? B[n * ldb + k] : B[k * ldb + n]
template <typename T>
struct Gemm {
    int M, N, K;
    T alpha, beta;
    const T *A; long lda;
    const T *B; long ldb;
    T *C; long ldc;
    bool ta, tb;
    Epilogue epi{};
    int splits = 0;   // split-K count: 0 = the launcher picks (split_count), 1 = none
};

// ---- status (runtime.hip) ----
int fail(int status, const char *fmt, ...);     // records status + message, returns status
int hip_fail(hipError_t e, const char *what);   // records UGEMM_ERR_HIP
#define UGEMM_HIP_TRY(expr)                                   \
    do {                                                      \
        hipError_t _e = (expr);                               \
        if (_e != hipSuccess) return ::ugemm::hip_fail(_e, #expr); \
    } while (0)

int ensure_device();        // lazily initialises the device; UGEMM_OK or error
int device_cu_count();

// ---- kernel families (one launcher per family) ----
// Variant tables: each family exposes a count, names, and a launcher.
int fp32_variant_count();
const char *fp32_variant_name(int v);
int fp32_pick_variant(const Gemm<float> &g);
int launch_fp32(const Gemm<float> &g, int variant, hipStream_t s);
// FP32 SGEMM on FP32 MFMA (sgemm_mfma32.hip).
int m32_variant_count();
const char *m32_variant_name(int v);
int m32_pick_variant(const Gemm<float> &g);
int launch_m32(const Gemm<float> &g, int variant, hipStream_t s);
int launch_conv_m32(const Gemm<float> &g, const ConvDesc &cd, int variant, hipStream_t s);

int mfma_variant_count(int mode);
const char *mfma_variant_name(int mode, int v);
int mfma_pick_variant(const Gemm<float> &g, int mode);
int launch_mfma(const Gemm<float> &g, int mode, int variant, hipStream_t s);
// Grouped split-form plans: tile shape of a variant, and the launch of
// the preparation (jobs) then the core (one block per tile).
// Grouped split-form variants: the single-GEMM table (planes of A and B
// prepared) plus variants that split B in the core from its FP32 source.
int mfma_grouped_variant_count(int mode);
const char *mfma_grouped_variant_name(int mode, int variant);
// 1 when the grouped variant splits B in the core (then every problem's B
// must have one layout: row-contiguous, 'N', or k-contiguous, 'T').
int mfma_grouped_splits_b(int mode, int variant);
int mfma_grouped_tile(int mode, int variant, int *bm, int *bn);
// b_kcont (variants that split B in the core): B is k-contiguous ('T').
// block_job[b]: the preparation job of block b.
int launch_mfma_grouped(int mode, int variant, int b_kcont, const PrepJob *jobs, const int *block_job,
                        int prep_blocks,
                        const GroupedSplitProblem *probs, const GroupedSplitTile *tiles, int ntiles, hipStream_t s,
                        const GroupedScales *scales = nullptr);
// Convolution in a split form: g.A = weights (g.M x g.K, k-contiguous), g.B
// unused -- the im2col operand of cd is gathered into planes.
int launch_mfma_conv(const Gemm<float> &g, const ConvDesc &cd, int mode, int variant, hipStream_t s);

int fp8_variant_count();
const char *fp8_variant_name(int v);
int launch_fp8(const Gemm<float> &g, int variant, hipStream_t s);

int dgemm_variant_count();
const char *dgemm_variant_name(int v);
int dgemm_pick_variant(const Gemm<double> &g);
int launch_dgemm(const Gemm<double> &g, int variant, hipStream_t s);

// C <- beta * C for the alpha == 0 / K == 0 cases (beta == 0 writes zeros),
// followed by the epilogue when one is given.
int launch_scale_f32(int M, int N, float beta, float *C, long ldc, hipStream_t s, const Epilogue &epi = Epilogue());

// Validates a convolution's shape and returns its output extent.
int conv_shape(int ich, int w, int h, int k, int pad, int stride, int ch, int *hcol, int *wcol);

// Implicit-GEMM convolution (FP32 tile kernel): g.B is unused, opB comes from cd.
int launch_conv_fp32(const Gemm<float> &g, const ConvDesc &cd, int variant, hipStream_t s);
int launch_scale_f64(int M, int N, double beta, double *C, long ldc, hipStream_t s);

// ---- BLAS-1/2 helpers (blas12.hip): device pointers, stream-ordered ----
int launch_saxpy(int N, float alpha, const float *x, int incx, float *y, int incy, hipStream_t s);
int launch_daxpy(int N, double alpha, const double *x, int incx, double *y, int incy, hipStream_t s);
// ---- device diagnostics (diag.hip) ----
// Sustained 16x16x32 MFMA rate (fp16 != 0: f16, else bf16) on operands taken
// from 1024 16-byte fragments at ops (device memory), in TFLOP/s.
int mfma_ceiling(const void *ops, int fp16, int rounds, double *tflops);

int launch_saxpy3(int N, float a, const float *x, const float *y, float *z, hipStream_t s);
int launch_daxpy3(int N, double a, const double *x, const double *y, double *z, hipStream_t s);
int launch_sgemv(char trans, int M, int N, float alpha, const float *A, int lda, const float *x, int incx, float beta,
                 float *y, int incy, hipStream_t s);
int launch_dgemv(char trans, int M, int N, double alpha, const double *A, int lda, const double *x, int incx,
                 double beta, double *y, int incy, hipStream_t s);

// ---- split-K (splitk.hip) ----
// A GEMM whose tile grid leaves CUs idle is split along K: block (tile, z)
// computes alpha = 1, beta = 0 partial products over K columns
// [z*kc, min(K, (z+1)*kc)) into workspace slice z (M x N, row length N), and
// a reduction adds the slices in z order -- a fixed order, so results are
// bitwise reproducible -- then applies alpha, beta and the epilogue.
// split_count: the number of splits for a grid of `tiles` tiles whose kernel
// wants per_cu resident blocks per CU (requested > 0: that many, capped so
// that every split has a K chunk); *kc is set to a multiple of `kalign`.
int split_count(int requested, long tiles, int K, int kalign, int *kc, int per_cu = 1);
int launch_splitk_reduce_f32(int M, int N, int splits, float alpha, const float *ws, float beta, float *C, long ldc,
                             const Epilogue &epi, hipStream_t s);
int launch_splitk_reduce_f64(int M, int N, int splits, double alpha, const double *ws, double beta, double *C,
                             long ldc, hipStream_t s);

// Tile variants of a grouped plan of `mode` (grouped.hip); 0 for a mode
// grouped plans do not provide.
int grouped_variant_count(int mode);

// ---- tuning (tune.hip) ----
// Tuning-table key of grouped plans (variants = grouped_variant_count(mode)).
constexpr int TUNE_KEY_GROUPED = 16;
// ta, tb: layout flags (grouped keys: convolution flag, split form).
int tuned_variant(int mode, int ta, int tb, int M, int N, int K, int *splits = nullptr);   // -1 if none
int forced_variant(int mode);                                         // -1 if none
int forced_splits(int mode);                                          // 0 if none
void tune_record(int mode, int ta, int tb, int M, int N, int K, int variant, float ms, int splits = 0);
// Merges a tuning cache file into the table (the device must be initialised);
// the number of entries taken, or a negative status.
int tune_load_file(const char *path);

// ---- launch timing (runtime.hip) ----
// When enabled (ugemm_hip_timing_enable), the dominant kernel of every call
// (the GEMM core, not the operand-preparation kernels) is bracketed by a pair
// of HIP events recorded on its own stream; ugemm_hip_timing_collect sums them.
void timing_begin(hipStream_t s);
void timing_end(hipStream_t s);

// ---- pinned staging for host-pointer calls (staging.hip) ----
// Pageable host memory <-> device through two pinned slots, chunked so that
// host-side copies (split over a few threads) overlap the DMA.  Stream-ordered
// on s; staged_d2h returns when the data is in dst.  Not thread-safe: callers
// hold the device mutex.
int staged_h2d(void *dst, const void *src, size_t bytes, hipStream_t s);
int staged_d2h(void *dst, const void *src, size_t bytes, hipStream_t s);
void release_staging();

// ---- device scratch (runtime.hip) ----
// Stream-ordered workspace for the operand-preparation kernels: allocated
// from the device's default memory pool on stream s (which keeps freed
// blocks for reuse), released in stream order after the consuming kernel.
int stream_alloc(void **ptr, size_t bytes, hipStream_t s);
int stream_free(void *ptr, hipStream_t s);

}  // namespace ugemm

// Alignment helpers used by launchers to select vectorised variants.
static inline bool ugemm_aligned16(const void *p) { return (((uintptr_t)p) & 15u) == 0; }
