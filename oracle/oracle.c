/*
 * oracle.c -- CPU oracle for the GEMM path.  TEST INFRASTRUCTURE ONLY: called
 * by tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of
 * bench.py, and never by the product library.
 *
 * What it restates (reference = yui0/ugemm):
 *   oracle_sgemm_naive / oracle_dgemm_naive
 *       the naive loops sgemm_cpu (ugemm.h:287-411) and dgemm_cpu
 *       (ugemm.h:162-286): for every (m, n) a sequential sum over k in the
 *       working precision, then C = alpha*sum + beta*C (C is read even when
 *       beta == 0, ugemm.h:313).  The 8 index formulas of ugemm.h:185-284 and
 *       :310-410 are expressed through element accessors.
 *   oracle_sgemm_f64 / oracle_dgemm_dd
 *       the same mathematics evaluated (far) more accurately: float inputs in
 *       FP64, double inputs in double-double (TwoProd + TwoSum), plus the
 *       elementwise error scale |alpha| (|A||B|)_ij + |beta| |C0_ij| that the
 *       acceptance bound 2*K*eps*scale uses.  Optional row subset for large
 *       shapes.
 *   oracle_sgemm_packed / oracle_dgemm_packed
 *       the GotoBLAS blocking of gemm_cpu.h:238-282 (MC = KC = 384, NC = 4096,
 *       MR = NR = 4; beta applied with the first KC block only,
 *       gemm_cpu.h:269), with per-call buffers instead of the static ones
 *       (gemm_cpu.h:9-11) and BLAS alpha == 0 semantics (row-major indexing,
 *       where gemm_cpu.h:308-323 always indexes column-major).
 *
 *   oracle_im2col / oracle_conv_lrelu
 *       im2col (sgemm_gl1.h:166-190) and the convolution + bias + LeakyReLU
 *       of gl_convolution_LReLU (sgemm_gl1.h:191-218) in double precision,
 *       with the elementwise error scale (end of this file).
 *
 * Pins: tests/test_oracle.py checks these functions against the known-answer
 * product of sgemm_test.c:186-200, exact small-integer products, numpy's FP64
 * matmul, brute force on tiny shapes, and the golden vectors produced by the
 * compiled reference (tests/golden/).
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* Element (m, k) of op(A) and (k, n) of op(B) for the two majors.
 * Row-major: 'N' A is M x K at A[m*lda + k], 'T' A at A[k*lda + m];
 *            'N' B is K x N at B[k*ldb + n], 'T' B at B[n*ldb + k];
 *            C at C[m*ldc + n]                     (ugemm.h:310,323,336,349).
 * Column-major: the same with the two indices of every operand swapped
 *            (ugemm.h:363,376,389,402). */
static inline long idx_a(int row, int ta, long lda, int m, int k)
{
	if (row) return ta ? (long)k * lda + m : (long)m * lda + k;
	return ta ? (long)m * lda + k : (long)k * lda + m;
}
static inline long idx_b(int row, int tb, long ldb, int k, int n)
{
	if (row) return tb ? (long)n * ldb + k : (long)k * ldb + n;
	return tb ? (long)k * ldb + n : (long)n * ldb + k;
}
static inline long idx_c(int row, long ldc, int m, int n)
{
	return row ? (long)m * ldc + n : (long)n * ldc + m;
}
static int is_t(char t) { return t == 'T'; }

/* Returns 0 when the flags are ones the reference's naive loops act on
 * ('R' or anything else meaning column-major, 'N' / 'T' transposes). */
static int flags_ok(char ta, char tb)
{
	return (ta == 'N' || ta == 'T') && (tb == 'N' || tb == 'T');
}

void oracle_sgemm_naive(char major, char ta, char tb, int M, int N, int K, float alpha,
                        const float *A, int lda, const float *B, int ldb, float beta, float *C, int ldc)
{
	if (!flags_ok(ta, tb)) return;
	const int row = major == 'R', a_t = is_t(ta), b_t = is_t(tb);
	for (int m = 0; m < M; m++) {
		for (int n = 0; n < N; n++) {
			float sum = 0.0f;
			for (int k = 0; k < K; k++) {
				float p = A[idx_a(row, a_t, lda, m, k)] * B[idx_b(row, b_t, ldb, k, n)];
				sum += p;
			}
			float *c = &C[idx_c(row, ldc, m, n)];
			*c = alpha * sum + beta * *c;
		}
	}
}

void oracle_dgemm_naive(char major, char ta, char tb, int M, int N, int K, double alpha,
                        const double *A, int lda, const double *B, int ldb, double beta, double *C, int ldc)
{
	if (!flags_ok(ta, tb)) return;
	const int row = major == 'R', a_t = is_t(ta), b_t = is_t(tb);
	for (int m = 0; m < M; m++) {
		for (int n = 0; n < N; n++) {
			double sum = 0.0;
			for (int k = 0; k < K; k++) {
				double p = A[idx_a(row, a_t, lda, m, k)] * B[idx_b(row, b_t, ldb, k, n)];
				sum += p;
			}
			double *c = &C[idx_c(row, ldc, m, n)];
			*c = alpha * sum + beta * *c;
		}
	}
}

/* ------------------------------------------------------------------ */
/* Accurate references, threaded over rows of C.                        */
/* ------------------------------------------------------------------ */
typedef struct {
	int row, a_t, b_t, M, N, K;
	int dbl;                 /* 0: float operands (FP64 sums); 1: double operands (double-double) */
	double alpha, beta;
	const void *A, *B, *C0;
	long lda, ldb, ldc;
	const int *rows;         /* NULL: all M rows; else nrows row indices */
	int nrows;
	double *out, *scale;     /* nrows x N, dense */
	int t, nt;
} ref_job;

/* Error-free transforms (Knuth TwoSum, FMA-based TwoProd). */
static inline void two_sum(double a, double b, double *s, double *e)
{
	double x = a + b, z = x - a;
	*s = x;
	*e = (a - (x - z)) + (b - z);
}

static void *ref_worker(void *arg)
{
	ref_job *j = (ref_job *)arg;
	const float *Af = j->A, *Bf = j->B, *Cf = j->C0;
	const double *Ad = j->A, *Bd = j->B, *Cd = j->C0;
	for (int r = j->t; r < j->nrows; r += j->nt) {
		const int m = j->rows ? j->rows[r] : r;
		for (int n = 0; n < j->N; n++) {
			double hi = 0.0, lo = 0.0, abs_sum = 0.0;
			for (int k = 0; k < j->K; k++) {
				long ia = idx_a(j->row, j->a_t, j->lda, m, k), ib = idx_b(j->row, j->b_t, j->ldb, k, n);
				if (!j->dbl) {
					/* float x float is exact in double; the sum is compensated too. */
					double p = (double)Af[ia] * (double)Bf[ib], s, e;
					two_sum(hi, p, &s, &e);
					hi = s;
					lo += e;
					abs_sum += fabs(p);
				} else {
					double a = Ad[ia], b = Bd[ib];
					double p = a * b, pe = fma(a, b, -p), s, e;
					two_sum(hi, p, &s, &e);
					hi = s;
					lo += e + pe;
					abs_sum += fabs(p);
				}
			}
			double c0 = 0.0;
			long ic = idx_c(j->row, j->ldc, m, n);
			if (j->beta != 0.0) c0 = j->dbl ? Cd[ic] : (double)Cf[ic];
			/* alpha*(hi+lo) + beta*c0 with the low part kept until the end. */
			double v = fma(j->alpha, hi, fma(j->alpha, lo, j->beta * c0));
			j->out[(long)r * j->N + n] = v;
			if (j->scale) j->scale[(long)r * j->N + n] = fabs(j->alpha) * abs_sum + fabs(j->beta) * fabs(c0);
		}
	}
	return NULL;
}

static void run_ref(ref_job *proto, int nthreads)
{
	if (nthreads <= 0) {
		long n = sysconf(_SC_NPROCESSORS_ONLN);
		nthreads = n > 0 ? (int)n : 1;
		if (nthreads > 16) nthreads = 16;
	}
	if (nthreads > proto->nrows) nthreads = proto->nrows > 0 ? proto->nrows : 1;
	pthread_t th[64];
	ref_job jobs[64];
	if (nthreads > 64) nthreads = 64;
	for (int t = 0; t < nthreads; t++) {
		jobs[t] = *proto;
		jobs[t].t = t;
		jobs[t].nt = nthreads;
		if (t) pthread_create(&th[t], NULL, ref_worker, &jobs[t]);
	}
	ref_worker(&jobs[0]);
	for (int t = 1; t < nthreads; t++) pthread_join(th[t], NULL);
}

/* out[r, n] = alpha * sum_k op(A)[m_r, k] op(B)[k, n] + beta * C0[m_r, n] in FP64
 * (compensated), scale[r, n] = |alpha| sum_k |op(A) op(B)| + |beta| |C0|.
 * rows == NULL means every row (nrows must then be M).  beta == 0 ignores C0.
 * Flags: major 'R' or 'C'; trans 'N' or 'T'.  Returns -1 on bad flags. */
int oracle_sgemm_f64(char major, char ta, char tb, int M, int N, int K, double alpha,
                     const float *A, int lda, const float *B, int ldb, double beta, const float *C0, int ldc,
                     const int *rows, int nrows, double *out, double *scale, int nthreads)
{
	if (!flags_ok(ta, tb)) return -1;
	ref_job j = {major == 'R', is_t(ta), is_t(tb), M, N, K, 0, alpha, beta, A, B, C0,
	             lda, ldb, ldc, rows, rows ? nrows : M, out, scale, 0, 1};
	run_ref(&j, nthreads);
	return 0;
}

int oracle_dgemm_dd(char major, char ta, char tb, int M, int N, int K, double alpha,
                    const double *A, int lda, const double *B, int ldb, double beta, const double *C0, int ldc,
                    const int *rows, int nrows, double *out, double *scale, int nthreads)
{
	if (!flags_ok(ta, tb)) return -1;
	ref_job j = {major == 'R', is_t(ta), is_t(tb), M, N, K, 1, alpha, beta, A, B, C0,
	             lda, ldb, ldc, rows, rows ? nrows : M, out, scale, 0, 1};
	run_ref(&j, nthreads);
	return 0;
}

/* ------------------------------------------------------------------ */
/* GotoBLAS-blocked GEMM with the blocking of gemm_cpu.h (generic over   */
/* the element type through a macro).                                    */
/* ------------------------------------------------------------------ */
#define P_MC 384
#define P_KC 384
#define P_NC 4096
#define P_MR 4
#define P_NR 4

#define DEFINE_PACKED(T, NAME)                                                                         \
static void NAME##_pack_a(int mc, int kc, const T *A, long rs, long cs, T *buf)                        \
{                                                                                                     \
	/* MR-row slivers, k-major inside a sliver, zero-padded below mc. */                              \
	for (int i0 = 0; i0 < mc; i0 += P_MR)                                                             \
		for (int k = 0; k < kc; k++)                                                                 \
			for (int i = 0; i < P_MR; i++)                                                           \
				*buf++ = (i0 + i < mc) ? A[(i0 + i) * rs + k * cs] : (T)0;                           \
}                                                                                                     \
static void NAME##_pack_b(int kc, int nc, const T *B, long rs, long cs, T *buf)                        \
{                                                                                                     \
	for (int j0 = 0; j0 < nc; j0 += P_NR)                                                             \
		for (int k = 0; k < kc; k++)                                                                 \
			for (int j = 0; j < P_NR; j++)                                                           \
				*buf++ = (j0 + j < nc) ? B[k * rs + (j0 + j) * cs] : (T)0;                           \
}                                                                                                     \
/* C[mr x nr] = beta*C + alpha*(A sliver x B sliver); beta == 0 does not read C. */                   \
static void NAME##_micro(int kc, int mr, int nr, T alpha, const T *a, const T *b, T beta,              \
                         T *C, long rs, long cs)                                                       \
{                                                                                                     \
	T ab[P_MR * P_NR] = {0};                                                                          \
	for (int k = 0; k < kc; k++, a += P_MR, b += P_NR)                                                \
		for (int j = 0; j < P_NR; j++)                                                               \
			for (int i = 0; i < P_MR; i++) ab[i + j * P_MR] += a[i] * b[j];                          \
	for (int j = 0; j < nr; j++)                                                                      \
		for (int i = 0; i < mr; i++) {                                                               \
			T *c = &C[i * rs + j * cs];                                                              \
			T v = (beta == (T)0) ? (T)0 : (beta == (T)1 ? *c : beta * *c);                           \
			*c = v + alpha * ab[i + j * P_MR];                                                       \
		}                                                                                            \
}                                                                                                     \
static void NAME##_nn(int m, int n, int k, T alpha, const T *A, long ars, long acs, const T *B,        \
                      long brs, long bcs, T beta, T *C, long crs, long ccs)                            \
{                                                                                                     \
	T *pa = malloc(sizeof(T) * P_MC * P_KC), *pb = malloc(sizeof(T) * P_KC * P_NC);                   \
	if (!pa || !pb) { free(pa); free(pb); return; }                                                    \
	for (int j0 = 0; j0 < n; j0 += P_NC) {                                                            \
		int nc = n - j0 < P_NC ? n - j0 : P_NC;                                                      \
		for (int l0 = 0; l0 < k; l0 += P_KC) {                                                       \
			int kc = k - l0 < P_KC ? k - l0 : P_KC;                                                  \
			T bb = (l0 == 0) ? beta : (T)1;                                                          \
			NAME##_pack_b(kc, nc, B + l0 * brs + j0 * bcs, brs, bcs, pb);                            \
			for (int i0 = 0; i0 < m; i0 += P_MC) {                                                   \
				int mc = m - i0 < P_MC ? m - i0 : P_MC;                                              \
				NAME##_pack_a(mc, kc, A + i0 * ars + l0 * acs, ars, acs, pa);                        \
				for (int jj = 0; jj < nc; jj += P_NR)                                                \
					for (int ii = 0; ii < mc; ii += P_MR)                                            \
						NAME##_micro(kc, mc - ii < P_MR ? mc - ii : P_MR,                            \
						             nc - jj < P_NR ? nc - jj : P_NR, alpha, pa + ii * kc,           \
						             pb + jj * kc, bb, C + (i0 + ii) * crs + (j0 + jj) * ccs, crs, ccs); \
			}                                                                                        \
		}                                                                                            \
	}                                                                                                 \
	free(pa);                                                                                         \
	free(pb);                                                                                         \
}                                                                                                     \
void NAME(char major, char ta, char tb, int M, int N, int K, T alpha, const T *A, int lda,             \
          const T *B, int ldb, T beta, T *C, int ldc)                                                  \
{                                                                                                     \
	/* Quick returns as gemm_cpu.h:303-305. */                                                        \
	if (M == 0 || N == 0 || ((alpha == (T)0 || K == 0) && beta == (T)1)) return;                      \
	const int row = major == 'R', a_t = ta != 'N', b_t = tb != 'N';                                   \
	/* Strides (row step, column step) of op(A) (M x K), op(B) (K x N), C (M x N). */                  \
	long ars, acs, brs, bcs, crs, ccs;                                                                \
	if (row) { crs = ldc; ccs = 1; } else { crs = 1; ccs = ldc; }                                     \
	if (row != a_t) { ars = lda; acs = 1; } else { ars = 1; acs = lda; }                              \
	if (row != b_t) { brs = ldb; bcs = 1; } else { brs = 1; bcs = ldb; }                              \
	if (alpha == (T)0 || K == 0) {                                                                    \
		for (int i = 0; i < M; i++)                                                                  \
			for (int j = 0; j < N; j++) {                                                            \
				T *c = &C[i * crs + j * ccs];                                                        \
				*c = (beta == (T)0) ? (T)0 : beta * *c;                                              \
			}                                                                                        \
		return;                                                                                      \
	}                                                                                                 \
	NAME##_nn(M, N, K, alpha, A, ars, acs, B, brs, bcs, beta, C, crs, ccs);                            \
}

DEFINE_PACKED(float, oracle_sgemm_packed)
DEFINE_PACKED(double, oracle_dgemm_packed)

/* ------------------------------------------------------------------------
 * Convolution (the reference's CNN path).
 *
 * oracle_im2col restates im2col of sgemm_gl1.h:166-190: channels_col =
 * channels * kernel_h * kernel_w rows; row c takes image channel
 * c / kernel_h / kernel_w at offset (h_offset, w_offset) = ((c / kernel_w) %
 * kernel_h, c % kernel_w); entries outside the image are 0.
 *
 * oracle_conv_lrelu restates gl_convolution_LReLU (sgemm_gl1.h:191-218) in
 * double precision: im2col, the row-major NN product weights (ch x K) times
 * col (K x N), then out += bias[i] and out = out > 0 ? out : slope * out
 * (act != 0; the reference's slope is 0.1).  scale (may be NULL) receives
 * sum_q |w[i][q]| |col[q][p]| + |bias[i]|, the elementwise error scale of a
 * single-precision evaluation.  Returns 0, or -1 when a buffer cannot be
 * allocated.
 * ------------------------------------------------------------------------ */
void oracle_im2col(const float *im, int channels, int height, int width, int kernel_h, int kernel_w, int pad_h,
                   int pad_w, int stride_h, int stride_w, float *col)
{
	int height_col = (height + 2 * pad_h - kernel_h) / stride_h + 1;
	int width_col = (width + 2 * pad_w - kernel_w) / stride_w + 1;
	int channels_col = channels * kernel_h * kernel_w;
	for (int c = 0; c < channels_col; c++) {
		int w_offset = c % kernel_w;
		int h_offset = (c / kernel_w) % kernel_h;
		int c_im = c / kernel_h / kernel_w;
		for (int h = 0; h < height_col; h++) {
			for (int w = 0; w < width_col; w++) {
				int h_pad = h * stride_h - pad_h + h_offset;
				int w_pad = w * stride_w - pad_w + w_offset;
				long o = ((long)c * height_col + h) * width_col + w;
				if (h_pad >= 0 && h_pad < height && w_pad >= 0 && w_pad < width)
					col[o] = im[((long)c_im * height + h_pad) * width + w_pad];
				else
					col[o] = 0.0f;
			}
		}
	}
}

int oracle_conv_lrelu(const float *inputs, int ich, int w, int h, const float *weights, int k, int pad, int stride,
                      double *outputs, int ch, const float *bias, int act, double slope, double *scale)
{
	int hcol = (h + 2 * pad - k) / stride + 1, wcol = (w + 2 * pad - k) / stride + 1;
	long K = (long)ich * k * k, N = (long)hcol * wcol;
	float *col = (float *)malloc(sizeof(float) * (size_t)(K * N > 0 ? K * N : 1));
	if (!col) return -1;
	oracle_im2col(inputs, ich, h, w, k, k, pad, pad, stride, stride, col);
	for (int i = 0; i < ch; i++) {
		for (long p = 0; p < N; p++) {
			double s = 0.0, a = 0.0;
			for (long q = 0; q < K; q++) {
				double t = (double)weights[(long)i * K + q] * (double)col[q * N + p];
				s += t;
				a += fabs(t);
			}
			if (bias) {
				s += (double)bias[i];
				a += fabs((double)bias[i]);
			}
			if (act && s <= 0.0) s *= slope;
			outputs[(long)i * N + p] = s;
			if (scale) scale[(long)i * N + p] = a;
		}
	}
	free(col);
	return 0;
}

/* ------------------------------------------------------------------------
 * BLAS-1/2 helpers (ugemm.h:76-160), restated in the working precision:
 *   oracle_saxpy / oracle_daxpy: y[n*incy] += alpha * x[n*incx]   (ugemm.h:76-101)
 *   oracle_sgemv / oracle_dgemv: sgemv_cpu / dgemv_cpu (ugemm.h:103-160),
 *     y[m*incy] = alpha * sum_n a(m, n) * x[n*xinc] + beta * y[m*incy],
 *     a(m, n) = trans == 'N' ? A[m + n*lda] : A[n + m*lda], the sum taken
 *     sequentially over n.  x_by_incy != 0 reproduces the reference exactly,
 *     which reads x with incy (xinc = incy, ugemm.h:116,124,147,155); 0 gives
 *     the intended xinc = incx.  The reference also reads y when beta == 0;
 *     so does this restatement.
 * ------------------------------------------------------------------------ */
#define DEFINE_AXPY(T, NAME)                                                              \
void NAME(int N, T alpha, const T *x, int incx, T *y, int incy)                          \
{                                                                                         \
	for (int n = 0; n < N; n++) y[(long)n * incy] += alpha * x[(long)n * incx];          \
}
DEFINE_AXPY(float, oracle_saxpy)
DEFINE_AXPY(double, oracle_daxpy)

/* oracle_saxpy3 / oracle_daxpy3: z[i] = a * x[i] + y[i] with one rounding,
 * i < N -- the FMA branch of saxpy_avx / daxpy_avx (ugemm.h:41-74), the branch
 * the reference's -march=native build takes (_mm256_fmadd_ps/pd, ugemm.h:50,
 * :66).  The reference processes whole vectors of 8 (4) elements, so it also
 * writes z up to the next multiple; this restatement writes exactly N. */
void oracle_saxpy3(int N, float a, const float *x, const float *y, float *z)
{
	for (int i = 0; i < N; i++) z[i] = fmaf(a, x[i], y[i]);
}
void oracle_daxpy3(int N, double a, const double *x, const double *y, double *z)
{
	for (int i = 0; i < N; i++) z[i] = fma(a, x[i], y[i]);
}

#define DEFINE_GEMV(T, NAME)                                                                  \
void NAME(char trans, int M, int N, T alpha, const T *A, int lda, const T *x, int incx,      \
          T beta, T *y, int incy, int x_by_incy)                                              \
{                                                                                             \
	const long xinc = x_by_incy ? incy : incx;                                               \
	for (int m = 0; m < M; m++) {                                                            \
		T sum = 0;                                                                           \
		for (int n = 0; n < N; n++) {                                                        \
			const T a = trans == 'N' ? A[m + (long)n * lda] : A[n + (long)m * lda];          \
			sum += a * x[n * xinc];                                                          \
		}                                                                                    \
		y[(long)m * incy] = alpha * sum + beta * y[(long)m * incy];                          \
	}                                                                                        \
}
DEFINE_GEMV(float, oracle_sgemv)
DEFINE_GEMV(double, oracle_dgemv)
