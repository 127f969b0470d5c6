// BLAS-1/2 helpers of the ugemm C API on gfx950: axpy and gemv, FP32 and
// FP64.  Both are HBM-bound (an axpy moves 3 words per element, a gemv reads
// the matrix once), so the kernels aim at coalesced, vectorised streaming,
// not at arithmetic.
//
// Semantics follow saxpy_cpu / daxpy_cpu (ugemm.h:76-101):
//   y[n*incy] += alpha * x[n*incx],  n < N
// saxpy_avx / daxpy_avx (ugemm.h:41-74): z[i] = fma(a, x[i], y[i]), i < N
// and sgemv_cpu / dgemv_cpu (ugemm.h:103-160), whose element mapping is
//   'N': y[m*incy] = alpha * sum_n A[m + n*lda] * x[n*incx] + beta * y[m*incy]
//   'T': y[m*incy] = alpha * sum_n A[n + m*lda] * x[n*incx] + beta * y[m*incy]
//   (m < M, n < N; any trans other than 'N'/'n' is 'T', as in the reference)
// with one deliberate difference: the reference reads x with incy
// (ugemm.h:116, :124, :147, :155), which is only right when incx == incy;
// here x uses incx.  Increments must be >= 1.  beta == 0 does not read y.
//
// gemv 'N' (column-major walk, A[m + n*lda] contiguous in m): threads own
// rows m and stream down the columns; N is split over up to 64 blocks in y's
// grid dimension, each writing a partial sum per row, and a second kernel
// adds the partials in a fixed order (deterministic, no atomics).
// gemv 'T' (A[n + m*lda] contiguous in n): one wave per output row, lanes
// stride along n, then a wave reduction (fixed shuffle order).
#include <algorithm>

#include "internal.h"
#include "kernel_common.h"

namespace ugemm {
namespace {

template <typename T>
__global__ __launch_bounds__(256) void axpy_kernel(long N, T alpha, const T *__restrict__ x, long incx,
                                                   T *__restrict__ y, long incy) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) y[i * incy] += alpha * x[i * incx];
}

// Unit strides, 16-byte aligned x and y: 16 bytes per thread per step.
template <typename T>
__global__ __launch_bounds__(256) void axpy_vec_kernel(long N, T alpha, const T *__restrict__ x, T *__restrict__ y) {
    typedef typename Vec16<T, 16 / sizeof(T)>::type V;
    constexpr int W = 16 / sizeof(T);
    const long nv = N / W;
    const long stride = (long)gridDim.x * blockDim.x;
    const V *xv = reinterpret_cast<const V *>(x);
    V *yv = reinterpret_cast<V *>(y);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride) {
        V a = xv[i], b = yv[i];
        T *pa = reinterpret_cast<T *>(&a), *pb = reinterpret_cast<T *>(&b);
#pragma unroll
        for (int j = 0; j < W; ++j) pb[j] = fma(alpha, pa[j], pb[j]);
        yv[i] = b;
    }
    // tail (fewer than W elements), one thread
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (long i = nv * W; i < N; ++i) y[i] = fma(alpha, x[i], y[i]);
}

// Three-operand axpy, the FMA branch of saxpy_avx / daxpy_avx
// (ugemm.h:41-74): z[i] = fma(a, x[i], y[i]), i < N, z written with
// non-temporal stores (the reference's _mm256_stream_ps/pd).  z may be y.
template <typename T>
__global__ __launch_bounds__(256) void axpy3_kernel(long N, T a, const T *__restrict__ x, const T *y, T *z) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride)
        __builtin_nontemporal_store(fma(a, x[i], y[i]), z + i);
}

// 16-byte aligned x, y, z: one 16-byte vector per thread.
template <typename T>
__global__ __launch_bounds__(256) void axpy3_vec_kernel(long N, T a, const T *__restrict__ x, const T *y, T *z) {
    constexpr int W = 16 / sizeof(T);
    typedef T V __attribute__((ext_vector_type(W)));
    const long nv = N / W;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride) {
        const V vx = reinterpret_cast<const V *>(x)[i], vy = reinterpret_cast<const V *>(y)[i];
        V vz;
#pragma unroll
        for (int j = 0; j < W; ++j) vz[j] = fma(a, vx[j], vy[j]);
        __builtin_nontemporal_store(vz, reinterpret_cast<V *>(z) + i);
    }
    // tail (fewer than W elements), one thread
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (long i = nv * W; i < N; ++i) z[i] = fma(a, x[i], y[i]);
}

// gemv 'N': block (bx, by) covers rows [256*bx, +256) and columns n in
// [by*chunk, min(N, (by+1)*chunk)); partial[by][m] (or y directly when there
// is one chunk).
template <typename T>
__global__ __launch_bounds__(256) void gemv_n_kernel(int M, int N, int chunk, T alpha, const T *__restrict__ A, long lda,
                                                     const T *__restrict__ x, long incx, T beta, T *__restrict__ y,
                                                     long incy, T *__restrict__ partial) {
    const int m = blockIdx.x * 256 + threadIdx.x;
    if (m >= M) return;
    const int n0 = blockIdx.y * chunk, n1 = min(N, n0 + chunk);
    T s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int n = n0;
    for (; n + 4 <= n1; n += 4) {
        s0 = fma(A[m + (long)n * lda], x[(long)n * incx], s0);
        s1 = fma(A[m + (long)(n + 1) * lda], x[(long)(n + 1) * incx], s1);
        s2 = fma(A[m + (long)(n + 2) * lda], x[(long)(n + 2) * incx], s2);
        s3 = fma(A[m + (long)(n + 3) * lda], x[(long)(n + 3) * incx], s3);
    }
    for (; n < n1; ++n) s0 = fma(A[m + (long)n * lda], x[(long)n * incx], s0);
    const T s = (s0 + s1) + (s2 + s3);
    if (partial) {
        partial[(long)blockIdx.y * M + m] = s;
    } else {
        T *py = y + (long)m * incy;
        *py = beta == T(0) ? alpha * s : fma(alpha, s, beta * *py);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void gemv_n_reduce(int M, int splits, T alpha, const T *__restrict__ partial, T beta,
                                                     T *__restrict__ y, long incy) {
    const int m = blockIdx.x * 256 + threadIdx.x;
    if (m >= M) return;
    T s = 0;
    for (int k = 0; k < splits; ++k) s += partial[(long)k * M + m];
    T *py = y + (long)m * incy;
    *py = beta == T(0) ? alpha * s : fma(alpha, s, beta * *py);
}

// gemv 'T': 4 waves per block, one row m per wave.
template <typename T>
__global__ __launch_bounds__(256) void gemv_t_kernel(int M, int N, T alpha, const T *__restrict__ A, long lda,
                                                     const T *__restrict__ x, long incx, T beta, T *__restrict__ y,
                                                     long incy) {
    const int lane = threadIdx.x & 63;
    const int m = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (m >= M) return;
    const T *row = A + (long)m * lda;
    // Four independent partial sums: four loads of the row (and of x) in
    // flight per step; with two, the wave waited on HBM after each pair.
    T s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int n = lane;
    for (; n + 192 < N; n += 256) {
        s0 = fma(row[n], x[(long)n * incx], s0);
        s1 = fma(row[n + 64], x[(long)(n + 64) * incx], s1);
        s2 = fma(row[n + 128], x[(long)(n + 128) * incx], s2);
        s3 = fma(row[n + 192], x[(long)(n + 192) * incx], s3);
    }
    for (; n < N; n += 64) s0 = fma(row[n], x[(long)n * incx], s0);
    T s = (s0 + s1) + (s2 + s3);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) {
        T *py = y + (long)m * incy;
        *py = beta == T(0) ? alpha * s : fma(alpha, s, beta * *py);
    }
}

template <typename T>
int axpy(int N, T alpha, const T *x, int incx, T *y, int incy, hipStream_t s) {
    if (N < 0) return fail(UGEMM_ERR_ARG, "axpy: N=%d", N);
    if (incx < 1 || incy < 1) return fail(UGEMM_ERR_ARG, "axpy: increments must be >= 1 (incx=%d incy=%d)", incx, incy);
    if (N == 0 || alpha == T(0)) return UGEMM_OK;
    if (!x || !y) return fail(UGEMM_ERR_ARG, "axpy: NULL vector");
    const int cus = device_cu_count();
    if (incx == 1 && incy == 1 && ugemm_aligned16(x) && ugemm_aligned16(y)) {
        // One 16-byte vector per thread (the grid-stride loop only matters
        // beyond 2^28 vectors): a grid capped at 8 blocks per CU looped
        // with two loads in flight per wave at a time.
        const long nv = N / (16 / (long)sizeof(T));
        const long blocks = std::max<long>(1, std::min<long>((nv + 255) / 256, 1L << 20));
        hipLaunchKernelGGL(axpy_vec_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s, (long)N, alpha, x, y);
    } else {
        const long blocks = std::max<long>(1, std::min<long>((N + 255L) / 256, 8L * cus));
        hipLaunchKernelGGL(axpy_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s, (long)N, alpha, x, (long)incx, y,
                           (long)incy);
    }
    UGEMM_HIP_TRY(hipGetLastError());
    return UGEMM_OK;
}

template <typename T>
int axpy3(int N, T a, const T *x, const T *y, T *z, hipStream_t s) {
    if (N < 0) return fail(UGEMM_ERR_ARG, "axpy3: N=%d", N);
    if (N == 0) return UGEMM_OK;
    if (!x || !y || !z) return fail(UGEMM_ERR_ARG, "axpy3: NULL vector");
    if (ugemm_aligned16(x) && ugemm_aligned16(y) && ugemm_aligned16(z)) {
        const long nv = N / (16 / (long)sizeof(T));
        const long blocks = std::max<long>(1, std::min<long>((nv + 255) / 256, 1L << 20));
        hipLaunchKernelGGL(axpy3_vec_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s, (long)N, a, x, y, z);
    } else {
        const long blocks = std::max<long>(1, std::min<long>((N + 255L) / 256, 8L * device_cu_count()));
        hipLaunchKernelGGL(axpy3_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s, (long)N, a, x, y, z);
    }
    UGEMM_HIP_TRY(hipGetLastError());
    return UGEMM_OK;
}

template <typename T>
int gemv(char trans, int M, int N, T alpha, const T *A, int lda, const T *x, int incx, T beta, T *y, int incy,
         hipStream_t s) {
    if (M < 0 || N < 0) return fail(UGEMM_ERR_ARG, "gemv: M=%d N=%d", M, N);
    if (incx < 1 || incy < 1) return fail(UGEMM_ERR_ARG, "gemv: increments must be >= 1 (incx=%d incy=%d)", incx, incy);
    const bool t = !(trans == 'N' || trans == 'n');
    if (lda < std::max(1, t ? N : M)) return fail(UGEMM_ERR_ARG, "gemv: lda=%d too small", lda);
    if (M == 0) return UGEMM_OK;
    if (!y || (N > 0 && (!A || !x))) return fail(UGEMM_ERR_ARG, "gemv: NULL operand");
    if (N == 0 || alpha == T(0)) {
        // y = beta * y (beta == 0 writes zeros): the gemv kernels with N = 0.
        hipLaunchKernelGGL(gemv_n_kernel<T>, dim3((unsigned)((M + 255) / 256), 1), dim3(256), 0, s, M, 0, 1, alpha, A,
                           (long)lda, x, (long)incx, beta, y, (long)incy, (T *)nullptr);
        UGEMM_HIP_TRY(hipGetLastError());
        return UGEMM_OK;
    }
    if (t) {
        hipLaunchKernelGGL(gemv_t_kernel<T>, dim3((unsigned)((M + 3) / 4)), dim3(256), 0, s, M, N, alpha, A,
                           (long)lda, x, (long)incx, beta, y, (long)incy);
        UGEMM_HIP_TRY(hipGetLastError());
        return UGEMM_OK;
    }
    // 'N': enough blocks to cover the chip -- split N when M alone is too small.
    const int cus = device_cu_count();
    const int mblocks = (M + 255) / 256;
    int splits = std::max(1, std::min(64, (4 * cus + mblocks - 1) / mblocks));
    splits = std::min(splits, std::max(1, N / 64));
    const int chunk = (N + splits - 1) / splits;
    splits = (N + chunk - 1) / chunk;
    if (splits == 1) {
        hipLaunchKernelGGL(gemv_n_kernel<T>, dim3((unsigned)mblocks, 1), dim3(256), 0, s, M, N, chunk, alpha, A,
                           (long)lda, x, (long)incx, beta, y, (long)incy, (T *)nullptr);
        UGEMM_HIP_TRY(hipGetLastError());
        return UGEMM_OK;
    }
    void *ws = nullptr;
    int rc = stream_alloc(&ws, (size_t)splits * M * sizeof(T), s);
    if (rc) return rc;
    hipLaunchKernelGGL(gemv_n_kernel<T>, dim3((unsigned)mblocks, (unsigned)splits), dim3(256), 0, s, M, N, chunk, alpha,
                       A, (long)lda, x, (long)incx, beta, y, (long)incy, (T *)ws);
    hipLaunchKernelGGL(gemv_n_reduce<T>, dim3((unsigned)mblocks), dim3(256), 0, s, M, splits, alpha, (const T *)ws, beta,
                       y, (long)incy);
    hipError_t e = hipGetLastError();
    rc = e == hipSuccess ? UGEMM_OK : hip_fail(e, "gemv launch");
    const int rc2 = stream_free(ws, s);
    return rc ? rc : rc2;
}

}  // namespace

int launch_saxpy(int N, float alpha, const float *x, int incx, float *y, int incy, hipStream_t s) {
    return axpy<float>(N, alpha, x, incx, y, incy, s);
}
int launch_daxpy(int N, double alpha, const double *x, int incx, double *y, int incy, hipStream_t s) {
    return axpy<double>(N, alpha, x, incx, y, incy, s);
}
int launch_saxpy3(int N, float a, const float *x, const float *y, float *z, hipStream_t s) {
    return axpy3<float>(N, a, x, y, z, s);
}
int launch_daxpy3(int N, double a, const double *x, const double *y, double *z, hipStream_t s) {
    return axpy3<double>(N, a, x, y, z, s);
}
int launch_sgemv(char trans, int M, int N, float alpha, const float *A, int lda, const float *x, int incx, float beta,
                 float *y, int incy, hipStream_t s) {
    return gemv<float>(trans, M, N, alpha, A, lda, x, incx, beta, y, incy, s);
}
int launch_dgemv(char trans, int M, int N, double alpha, const double *A, int lda, const double *x, int incx,
                 double beta, double *y, int incy, hipStream_t s) {
    return gemv<double>(trans, M, N, alpha, A, lda, x, incx, beta, y, incy, s);
}

}  // namespace ugemm
