// Split-K support shared by the FP32, split-form and FP64 kernels: the split
// count heuristic and the fixed-order reduction of the partial products.
#include <algorithm>

#include "internal.h"

namespace ugemm {
namespace {

// C[m, n] = alpha * sum_{z < S} ws[z][m, n] + beta * C[m, n] (C not read when
// beta == 0), then the epilogue; the slices are added in z order.
template <typename T>
__global__ __launch_bounds__(256) void splitk_reduce(int M, int N, int S, T alpha, const T *__restrict__ ws, T beta,
                                                     T *__restrict__ C, long ldc, Epilogue epi) {
    const long total = (long)M * N, slice = total;
    const bool has_epi = epi.active();
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        T acc = ws[i];
        for (int z = 1; z < S; ++z) acc += ws[z * slice + i];
        const int m = (int)(i / N), n = (int)(i - (long)m * N);
        T *cp = C + (long)m * ldc + n;
        T v = beta == T(0) ? alpha * acc : fma(alpha, acc, beta * *cp);
        if constexpr (sizeof(T) == 4) {
            if (has_epi) v = apply_epilogue(epi, v, m, n);
        }
        *cp = v;
    }
}

template <typename T>
int reduce(int M, int N, int S, T alpha, const T *ws, T beta, T *C, long ldc, const Epilogue &epi, hipStream_t s) {
    const long total = (long)M * N;
    const long blocks = std::max<long>(1, std::min<long>((total + 255) / 256, 16L * device_cu_count()));
    hipLaunchKernelGGL(splitk_reduce<T>, dim3((unsigned)blocks), dim3(256), 0, s, M, N, S, alpha, ws, beta, C, ldc,
                       epi);
    UGEMM_HIP_TRY(hipGetLastError());
    return UGEMM_OK;
}

}  // namespace

int split_count(int requested, long tiles, int K, int kalign, int *kc, int per_cu) {
    int S = requested;
    if (S <= 0) {
        // Auto: split only when the grid fills at most half of the wanted
        // per_cu blocks on every CU, enough to reach that, with at least
        // 512 K per split.
        const long want = (long)per_cu * device_cu_count();
        S = 1;
        if (tiles * 2 <= want) S = (int)std::min<long>({want / std::max<long>(1, tiles), (long)K / 512, 16L});
        S = std::max(S, 1);
    }
    int chunk = (K + S - 1) / S;
    chunk = (chunk + kalign - 1) / kalign * kalign;
    if (chunk <= 0) chunk = kalign;
    S = std::max(1, (K + chunk - 1) / chunk);   // every split has a non-empty K range
    *kc = chunk;
    return S;
}

int launch_splitk_reduce_f32(int M, int N, int splits, float alpha, const float *ws, float beta, float *C, long ldc,
                             const Epilogue &epi, hipStream_t s) {
    return reduce<float>(M, N, splits, alpha, ws, beta, C, ldc, epi, s);
}
int launch_splitk_reduce_f64(int M, int N, int splits, double alpha, const double *ws, double beta, double *C,
                             long ldc, hipStream_t s) {
    return reduce<double>(M, N, splits, alpha, ws, beta, C, ldc, Epilogue(), s);
}

}  // namespace ugemm
