// FP32 SGEMM on the FP32 matrix cores of gfx950 (v_mfma_f32_16x16x4_f32):
// FP32 operands as given, FP32 products, FP32 accumulation -- no operand
// split or conversion, so no preparation kernels and the error class of the
// plain FP32 kernel.  MI355X's FP32 MFMA peak (157.3 TF/s) equals its FP32
// vector peak, but the matrix core is fed one register per operand per 2048
// flop instead of one per 2 flop, so it runs much closer to that peak.
// The tile computation (LDS layout, k-permuted fragments) is m32_tile in
// mfma32_tile.h.
//
// Split-K (gridDim.y > 1): block (tile, z) multiplies K columns
// [z*kc, min(K, (z+1)*kc)) into the partial slice ws + z*M*N; the partials
// are added in fixed order by the split-K reduction, which also applies
// alpha, beta and the epilogue.  Repeated calls are bitwise identical.
#include "internal.h"
#include "kernel_common.h"
#include "mfma32_tile.h"

namespace ugemm {
namespace {

template <class CF, bool TA, bool TB, int VEC>
__global__ __launch_bounds__(CF::NT, CF::MINB) void sgemm_mfma32_kernel(
    int M, int N, int K_, float alpha_, const float *__restrict__ A_, long lda, const float *__restrict__ B_,
    long ldb, float beta_, float *__restrict__ C_, long ldc_, int tiles_m, int tiles_n, Epilogue epi, int kc,
    float *__restrict__ ws) {
    // This layout's two stages of both operands (at most CF::kLdsFloats).
    constexpr int kLds = 2 * (CF::template stage<!TA, CF::BM>() + CF::template stage<TB, CF::BN>());
    __shared__ __attribute__((aligned(16))) float lds[kLds];
    const bool part = gridDim.y > 1;
    const int kz0 = part ? (int)blockIdx.y * kc : 0;
    const int K = part ? min(kc, K_ - kz0) : K_;
    const float *A = TA ? A_ + (long)kz0 * lda : A_ + kz0;
    const float *B = TB ? B_ + kz0 : B_ + (long)kz0 * ldb;
    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
    if (part)
        m32_tile<CF, TA, TB, VEC>(M, N, K, 1.0f, A, lda, B, ldb, 0.0f, ws + (long)blockIdx.y * M * N, (long)N,
                                  tm * CF::BM, tn * CF::BN, lds);
    else
        m32_tile<CF, TA, TB, VEC>(M, N, K, alpha_, A, lda, B, ldb, beta_, C_, ldc_, tm * CF::BM, tn * CF::BN, lds,
                                  true, epi);
}

// Implicit-GEMM convolution on the FP32 matrix cores: C (M = output
// channels x N = output pixels) = W (M x K, row-major weights) * im2col(in),
// then the epilogue; the column matrix is gathered in the tile loader
// (ConvDesc).  Split-K: the chunk's rows of the offset table and its weight
// columns, as in the GEMM kernel.
template <class CF>
__global__ __launch_bounds__(CF::NT, CF::MINB) void conv_mfma32_kernel(
    int M, int N, int K_, const float *__restrict__ A_, long lda, float *__restrict__ C_, long ldc_, int tiles_m,
    int tiles_n, Epilogue epi, ConvDesc cd, bool vec, int kc, float *__restrict__ ws) {
    constexpr int kLds = 2 * (CF::template stage<true, CF::BM>() + CF::template stage<false, CF::BN>());
    __shared__ __attribute__((aligned(16))) float lds[kLds];
    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
    if (gridDim.y > 1) {
        const int k0 = blockIdx.y * kc, kz = min(kc, K_ - k0);
        ConvDesc cz = cd;
        cz.ktab = cd.ktab + k0;
        m32_tile<CF, false, false, 4, true>(M, N, kz, 1.0f, A_ + k0, lda, nullptr, 0, 0.0f,
                                            ws + (long)blockIdx.y * M * N, (long)N, tm * CF::BM, tn * CF::BN, lds, vec,
                                            Epilogue(), cz);
    } else {
        m32_tile<CF, false, false, 4, true>(M, N, K_, 1.0f, A_, lda, nullptr, 0, 0.0f, C_, ldc_, tm * CF::BM,
                                            tn * CF::BN, lds, vec, epi, cd);
    }
}

using V0 = M32Cfg<128, 128, 16, 2, 2, 2>;
using V1 = M32Cfg<256, 128, 16, 4, 2>;
using V2 = M32Cfg<128, 256, 16, 2, 4>;
using V3 = M32Cfg<64, 64, 16, 2, 2, 2>;
using V4 = M32Cfg<256, 256, 16, 2, 4>;
using V5 = M32Cfg<128, 128, 32, 2, 2, 2>;
const char *kNames[] = {"f32_mfma_128x128x16_w2x2", "f32_mfma_256x128x16_w4x2", "f32_mfma_128x256x16_w2x4",
                        "f32_mfma_64x64x16_w2x2",   "f32_mfma_256x256x16_w2x4", "f32_mfma_128x128x32_w2x2"};
constexpr int kNum = 6;

template <class CF>
int launch_cfg(const Gemm<float> &g, hipStream_t s) {
    const int tiles_m = (g.M + CF::BM - 1) / CF::BM, tiles_n = (g.N + CF::BN - 1) / CF::BN;
    const long nblk = (long)tiles_m * tiles_n;
    if (nblk > 0x7fffffffL) return fail(UGEMM_ERR_ARG, "sgemm fp32 mfma: too many tiles (%ld)", nblk);
    const bool vec = ugemm_aligned16(g.A) && ugemm_aligned16(g.B) && (g.lda % 4 == 0) && (g.ldb % 4 == 0);
    int kc = g.K;
    const int splits = split_count(g.splits, nblk, g.K, 64, &kc, 2);
    float *ws = nullptr;
    if (splits > 1) {
        const int rc = stream_alloc((void **)&ws, (size_t)splits * g.M * g.N * sizeof(float), s);
        if (rc) return rc;
    }
    dim3 grid((unsigned)nblk, (unsigned)splits), block(CF::NT);
#define UGEMM_L(TA, TB, VEC)                                                                                   \
    hipLaunchKernelGGL((sgemm_mfma32_kernel<CF, TA, TB, VEC>), grid, block, 0, s, g.M, g.N, g.K, g.alpha, g.A, \
                       g.lda, g.B, g.ldb, g.beta, g.C, g.ldc, tiles_m, tiles_n, g.epi, kc, ws)
    timing_begin(s);
    if (vec) {
        if (!g.ta && !g.tb) UGEMM_L(false, false, 4);
        else if (!g.ta && g.tb) UGEMM_L(false, true, 4);
        else if (g.ta && !g.tb) UGEMM_L(true, false, 4);
        else UGEMM_L(true, true, 4);
    } else {
        if (!g.ta && !g.tb) UGEMM_L(false, false, 1);
        else if (!g.ta && g.tb) UGEMM_L(false, true, 1);
        else if (g.ta && !g.tb) UGEMM_L(true, false, 1);
        else UGEMM_L(true, true, 1);
    }
    timing_end(s);
#undef UGEMM_L
    hipError_t e = hipGetLastError();
    int rc = e == hipSuccess ? UGEMM_OK : hip_fail(e, "sgemm fp32 mfma launch");
    if (splits > 1) {
        if (!rc) rc = launch_splitk_reduce_f32(g.M, g.N, splits, g.alpha, ws, g.beta, g.C, g.ldc, g.epi, s);
        const int rc2 = stream_free(ws, s);
        if (!rc) rc = rc2;
    }
    return rc;
}

template <class CF>
int launch_conv_cfg(const Gemm<float> &g, const ConvDesc &cd, hipStream_t s) {
    const int tiles_m = (g.M + CF::BM - 1) / CF::BM, tiles_n = (g.N + CF::BN - 1) / CF::BN;
    const long nblk = (long)tiles_m * tiles_n;
    if (nblk > 0x7fffffffL) return fail(UGEMM_ERR_ARG, "conv fp32 mfma: too many tiles (%ld)", nblk);
    const bool vec = ugemm_aligned16(g.A) && (g.lda % 4 == 0);
    int kc = g.K;
    const int splits = split_count(g.splits, nblk, g.K, 64, &kc, 2);
    float *ws = nullptr;
    if (splits > 1) {
        const int rc = stream_alloc((void **)&ws, (size_t)splits * g.M * g.N * sizeof(float), s);
        if (rc) return rc;
    }
    timing_begin(s);
    hipLaunchKernelGGL((conv_mfma32_kernel<CF>), dim3((unsigned)nblk, (unsigned)splits), dim3(CF::NT), 0, s, g.M,
                       g.N, g.K, g.A, g.lda, g.C, g.ldc, tiles_m, tiles_n, g.epi, cd, vec, kc, ws);
    timing_end(s);
    hipError_t e = hipGetLastError();
    int rc = e == hipSuccess ? UGEMM_OK : hip_fail(e, "conv fp32 mfma launch");
    if (splits > 1) {
        if (!rc) rc = launch_splitk_reduce_f32(g.M, g.N, splits, 1.0f, ws, 0.0f, g.C, g.ldc, g.epi, s);
        const int rc2 = stream_free(ws, s);
        if (!rc) rc = rc2;
    }
    return rc;
}

}  // namespace

int m32_variant_count() { return kNum; }
const char *m32_variant_name(int v) { return (v >= 0 && v < kNum) ? kNames[v] : nullptr; }

// 256x128 tiles of 8 waves wherever they fill every CU (measured on MI355X at
// 8192^3 NN: 137 TF/s against 122 for 128x128, and 131 against 106 on
// 65536x256x4096); 128x128 when those fill it; 64x64 below (with split-K
// on a grid that is still short).
int m32_pick_variant(const Gemm<float> &g) {
    const long cus = device_cu_count();
    const long t256 = (long)((g.M + 255) / 256) * ((g.N + 127) / 128);
    if (t256 >= cus) return 1;
    const long t128 = (long)((g.M + 127) / 128) * ((g.N + 127) / 128);
    return t128 >= cus ? 0 : 3;
}

// g: the convolution as a GEMM (alpha 1, beta 0, A the weights, B unused).
int launch_conv_m32(const Gemm<float> &g, const ConvDesc &cd, int variant, hipStream_t s) {
    if (variant < 0) variant = m32_pick_variant(g);
    switch (variant) {
        case 0: return launch_conv_cfg<V0>(g, cd, s);
        case 1: return launch_conv_cfg<V1>(g, cd, s);
        case 2: return launch_conv_cfg<V2>(g, cd, s);
        case 3: return launch_conv_cfg<V3>(g, cd, s);
        case 4: return launch_conv_cfg<V4>(g, cd, s);
        case 5: return launch_conv_cfg<V5>(g, cd, s);
        default: return fail(UGEMM_ERR_ARG, "conv fp32 mfma: no variant %d", variant);
    }
}

int launch_m32(const Gemm<float> &g, int variant, hipStream_t s) {
    switch (variant) {
        case 0: return launch_cfg<V0>(g, s);
        case 1: return launch_cfg<V1>(g, s);
        case 2: return launch_cfg<V2>(g, s);
        case 3: return launch_cfg<V3>(g, s);
        case 4: return launch_cfg<V4>(g, s);
        case 5: return launch_cfg<V5>(g, s);
        default: return fail(UGEMM_ERR_ARG, "sgemm fp32 mfma: no variant %d", variant);
    }
}

}  // namespace ugemm
