// Plain FP32 SGEMM for gfx950: vector FMA with FP32 accumulation, one
// workgroup per C tile (tile body in fp32_tile.h).  Tiles are visited in the
// XCD-aware order of tile_coords (kernel_common.h).
#include "internal.h"
#include "fp32_tile.h"

namespace ugemm {
namespace {

// Split-K (gridDim.y > 1): block (tile, z) multiplies K columns
// [z*kc, min(K, (z+1)*kc)) into the partial slice ws + z*M*N (row length N).
template <class CF, bool TA, bool TB, int VEC>
__global__ __launch_bounds__(CF::NT, CF::MINB) void sgemm_fp32_kernel(int M, int N, int K, float alpha,
                                                            const float *__restrict__ A, long lda,
                                                            const float *__restrict__ B, long ldb, float beta,
                                                            float *__restrict__ C, long ldc, int tiles_m, int tiles_n,
                                                            Epilogue epi, int kc, float *__restrict__ ws) {
    __shared__ __attribute__((aligned(16))) float lds[2 * CF::BK * (CF::LDA + CF::LDB)];
    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
    if (gridDim.y > 1) {
        const int k0 = blockIdx.y * kc, kz = min(kc, K - k0);
        const float *Az = TA ? A + (long)k0 * lda : A + k0;
        const float *Bz = TB ? B + k0 : B + (long)k0 * ldb;
        fp32_tile<CF, TA, TB, VEC>(M, N, kz, 1.0f, Az, lda, Bz, ldb, 0.0f, ws + (long)blockIdx.y * M * N, N,
                                   tm * CF::BM, tn * CF::BN, lds, true, Epilogue());
    } else {
        fp32_tile<CF, TA, TB, VEC>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tm * CF::BM, tn * CF::BN, lds, true,
                                   epi);
    }
}

// Implicit-GEMM convolution: C (M = output channels x N = output pixels) =
// alpha * W (M x K, row-major weights) * im2col(in) + beta * C, then the
// epilogue.  The column matrix is gathered in the tile loader (ConvDesc).
// Split-K as in sgemm_fp32_kernel: the chunk's rows of the offset table
// (absolute image offsets) and its weight columns.
template <class CF, bool TA, int VEC>
__global__ __launch_bounds__(CF::NT) void conv_fp32_kernel(int M, int N, int K, float alpha,
                                                           const float *__restrict__ A, long lda, float beta,
                                                           float *__restrict__ C, long ldc, int tiles_m, int tiles_n,
                                                           Epilogue epi, ConvDesc cd, int kc, float *__restrict__ ws) {
    __shared__ __attribute__((aligned(16))) float lds[2 * CF::BK * (CF::LDA + CF::LDB)];
    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
    if (gridDim.y > 1) {
        const int k0 = blockIdx.y * kc, kz = min(kc, K - k0);
        ConvDesc cz = cd;
        cz.ktab = cd.ktab + k0;
        const float *Az = TA ? A + (long)k0 * lda : A + k0;
        fp32_tile<CF, TA, false, VEC, true>(M, N, kz, 1.0f, Az, lda, nullptr, 0, 0.0f, ws + (long)blockIdx.y * M * N,
                                            N, tm * CF::BM, tn * CF::BN, lds, true, Epilogue(), cz);
    } else {
        fp32_tile<CF, TA, false, VEC, true>(M, N, K, alpha, A, lda, nullptr, 0, beta, C, ldc, tm * CF::BM,
                                            tn * CF::BN, lds, true, epi, cd);
    }
}

// ---------------------------------------------------------------------------
// Variant table.
using V0 = Fp32Cfg<128, 128, 16, 8, 8>;
using V1 = Fp32Cfg<64, 64, 16, 4, 4>;
using V2 = Fp32Cfg<128, 64, 16, 8, 4>;
using V3 = Fp32Cfg<64, 128, 16, 4, 8>;
using V4 = Fp32Cfg<256, 128, 16, 16, 8>;
using V5 = Fp32Cfg<128, 256, 16, 8, 16>;

const char *kNames[] = {"fp32_128x128x16_t8x8", "fp32_64x64x16_t4x4", "fp32_128x64x16_t8x4",
                        "fp32_64x128x16_t4x8", "fp32_256x128x16_t16x8", "fp32_128x256x16_t8x16"};
constexpr int kNumVariants = 6;

template <class CF>
int launch_cfg(const Gemm<float> &g, hipStream_t s) {
    const int tiles_m = (g.M + CF::BM - 1) / CF::BM;
    const int tiles_n = (g.N + CF::BN - 1) / CF::BN;
    const long nblk = (long)tiles_m * tiles_n;
    if (nblk > 0x7fffffffL) return fail(UGEMM_ERR_ARG, "sgemm: too many tiles (%ld)", nblk);
    // 128-bit loads need 16-byte aligned bases and leading dimensions that are multiples of 4
    // (split-K chunk offsets are multiples of 64 columns, which keeps that).
    const bool vec = ugemm_aligned16(g.A) && ugemm_aligned16(g.B) && (g.lda % 4 == 0) && (g.ldb % 4 == 0);
    int kc = g.K;
    const int splits = split_count(g.splits, nblk, g.K, 64, &kc, 2);
    float *ws = nullptr;
    if (splits > 1) {
        const int rc = stream_alloc((void **)&ws, (size_t)splits * g.M * g.N * sizeof(float), s);
        if (rc) return rc;
    }
    dim3 grid((unsigned)nblk, (unsigned)splits), block(CF::NT);
#define UGEMM_L(TA, TB, VEC)                                                                             \
    hipLaunchKernelGGL((sgemm_fp32_kernel<CF, TA, TB, VEC>), grid, block, 0, s, g.M, g.N, g.K, g.alpha, g.A, \
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
    int rc = e == hipSuccess ? UGEMM_OK : hip_fail(e, "sgemm launch");
    if (splits > 1) {
        if (!rc) rc = launch_splitk_reduce_f32(g.M, g.N, splits, g.alpha, ws, g.beta, g.C, g.ldc, g.epi, s);
        const int rc2 = stream_free(ws, s);
        if (!rc) rc = rc2;
    }
    return rc;
}

template <class CF>
int launch_conv_cfg(const Gemm<float> &g, const ConvDesc &cd, hipStream_t s) {
    const int tiles_m = (g.M + CF::BM - 1) / CF::BM;
    const int tiles_n = (g.N + CF::BN - 1) / CF::BN;
    const long nblk = (long)tiles_m * tiles_n;
    if (nblk > 0x7fffffffL) return fail(UGEMM_ERR_ARG, "conv: too many tiles (%ld)", nblk);
    const bool vec = ugemm_aligned16(g.A) && (g.lda % 4 == 0);
    int kc = g.K;
    const int splits = split_count(g.splits, nblk, g.K, 64, &kc, 2);
    float *ws = nullptr;
    if (splits > 1) {
        const int rc = stream_alloc((void **)&ws, (size_t)splits * g.M * g.N * sizeof(float), s);
        if (rc) return rc;
    }
    dim3 grid((unsigned)nblk, (unsigned)splits), block(CF::NT);
    timing_begin(s);
    if (vec && !g.ta)
        hipLaunchKernelGGL((conv_fp32_kernel<CF, false, 4>), grid, block, 0, s, g.M, g.N, g.K, g.alpha, g.A, g.lda,
                           g.beta, g.C, g.ldc, tiles_m, tiles_n, g.epi, cd, kc, ws);
    else if (!g.ta)
        hipLaunchKernelGGL((conv_fp32_kernel<CF, false, 1>), grid, block, 0, s, g.M, g.N, g.K, g.alpha, g.A, g.lda,
                           g.beta, g.C, g.ldc, tiles_m, tiles_n, g.epi, cd, kc, ws);
    else
        hipLaunchKernelGGL((conv_fp32_kernel<CF, true, 1>), grid, block, 0, s, g.M, g.N, g.K, g.alpha, g.A, g.lda,
                           g.beta, g.C, g.ldc, tiles_m, tiles_n, g.epi, cd, kc, ws);
    timing_end(s);
    hipError_t e = hipGetLastError();
    int rc = e == hipSuccess ? UGEMM_OK : hip_fail(e, "conv launch");
    if (splits > 1) {
        if (!rc) rc = launch_splitk_reduce_f32(g.M, g.N, splits, g.alpha, ws, g.beta, g.C, g.ldc, g.epi, s);
        const int rc2 = stream_free(ws, s);
        if (!rc) rc = rc2;
    }
    return rc;
}

}  // namespace

int fp32_variant_count() { return kNumVariants; }
const char *fp32_variant_name(int v) { return (v >= 0 && v < kNumVariants) ? kNames[v] : nullptr; }

int fp32_pick_variant(const Gemm<float> &g) {
    // Enough 128x256 tiles to give every CU two: the 8x16 micro-tile, whose
    // k-step reads 24 floats from LDS per 128 FMAs (8x8: 16 per 64), the LDS
    // bandwidth that bounds the 8x8 kernel (measured 6-8% faster at 8192^3).
    // Else enough 128x128 tiles: those; otherwise shrink the tile so that the
    // grid still covers the chip.
    // Row-major NN (A k-contiguous, B row-contiguous) runs best on the 16x8
    // micro-tile: 98-100 against 92-93 TF/s for 8x16 at 4096^3 and 8192^3,
    // interleaved on one box; NT and TN stay on 8x16 (within 1-3% of 8x8).
    // A grid of 128x256 tiles covering half the CUs still takes 8x16, the
    // launcher splitting K (2048^3: 77 against 74 TF/s on 64x64 tiles).
    const int cus = device_cu_count();
    const long t128x256 = (long)((g.M + 127) / 128) * ((g.N + 255) / 256);
    const long t256x128 = (long)((g.M + 255) / 256) * ((g.N + 127) / 128);
    if (!g.ta && !g.tb && t256x128 >= 2L * cus) return 4;
    if (t128x256 >= 2L * cus) return 5;
    const long t128 = (long)((g.M + 127) / 128) * ((g.N + 127) / 128);
    if (t128 >= 2L * cus) return 0;
    if (2 * t128x256 >= cus && g.K >= 1024 && g.splits != 1) return 5;
    const long t64 = (long)((g.M + 63) / 64) * ((g.N + 63) / 64);
    if (g.M >= 4 * g.N && t64 >= cus) return 2;
    if (g.N >= 4 * g.M && t64 >= cus) return 3;
    return 1;
}

int launch_conv_fp32(const Gemm<float> &g, const ConvDesc &cd, int variant, hipStream_t s) {
    // Convolutions are compiled for the four 8x8-or-smaller tiles.
    if (variant < 0) variant = fp32_pick_variant(g);
    if (variant >= 4) variant = 0;
    switch (variant) {
        case 0: return launch_conv_cfg<V0>(g, cd, s);
        case 1: return launch_conv_cfg<V1>(g, cd, s);
        case 2: return launch_conv_cfg<V2>(g, cd, s);
        case 3: return launch_conv_cfg<V3>(g, cd, s);
        default: return fail(UGEMM_ERR_ARG, "conv: no variant %d", variant);
    }
}

int launch_fp32(const Gemm<float> &g, int variant, hipStream_t s) {
    switch (variant) {
        case 0: return launch_cfg<V0>(g, s);
        case 1: return launch_cfg<V1>(g, s);
        case 2: return launch_cfg<V2>(g, s);
        case 3: return launch_cfg<V3>(g, s);
        case 4: return launch_cfg<V4>(g, s);
        case 5: return launch_cfg<V5>(g, s);
        default: return fail(UGEMM_ERR_ARG, "fp32: no variant %d", variant);
    }
}

}  // namespace ugemm
