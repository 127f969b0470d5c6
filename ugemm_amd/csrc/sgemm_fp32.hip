// Plain FP32 SGEMM for gfx950: vector FMA with FP32 accumulation.
//
// Design (one workgroup = one BM x BN tile of C):
//   * global -> registers -> LDS, two LDS stages: tile k+1 is fetched into
//     registers while tile k is multiplied from LDS, so one barrier per BK step;
//   * LDS holds both operands K-major: As[k][m], Bs[k][n]; the transposed
//     layouts are turned around while writing LDS (no separate transpose pass);
//   * each thread owns a TM x TN register micro-tile built from 4-wide groups,
//     read from LDS with 128-bit loads; a wave is an 8 x 8 grid of threads, so a
//     k-step reads 128 contiguous bytes of As and of Bs per wave;
//   * tiles are visited XCD-aware: the linear block id is re-mapped so that the
//     blocks resident on one XCD (one L2) cover a compact group of C tiles.
// Semantics follow the reference's row-major kernel gemm_rnn (sgemm_ocl2.h:17-90):
// C = alpha * acc + beta * C, C is not read when beta == 0 (sgemm_ocl2.h:86-88).
#include "internal.h"
#include "kernel_common.h"

namespace ugemm {
namespace {

constexpr int LDS_PAD = 4;   // floats; keeps rows 16-byte aligned

// Empty register constraints on the accumulators: the FMAs of one k-step
// must complete before the next step's code, which bounds live ranges.
template <int TM, int TN>
__device__ __forceinline__ void pin_step(float (&acc)[TM][TN]) {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; j += 4)
            asm volatile("" : "+v"(acc[i][j]), "+v"(acc[i][j + 1]), "+v"(acc[i][j + 2]), "+v"(acc[i][j + 3]));
}

template <int BM_, int BN_, int BK_, int TM_, int TN_>
struct Fp32Cfg {
    static constexpr int BM = BM_, BN = BN_, BK = BK_, TM = TM_, TN = TN_;
    static constexpr int TX = BN / TN, TY = BM / TM;
    static constexpr int NT = TX * TY;
    static_assert(TM % 4 == 0 && TN % 4 == 0, "micro-tile built from 4-wide groups");
    static_assert(TX % 8 == 0 && TY % 8 == 0 && NT % 64 == 0, "waves are 8x8 thread grids");
    static constexpr int LDA = BM + LDS_PAD, LDB = BN + LDS_PAD;
};

// One workgroup computes the C tile at (m0, n0) of g; shared by the plain and
// the grouped (persistent) kernels. lds must hold 2*(BK*LDA + BK*LDB) floats.
template <class CF, bool TA, bool TB, int VEC>
__device__ __forceinline__ void fp32_tile(const int M, const int N, const int K, const float alpha,
                                          const float *__restrict__ A, const long lda,
                                          const float *__restrict__ B, const long ldb, const float beta,
                                          float *__restrict__ C, const long ldc, const int m0, const int n0,
                                          float *lds) {
    constexpr int BM = CF::BM, BN = CF::BN, BK = CF::BK, TM = CF::TM, TN = CF::TN, NT = CF::NT;
    constexpr int LDA = CF::LDA, LDB = CF::LDB;
    float *As = lds;                    // [2][BK][LDA]
    float *Bs = lds + 2 * BK * LDA;     // [2][BK][LDB]

    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    constexpr int WX = CF::TX / 8;
    const int tx = (lane & 7) + 8 * (w % WX);
    const int ty = (lane >> 3) + 8 * (w / WX);

    // A tile: TA -> stored K x M (contiguous m); else M x K (contiguous k).
    using ARegs = TileRegs<float, TA ? BM : BK, TA ? BK : BM, NT, VEC>;
    using BRegs = TileRegs<float, TB ? BK : BN, TB ? BN : BK, NT, VEC>;
    ARegs ra;
    BRegs rb;

    const bool interior = (m0 + BM <= M) && (n0 + BN <= N);
    const int ktiles = (K + BK - 1) / BK;
    const int kfull = K / BK;   // tiles with no K tail

    auto load_tile = [&](int kt) {
        const int k0 = kt * BK;
        const float *a0 = TA ? A + (long)k0 * lda + m0 : A + (long)m0 * lda + k0;
        const float *b0 = TB ? B + (long)n0 * ldb + k0 : B + (long)k0 * ldb + n0;
        if (interior && kt < kfull) {
            ra.template load<false>(a0, lda, t, 0, 0);
            rb.template load<false>(b0, ldb, t, 0, 0);
        } else {
            const int mr = M - m0, nr = N - n0, kr = K - k0;
            ra.template load<true>(a0, lda, t, TA ? mr : kr, TA ? kr : mr);
            rb.template load<true>(b0, ldb, t, TB ? kr : nr, TB ? nr : kr);
        }
    };
    auto store_tile = [&](int buf) {
        ra.template store<TA, LDA>(As + buf * BK * LDA, t);
        rb.template store<!TB, LDB>(Bs + buf * BK * LDB, t);
    };

    float acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = 0.0f;

    if (ktiles > 0) {
        load_tile(0);
        store_tile(0);
    }
    __syncthreads();

    for (int kt = 0; kt < ktiles; ++kt) {
        const int buf = kt & 1;
        const bool more = kt + 1 < ktiles;
        if (more) load_tile(kt + 1);
        const float *as = As + buf * BK * LDA + ty * 4;
        const float *bs = Bs + buf * BK * LDB + tx * 4;
        // Register double-buffering of the k-step fragments: the LDS reads
        // of step k+1 are issued before the FMAs of step k. pin_step ends
        // each k-step: without it the compiler sinks all FMAs of the tile
        // below all of its LDS reads (4x the registers, spills).
        float a[2][TM], b[2][TN];
        auto read_frag = [&](int k, float *fa, float *fb) {
#pragma unroll
            for (int g = 0; g < TM / 4; ++g) {
                float4 x = *reinterpret_cast<const float4 *>(&as[k * LDA + g * CF::TY * 4]);
                fa[g * 4 + 0] = x.x; fa[g * 4 + 1] = x.y; fa[g * 4 + 2] = x.z; fa[g * 4 + 3] = x.w;
            }
#pragma unroll
            for (int g = 0; g < TN / 4; ++g) {
                float4 x = *reinterpret_cast<const float4 *>(&bs[k * LDB + g * CF::TX * 4]);
                fb[g * 4 + 0] = x.x; fb[g * 4 + 1] = x.y; fb[g * 4 + 2] = x.z; fb[g * 4 + 3] = x.w;
            }
        };
        read_frag(0, a[0], b[0]);
#pragma unroll
        for (int k = 0; k < BK; ++k) {
            if (k + 1 < BK) read_frag(k + 1, a[(k + 1) & 1], b[(k + 1) & 1]);
            const float *fa = a[k & 1], *fb = b[k & 1];
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) acc[i][j] = fmaf(fa[i], fb[j], acc[i][j]);
            pin_step(acc);
        }
        if (more) store_tile(buf ^ 1);
        __syncthreads();
    }

    // Epilogue: C = alpha * acc (+ beta * C).
    const bool vec_c = interior && ((ldc & 3) == 0) && ((((uintptr_t)C) & 15) == 0);
#pragma unroll
    for (int i = 0; i < TM; ++i) {
        const int m = m0 + (i / 4) * CF::TY * 4 + ty * 4 + (i % 4);
#pragma unroll
        for (int g = 0; g < TN / 4; ++g) {
            const int n = n0 + g * CF::TX * 4 + tx * 4;
            float *cp = C + (long)m * ldc + n;
            if (vec_c) {
                float4 o;
                if (beta == 0.0f) {
                    o = make_float4(alpha * acc[i][g * 4 + 0], alpha * acc[i][g * 4 + 1],
                                    alpha * acc[i][g * 4 + 2], alpha * acc[i][g * 4 + 3]);
                } else {
                    const float4 c = *reinterpret_cast<const float4 *>(cp);
                    o = make_float4(fmaf(alpha, acc[i][g * 4 + 0], beta * c.x), fmaf(alpha, acc[i][g * 4 + 1], beta * c.y),
                                    fmaf(alpha, acc[i][g * 4 + 2], beta * c.z), fmaf(alpha, acc[i][g * 4 + 3], beta * c.w));
                }
                *reinterpret_cast<float4 *>(cp) = o;
            } else if (m < M) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (n + j < N) {
                        const float r = alpha * acc[i][g * 4 + j];
                        cp[j] = (beta == 0.0f) ? r : fmaf(alpha, acc[i][g * 4 + j], beta * cp[j]);
                    }
                }
            }
        }
    }
}

template <class CF, bool TA, bool TB, int VEC>
__global__ __launch_bounds__(CF::NT) void sgemm_fp32_kernel(int M, int N, int K, float alpha,
                                                            const float *__restrict__ A, long lda,
                                                            const float *__restrict__ B, long ldb, float beta,
                                                            float *__restrict__ C, long ldc, int tiles_m, int tiles_n) {
    __shared__ __attribute__((aligned(16))) float lds[2 * CF::BK * (CF::LDA + CF::LDB)];
    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
    fp32_tile<CF, TA, TB, VEC>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tm * CF::BM, tn * CF::BN, lds);
}

// ---------------------------------------------------------------------------
// Variant table.
using V0 = Fp32Cfg<128, 128, 16, 8, 8>;
using V1 = Fp32Cfg<64, 64, 16, 4, 4>;
using V2 = Fp32Cfg<128, 64, 16, 8, 4>;
using V3 = Fp32Cfg<64, 128, 16, 4, 8>;

const char *kNames[] = {"fp32_128x128x16_t8x8", "fp32_64x64x16_t4x4", "fp32_128x64x16_t8x4",
                        "fp32_64x128x16_t4x8"};
constexpr int kNumVariants = 4;

template <class CF>
int launch_cfg(const Gemm<float> &g, hipStream_t s) {
    const int tiles_m = (g.M + CF::BM - 1) / CF::BM;
    const int tiles_n = (g.N + CF::BN - 1) / CF::BN;
    const long nblk = (long)tiles_m * tiles_n;
    if (nblk > 0x7fffffffL) return fail(UGEMM_ERR_ARG, "sgemm: too many tiles (%ld)", nblk);
    // 128-bit loads need 16-byte aligned bases and leading dimensions that are multiples of 4.
    const bool vec = ugemm_aligned16(g.A) && ugemm_aligned16(g.B) && (g.lda % 4 == 0) && (g.ldb % 4 == 0);
    dim3 grid((unsigned)nblk), block(CF::NT);
#define UGEMM_L(TA, TB, VEC)                                                                             \
    hipLaunchKernelGGL((sgemm_fp32_kernel<CF, TA, TB, VEC>), grid, block, 0, s, g.M, g.N, g.K, g.alpha, g.A, \
                       g.lda, g.B, g.ldb, g.beta, g.C, g.ldc, tiles_m, tiles_n)
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
#undef UGEMM_L
    UGEMM_HIP_TRY(hipGetLastError());
    return UGEMM_OK;
}

}  // namespace

int fp32_variant_count() { return kNumVariants; }
const char *fp32_variant_name(int v) { return (v >= 0 && v < kNumVariants) ? kNames[v] : nullptr; }

int fp32_pick_variant(const Gemm<float> &g) {
    // Enough 128x128 tiles to give every CU a few blocks: use them; otherwise
    // shrink the tile so that the grid still covers the chip.
    const long t128 = (long)((g.M + 127) / 128) * ((g.N + 127) / 128);
    const int cus = device_cu_count();
    if (t128 >= 2L * cus) return 0;
    const long t64 = (long)((g.M + 63) / 64) * ((g.N + 63) / 64);
    if (g.M >= 4 * g.N && t64 >= cus) return 2;
    if (g.N >= 4 * g.M && t64 >= cus) return 3;
    return 1;
}

int launch_fp32(const Gemm<float> &g, int variant, hipStream_t s) {
    switch (variant) {
        case 0: return launch_cfg<V0>(g, s);
        case 1: return launch_cfg<V1>(g, s);
        case 2: return launch_cfg<V2>(g, s);
        case 3: return launch_cfg<V3>(g, s);
        default: return fail(UGEMM_ERR_ARG, "fp32: no variant %d", variant);
    }
}

}  // namespace ugemm
