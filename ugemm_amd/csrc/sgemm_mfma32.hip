// FP32 SGEMM on the FP32 matrix cores of gfx950 (v_mfma_f32_16x16x4_f32):
// FP32 operands as given, FP32 products, FP32 accumulation -- no operand
// split or conversion, so no preparation kernels and the error class of the
// plain FP32 kernel.  MI355X's FP32 MFMA peak (157.3 TF/s) equals its FP32
// vector peak, but the matrix core is fed one register per operand per 2048
// flop instead of one per 2 flop, so it runs much closer to that peak.
//
// A workgroup of WGM x WGN waves computes a BM x BN tile of C; each wave owns
// a (BM/WGM) x (BN/WGN) grid of 16x16 fragments.  Operand tiles of BK move
// global -> registers -> LDS, two LDS stages, one barrier per BK step.  Each
// operand keeps its global orientation in LDS, so every LDS store is a
// 16-byte vector store (no transposing writes):
//   k-contiguous operand (A when op(A) = A, B when op(B) = B^T): L[row][k],
//     row length BK + 4 floats; a lane reads its four k values as one
//     16-byte ds_read;
//   row-contiguous operand: L[k][row], row length rows + 4 floats; a lane
//     reads four single floats, 16 lanes over 64 contiguous bytes.
// The 16x16x4 MFMA takes from lane l the element (row l%16, k l/16).  Within
// a 16-deep k slab the lane's k group g = l/16 supplies k = 4g + q to the
// q-th MFMA (q = 0..3): both operands use this same permutation of k, so
// the sum over the slab is unchanged, and a lane's four k values are
// adjacent in memory for the k-contiguous layout.  The D fragment holds rows
// 4*(l/16) + r, column l%16.
//
// Split-K (gridDim.y > 1): block (tile, z) multiplies K columns
// [z*kc, min(K, (z+1)*kc)) into the partial slice ws + z*M*N; the partials
// are added in fixed order by the split-K reduction, which also applies
// alpha, beta and the epilogue.  Repeated calls are bitwise identical.
// Semantics as sgemm_cpu (ugemm.h:287-411) with the BLAS beta == 0 rule
// (C is not read), as gemm_cpu.h:119-123.
#include "internal.h"
#include "kernel_common.h"

namespace ugemm {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

template <int BM_, int BN_, int BK_, int WGM_, int WGN_, int MINB_ = 1>
struct M32Cfg {
    static constexpr int BM = BM_, BN = BN_, BK = BK_, WGM = WGM_, WGN = WGN_, MINB = MINB_;
    static constexpr int NT = 64 * WGM * WGN;
    static constexpr int WTM = BM / WGM, WTN = BN / WGN, FM = WTM / 16, FN = WTN / 16;
    static_assert(BK % 16 == 0 && WTM % 16 == 0 && WTN % 16 == 0, "fragment-aligned tile");
    // LDS row length and stage size (floats) of an operand tile of ROWS rows.
    template <bool KC, int ROWS>
    static constexpr int ld() { return KC ? BK + 4 : ROWS + 4; }
    template <bool KC, int ROWS>
    static constexpr int stage() { return KC ? ROWS * (BK + 4) : BK * (ROWS + 4); }
};

// The lane's four k values (k .. k+3) of fragment row `row`.
template <bool KC, int LD>
__device__ __forceinline__ floatx4 frag4(const float *L, int row, int k) {
    if constexpr (KC) {
        return *reinterpret_cast<const floatx4 *>(L + row * LD + k);
    } else {
        return floatx4{L[k * LD + row], L[(k + 1) * LD + row], L[(k + 2) * LD + row], L[(k + 3) * LD + row]};
    }
}

template <class CF, bool TA, bool TB, int VEC>
__global__ __launch_bounds__(CF::NT, CF::MINB) void sgemm_mfma32_kernel(
    int M, int N, int K_, float alpha_, const float *__restrict__ A_, long lda, const float *__restrict__ B_,
    long ldb, float beta_, float *__restrict__ C_, long ldc_, int tiles_m, int tiles_n, Epilogue epi, int kc,
    float *__restrict__ ws) {
    constexpr bool KCA = !TA, KCB = TB;
    constexpr int BM = CF::BM, BN = CF::BN, BK = CF::BK, NT = CF::NT, FM = CF::FM, FN = CF::FN;
    constexpr int LDA = CF::template ld<KCA, BM>(), LDB = CF::template ld<KCB, BN>();
    constexpr int SA = CF::template stage<KCA, BM>(), SB = CF::template stage<KCB, BN>();
    __shared__ __attribute__((aligned(16))) float lds[2 * (SA + SB)];
    float *As = lds, *Bs = lds + 2 * SA;

    const bool part = gridDim.y > 1;
    const int kz0 = part ? (int)blockIdx.y * kc : 0;
    const int K = part ? min(kc, K_ - kz0) : K_;
    const float *__restrict__ A = TA ? A_ + (long)kz0 * lda : A_ + kz0;
    const float *__restrict__ B = TB ? B_ + kz0 : B_ + (long)kz0 * ldb;
    float *__restrict__ C = part ? ws + (long)blockIdx.y * M * N : C_;
    const long ldc = part ? (long)N : ldc_;
    const float alpha = part ? 1.0f : alpha_, beta = part ? 0.0f : beta_;
    const bool has_epi = !part && epi.active();

    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w / CF::WGN, wn = w % CF::WGN;

    using ARegs = TileRegs<float, KCA ? BK : BM, KCA ? BM : BK, NT, VEC>;
    using BRegs = TileRegs<float, KCB ? BK : BN, KCB ? BN : BK, NT, VEC>;
    ARegs ra;
    BRegs rb;
    const bool interior = (m0 + BM <= M) && (n0 + BN <= N);
    const int ktiles = (K + BK - 1) / BK, kfull = K / BK;

    auto load_tile = [&](int kt) {
        const int k0 = kt * BK;
        const float *a0 = TA ? A + (long)k0 * lda + m0 : A + (long)m0 * lda + k0;
        const float *b0 = TB ? B + (long)n0 * ldb + k0 : B + (long)k0 * ldb + n0;
        if (interior && kt < kfull) {
            ra.template load<false>(a0, lda, t, 0, 0);
            rb.template load<false>(b0, ldb, t, 0, 0);
        } else {
            const int mr = M - m0, nr = N - n0, kr = K - k0;
            ra.template load<true>(a0, lda, t, KCA ? kr : mr, KCA ? mr : kr);
            rb.template load<true>(b0, ldb, t, KCB ? kr : nr, KCB ? nr : kr);
        }
    };
    auto store_tile = [&](int buf) {
        ra.template store<true, LDA>(As + buf * SA, t);
        rb.template store<true, LDB>(Bs + buf * SB, t);
    };

    floatx4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

    if (ktiles > 0) {
        load_tile(0);
        store_tile(0);
    }
    __syncthreads();

    const int frow = lane & 15, g4 = (lane >> 4) * 4;
    for (int kt = 0; kt < ktiles; ++kt) {
        const int buf = kt & 1;
        const bool more = kt + 1 < ktiles;
        if (more) load_tile(kt + 1);
        const float *as = As + buf * SA, *bs = Bs + buf * SB;
#pragma unroll
        for (int s = 0; s < BK / 16; ++s) {
            floatx4 fa[FM], fb[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) fa[i] = frag4<KCA, LDA>(as, wm * CF::WTM + i * 16 + frow, s * 16 + g4);
#pragma unroll
            for (int j = 0; j < FN; ++j) fb[j] = frag4<KCB, LDB>(bs, wn * CF::WTN + j * 16 + frow, s * 16 + g4);
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int j = 0; j < FN; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[i][q], fb[j][q], acc[i][j], 0, 0, 0);
        }
        if (more) store_tile(buf ^ 1);
        __syncthreads();
    }

    // C values of one fragment row are loaded before any is stored (a load
    // cannot pass a possibly aliasing store: one latency per row, not per element).
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        const int mb = m0 + wm * CF::WTM + i * 16 + 4 * (lane >> 4);
        float cv[FN][4];
        if (beta != 0.0f) {
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int n = n0 + wn * CF::WTN + j * 16 + (lane & 15);
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    cv[j][r] = (interior || (mb + r < M && n < N)) ? C[(long)(mb + r) * ldc + n] : 0.0f;
            }
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int n = n0 + wn * CF::WTN + j * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = mb + r;
                if (interior || (m < M && n < N)) {
                    float v = (beta == 0.0f) ? alpha * acc[i][j][r] : fmaf(alpha, acc[i][j][r], beta * cv[j][r]);
                    if (has_epi) v = apply_epilogue(epi, v, m, n);
                    C[(long)m * ldc + n] = v;
                }
            }
        }
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
