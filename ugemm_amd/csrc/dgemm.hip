// DGEMM for gfx950 on FP64 MFMA (v_mfma_f64_16x16x4_f64), FP64 accumulation.
//
// One workgroup computes a BM x BN tile of C with 4 waves (2 x 2); each wave
// owns a (BM/2) x (BN/2) grid of 16x16 fragments.  Operand tiles of BK = 16
// move global -> registers -> LDS with the layout-aware loader shared with
// the FP32 kernel (transposed layouts are turned around while writing LDS),
// two LDS stages, one barrier per BK step.  LDS holds both operands K-major
// (As[k][m], Bs[k][n]) so that a fragment read -- lane l takes row l%16 of
// k-slice l/16 -- is one conflict-free ds_read_b64 over 128 contiguous bytes
// per 16 lanes.  The D fragment of the FP64 MFMA holds rows l/16 + 4*i,
// column l%16 (measured on gfx950: unlike the FP32-output forms, whose rows
// are 4*(l/16) + i).
//
// Determinism: every element of C is a fixed sequence of MFMAs over k, and a
// split-K call adds its partial slices in a fixed order (no atomics), so
// repeated calls are bitwise identical.
// Semantics as dgemm_cpu (ugemm.h:162-286) with the BLAS beta == 0 rule.
#include "internal.h"
#include "kernel_common.h"

namespace ugemm {
namespace {

typedef double doublex4 __attribute__((ext_vector_type(4)));

template <int BM_, int BN_>
struct DCfg {
    static constexpr int BM = BM_, BN = BN_, BK = 16, NT = 256;
    static constexpr int WTM = BM / 2, WTN = BN / 2, FM = WTM / 16, FN = WTN / 16;
    static constexpr int LDA = BM + 2, LDB = BN + 2;   // doubles; +2 keeps 16-byte alignment
};

// Split-K (gridDim.y > 1): block (tile, z) multiplies K columns
// [z*kc, min(K, (z+1)*kc)) into the partial slice ws + z*M*N (row length N).
template <class CF, bool TA, bool TB, int VEC>
__global__ __launch_bounds__(CF::NT) __attribute__((amdgpu_waves_per_eu(2))) void dgemm_mfma_kernel(
    int M, int N, int K_, double alpha_, const double *__restrict__ A_, long lda, const double *__restrict__ B_,
    long ldb, double beta_, double *__restrict__ C_, long ldc_, int tiles_m, int tiles_n, int kc,
    double *__restrict__ ws) {
    const bool part = gridDim.y > 1;
    const int kz0 = part ? (int)blockIdx.y * kc : 0;
    const int K = part ? min(kc, K_ - kz0) : K_;
    const double *__restrict__ A = TA ? A_ + (long)kz0 * lda : A_ + kz0;
    const double *__restrict__ B = TB ? B_ + kz0 : B_ + (long)kz0 * ldb;
    double *__restrict__ C = part ? ws + (long)blockIdx.y * M * N : C_;
    const long ldc = part ? (long)N : ldc_;
    const double alpha = part ? 1.0 : alpha_, beta = part ? 0.0 : beta_;
    constexpr int BM = CF::BM, BN = CF::BN, BK = CF::BK, NT = CF::NT, FM = CF::FM, FN = CF::FN;
    constexpr int LDA = CF::LDA, LDB = CF::LDB;
    __shared__ __attribute__((aligned(16))) double lds[2 * BK * (LDA + LDB)];
    double *As = lds, *Bs = lds + 2 * BK * LDA;

    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w >> 1, wn = w & 1;

    using ARegs = TileRegs<double, TA ? BM : BK, TA ? BK : BM, NT, VEC>;
    using BRegs = TileRegs<double, TB ? BK : BN, TB ? BN : BK, NT, VEC>;
    ARegs ra;
    BRegs rb;
    const bool interior = (m0 + BM <= M) && (n0 + BN <= N);
    const int ktiles = (K + BK - 1) / BK, kfull = K / BK;

    auto load_tile = [&](int kt) {
        const int k0 = kt * BK;
        const double *a0 = TA ? A + (long)k0 * lda + m0 : A + (long)m0 * lda + k0;
        const double *b0 = TB ? B + (long)n0 * ldb + k0 : B + (long)k0 * ldb + n0;
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

    doublex4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = doublex4{0.0, 0.0, 0.0, 0.0};

    if (ktiles > 0) {
        load_tile(0);
        store_tile(0);
    }
    __syncthreads();

    const int frow = lane & 15, fk = lane >> 4;
    for (int kt = 0; kt < ktiles; ++kt) {
        const int buf = kt & 1;
        const bool more = kt + 1 < ktiles;
        if (more) load_tile(kt + 1);
        const double *as = As + buf * BK * LDA + fk * LDA + wm * CF::WTM + frow;
        const double *bs = Bs + buf * BK * LDB + fk * LDB + wn * CF::WTN + frow;
#pragma unroll
        for (int s = 0; s < BK / 4; ++s) {
            double fa[FM], fb[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) fa[i] = as[s * 4 * LDA + i * 16];
#pragma unroll
            for (int j = 0; j < FN; ++j) fb[j] = bs[s * 4 * LDB + j * 16];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
        if (more) store_tile(buf ^ 1);
        __syncthreads();
    }

    // C values of one fragment row are loaded before any is stored (a load
    // cannot pass a possibly aliasing store: one latency per row, not per element).
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        const int mb = m0 + wm * CF::WTM + i * 16 + (lane >> 4);
        double cv[FN][4];
        if (beta != 0.0) {
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int n = n0 + wn * CF::WTN + j * 16 + (lane & 15);
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    cv[j][r] = (interior || (mb + 4 * r < M && n < N)) ? C[(long)(mb + 4 * r) * ldc + n] : 0.0;
            }
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int n = n0 + wn * CF::WTN + j * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = mb + 4 * r;
                if (interior || (m < M && n < N))
                    C[(long)m * ldc + n] = (beta == 0.0) ? alpha * acc[i][j][r]
                                                         : fma(alpha, acc[i][j][r], beta * cv[j][r]);
            }
        }
    }
}

using D0 = DCfg<128, 128>;
using D1 = DCfg<64, 64>;
using D2 = DCfg<128, 64>;
using D3 = DCfg<64, 128>;
const char *kNames[] = {"f64_mfma_128x128x16", "f64_mfma_64x64x16", "f64_mfma_128x64x16", "f64_mfma_64x128x16"};
constexpr int kNum = 4;

template <class CF>
int launch_cfg(const Gemm<double> &g, hipStream_t s) {
    const int tiles_m = (g.M + CF::BM - 1) / CF::BM, tiles_n = (g.N + CF::BN - 1) / CF::BN;
    const long nblk = (long)tiles_m * tiles_n;
    if (nblk > 0x7fffffffL) return fail(UGEMM_ERR_ARG, "dgemm: too many tiles (%ld)", nblk);
    const bool vec = ugemm_aligned16(g.A) && ugemm_aligned16(g.B) && (g.lda % 2 == 0) && (g.ldb % 2 == 0);
    int kc = g.K;
    const int splits = split_count(g.splits, nblk, g.K, 64, &kc, 2);
    double *ws = nullptr;
    if (splits > 1) {
        const int rc = stream_alloc((void **)&ws, (size_t)splits * g.M * g.N * sizeof(double), s);
        if (rc) return rc;
    }
    dim3 grid((unsigned)nblk, (unsigned)splits), block(CF::NT);
#define UGEMM_L(TA, TB, VEC)                                                                                \
    hipLaunchKernelGGL((dgemm_mfma_kernel<CF, TA, TB, VEC>), grid, block, 0, s, g.M, g.N, g.K, g.alpha, g.A, \
                       g.lda, g.B, g.ldb, g.beta, g.C, g.ldc, tiles_m, tiles_n, kc, ws)
    timing_begin(s);
    if (vec) {
        if (!g.ta && !g.tb) UGEMM_L(false, false, 2);
        else if (!g.ta && g.tb) UGEMM_L(false, true, 2);
        else if (g.ta && !g.tb) UGEMM_L(true, false, 2);
        else UGEMM_L(true, true, 2);
    } else {
        if (!g.ta && !g.tb) UGEMM_L(false, false, 1);
        else if (!g.ta && g.tb) UGEMM_L(false, true, 1);
        else if (g.ta && !g.tb) UGEMM_L(true, false, 1);
        else UGEMM_L(true, true, 1);
    }
    timing_end(s);
#undef UGEMM_L
    hipError_t e = hipGetLastError();
    int rc = e == hipSuccess ? UGEMM_OK : hip_fail(e, "dgemm launch");
    if (splits > 1) {
        if (!rc) rc = launch_splitk_reduce_f64(g.M, g.N, splits, g.alpha, ws, g.beta, g.C, g.ldc, s);
        const int rc2 = stream_free(ws, s);
        if (!rc) rc = rc2;
    }
    return rc;
}

}  // namespace

int dgemm_variant_count() { return kNum; }
const char *dgemm_variant_name(int v) { return (v >= 0 && v < kNum) ? kNames[v] : nullptr; }

int dgemm_pick_variant(const Gemm<double> &g) {
    const long cus = device_cu_count();
    const long t128 = (long)((g.M + 127) / 128) * ((g.N + 127) / 128);
    if (t128 >= 2 * cus) return 0;
    const long t64 = (long)((g.M + 63) / 64) * ((g.N + 63) / 64);
    if (g.M >= 4L * g.N && t64 >= cus) return 2;
    if (g.N >= 4L * g.M && t64 >= cus) return 3;
    return 1;
}

int launch_dgemm(const Gemm<double> &g, int variant, hipStream_t s) {
    switch (variant) {
        case 0: return launch_cfg<D0>(g, s);
        case 1: return launch_cfg<D1>(g, s);
        case 2: return launch_cfg<D2>(g, s);
        case 3: return launch_cfg<D3>(g, s);
        default: return fail(UGEMM_ERR_ARG, "dgemm: no variant %d", variant);
    }
}

}  // namespace ugemm
