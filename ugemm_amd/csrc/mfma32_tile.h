// The FP32-MFMA tile computation (v_mfma_f32_16x16x4_f32), shared by the
// single-GEMM kernel (sgemm_mfma32.hip) and the grouped persistent kernel
// (grouped.hip).  FP32 operands as given, FP32 products, FP32 accumulation.
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
// Semantics as sgemm_cpu (ugemm.h:287-411) with the BLAS beta == 0 rule
// (C is not read), as gemm_cpu.h:119-123.
#pragma once

#include <type_traits>

#include "internal.h"
#include "kernel_common.h"

namespace ugemm {
namespace {

typedef float m32_floatx4 __attribute__((ext_vector_type(4)));

// MINB_: workgroups per CU the kernel is compiled for (register budget).
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
    // LDS floats of the largest layout (both operands k-contiguous).
    static constexpr int kLdsFloats = 2 * (BM + BN) * (BK + 4);
};

template <class CF>
struct IsM32 : std::false_type {};
template <int BM, int BN, int BK, int WGM, int WGN, int MINB>
struct IsM32<M32Cfg<BM, BN, BK, WGM, WGN, MINB>> : std::true_type {};

// The lane's four k values (k .. k+3) of fragment row `row`.
template <bool KC, int LD>
__device__ __forceinline__ m32_floatx4 m32_frag4(const float *L, int row, int k) {
    if constexpr (KC) {
        return *reinterpret_cast<const m32_floatx4 *>(L + row * LD + k);
    } else {
        return m32_floatx4{L[k * LD + row], L[(k + 1) * LD + row], L[(k + 2) * LD + row], L[(k + 3) * LD + row]};
    }
}

// C[m0:m0+BM, n0:n0+BN] = alpha * op(A) op(B) + beta * C, then the epilogue.
// A and B point at the operands' origins (k = 0); vec_ok == false (VEC 4
// only): the operands are not 16-byte aligned, every chunk is read with
// scalar loads.  lds: CF::kLdsFloats floats, 16-byte aligned.
// CONV: op(B) is the implicit im2col matrix of cd (B and TB unused; the
// gather of sgemm_gl1.h:166-190): every thread keeps one column n of the B
// tile, so the pixel's origin is computed once per tile and each element
// costs one table read and two compares; the B stage is [k][n].
template <class CF, bool TA, bool TB, int VEC, bool CONV = false>
__device__ __forceinline__ void m32_tile(const int M, const int N, const int K, const float alpha,
                                         const float *__restrict__ A, const long lda, const float *__restrict__ B,
                                         const long ldb, const float beta, float *__restrict__ C, const long ldc,
                                         const int m0, const int n0, float *lds, const bool vec_ok = true,
                                         const Epilogue &epi = Epilogue(), const ConvDesc &cd = ConvDesc()) {
    constexpr bool KCA = !TA, KCB = TB;
    constexpr int BM = CF::BM, BN = CF::BN, BK = CF::BK, NT = CF::NT, FM = CF::FM, FN = CF::FN;
    constexpr int LDA = CF::template ld<KCA, BM>(), LDB = CF::template ld<KCB, BN>();
    constexpr int SA = CF::template stage<KCA, BM>(), SB = CF::template stage<KCB, BN>();
    static_assert(2 * (SA + SB) <= CF::kLdsFloats, "LDS budget");
    float *As = lds, *Bs = lds + 2 * SA;
    const bool has_epi = epi.active();

    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w / CF::WGN, wn = w % CF::WGN;

    using ARegs = TileRegs<float, KCA ? BK : BM, KCA ? BM : BK, NT, VEC>;
    using BRegs = TileRegs<float, KCB ? BK : BN, KCB ? BN : BK, NT, VEC>;
    ARegs ra;
    BRegs rb;
    const bool interior = (m0 + BM <= M) && (n0 + BN <= N);
    const bool fast = interior && vec_ok;
    const int ktiles = (K + BK - 1) / BK, kfull = K / BK;

    static_assert(!CONV || (!TB && NT % BN == 0 && BN % 64 == 0), "implicit operand layout");
    constexpr int CONV_N = CONV ? BK * BN / NT : 1;
    float cvals[CONV_N];
    long c_base = 0;
    int c_h0 = 0, c_w0 = 0;
    bool c_ncol = false;
    if constexpr (CONV) {
        const int n = n0 + t % BN;
        c_ncol = n < N;
        const int ho = c_ncol ? n / cd.wcol : 0, wo = c_ncol ? n - ho * cd.wcol : 0;
        c_h0 = ho * cd.stride - cd.pad;
        c_w0 = wo * cd.stride - cd.pad;
        c_base = (long)c_h0 * cd.W + c_w0;
    }

    auto load_tile = [&](int kt) {
        const int k0 = kt * BK;
        const float *a0 = TA ? A + (long)k0 * lda + m0 : A + (long)m0 * lda + k0;
        const int mr = M - m0, nr = N - n0, kr = K - k0;
        if (fast && kt < kfull) ra.template load<false>(a0, lda, t, 0, 0);
        else ra.template load<true>(a0, lda, t, KCA ? kr : mr, KCA ? mr : kr, vec_ok);
        if constexpr (CONV) {
#pragma unroll
            for (int i = 0; i < CONV_N; ++i) {
                // k is the same for the 64 lanes of a wave (BN is a multiple
                // of 64): one scalar load of the table entry per wave.
                const int k = __builtin_amdgcn_readfirstlane(k0 + t / BN + i * (NT / BN));
                float v = 0.0f;
                if (c_ncol && k < K) {
                    const int2 e = cd.ktab[k];
                    const int hi = c_h0 + (e.y >> 16), wi = c_w0 + (e.y & 0xffff);
                    if ((unsigned)hi < (unsigned)cd.H && (unsigned)wi < (unsigned)cd.W) v = cd.in[c_base + e.x];
                }
                cvals[i] = v;
            }
        } else {
            const float *b0 = TB ? B + (long)n0 * ldb + k0 : B + (long)k0 * ldb + n0;
            if (fast && kt < kfull) rb.template load<false>(b0, ldb, t, 0, 0);
            else rb.template load<true>(b0, ldb, t, KCB ? kr : nr, KCB ? nr : kr, vec_ok);
        }
    };
    auto store_tile = [&](int buf) {
        ra.template store<true, LDA>(As + buf * SA, t);
        if constexpr (CONV) {
#pragma unroll
            for (int i = 0; i < CONV_N; ++i) Bs[buf * SB + (t / BN + i * (NT / BN)) * LDB + t % BN] = cvals[i];
        } else {
            rb.template store<true, LDB>(Bs + buf * SB, t);
        }
    };

    m32_floatx4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = m32_floatx4{0.0f, 0.0f, 0.0f, 0.0f};

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
            m32_floatx4 fa[FM], fb[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) fa[i] = m32_frag4<KCA, LDA>(as, wm * CF::WTM + i * 16 + frow, s * 16 + g4);
#pragma unroll
            for (int j = 0; j < FN; ++j) fb[j] = m32_frag4<KCB, LDB>(bs, wn * CF::WTN + j * 16 + frow, s * 16 + g4);
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

}  // namespace
}  // namespace ugemm
