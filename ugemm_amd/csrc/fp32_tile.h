// The plain FP32 tile computation, shared by the single-GEMM kernel
// (sgemm_fp32.hip) and the grouped persistent kernel (grouped.hip).
//
// One workgroup computes one BM x BN tile of C:
//   * global -> registers -> LDS, two LDS stages: tile k+1 is fetched into
//     registers while tile k is multiplied from LDS, one barrier per BK step;
//   * LDS holds both operands K-major: As[k][m], Bs[k][n]; the transposed
//     layouts are turned around while writing LDS (no separate transpose pass);
//   * each thread owns a TM x TN register micro-tile built from 4-wide groups,
//     read from LDS with 128-bit loads; a wave is an 8 x 8 grid of threads, so a
//     k-step reads 128 contiguous bytes of As and of Bs per wave.
// Semantics follow the reference's row-major kernel gemm_rnn (sgemm_ocl2.h:17-90):
// C = alpha * acc + beta * C, C is not read when beta == 0 (sgemm_ocl2.h:86-88).
#pragma once

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

// MINB_: workgroups per CU the kernel is compiled for (register budget).
template <int BM_, int BN_, int BK_, int TM_, int TN_, int MINB_ = 1>
struct Fp32Cfg {
    static constexpr int BM = BM_, BN = BN_, BK = BK_, TM = TM_, TN = TN_, MINB = MINB_;
    static constexpr int TX = BN / TN, TY = BM / TM;
    static constexpr int NT = TX * TY;
    static_assert(TM % 4 == 0 && TN % 4 == 0, "micro-tile built from 4-wide groups");
    static_assert(TX % 8 == 0 && TY % 8 == 0 && NT % 64 == 0, "waves are 8x8 thread grids");
    static constexpr int LDA = BM + LDS_PAD, LDB = BN + LDS_PAD;
};

// One workgroup computes the C tile at (m0, n0) of g; shared by the plain and
// the grouped (persistent) kernels. lds must hold 2*(BK*LDA + BK*LDB) floats.
// aligned == false (with VEC == 4) reads A and B with scalar loads: the
// grouped kernel mixes problems whose operands allow 128-bit loads with
// problems whose operands do not.
// CONV: opB is the implicit im2col matrix of cd (B and TB unused): every
// thread keeps one column n of the B tile, so the pixel's origin is computed
// once per tile and each element costs one table read and two compares.
// epi: bias and activation after alpha/beta (Epilogue, internal.h).
template <class CF, bool TA, bool TB, int VEC, bool CONV = false>
__device__ __forceinline__ void fp32_tile(const int M, const int N, const int K, const float alpha,
                                          const float *__restrict__ A, const long lda,
                                          const float *__restrict__ B, const long ldb, const float beta,
                                          float *__restrict__ C, const long ldc, const int m0, const int n0,
                                          float *lds, const bool aligned = true, const Epilogue &epi = Epilogue(),
                                          const ConvDesc &cd = ConvDesc()) {
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

    // Implicit im2col: thread t always holds column nl = t % BN of the B
    // tile (NT is a multiple of BN) and rows kl = t / BN + i * (NT / BN).
    static_assert(!CONV || (NT % BN == 0 && BN % 64 == 0 && !TB), "implicit operand layout");
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
        if (interior && kt < kfull && aligned) ra.template load<false>(a0, lda, t, 0, 0);
        else ra.template load<true>(a0, lda, t, TA ? M - m0 : K - k0, TA ? K - k0 : M - m0, aligned);
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
            if (interior && kt < kfull && aligned) rb.template load<false>(b0, ldb, t, 0, 0);
            else rb.template load<true>(b0, ldb, t, TB ? K - k0 : N - n0, TB ? N - n0 : K - k0, aligned);
        }
    };
    auto store_tile = [&](int buf) {
        ra.template store<TA, LDA>(As + buf * BK * LDA, t);
        if constexpr (CONV) {
#pragma unroll
            for (int i = 0; i < CONV_N; ++i) Bs[buf * BK * LDB + (t / BN + i * (NT / BN)) * LDB + t % BN] = cvals[i];
        } else {
            rb.template store<!TB, LDB>(Bs + buf * BK * LDB, t);
        }
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

    // Epilogue: C = alpha * acc (+ beta * C), then bias and activation.
    const bool vec_c = interior && ((ldc & 3) == 0) && ((((uintptr_t)C) & 15) == 0);
    const bool has_epi = epi.active();
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
                if (has_epi) {
                    o.x = apply_epilogue(epi, o.x, m, n);
                    o.y = apply_epilogue(epi, o.y, m, n + 1);
                    o.z = apply_epilogue(epi, o.z, m, n + 2);
                    o.w = apply_epilogue(epi, o.w, m, n + 3);
                }
                *reinterpret_cast<float4 *>(cp) = o;
            } else if (m < M) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (n + j < N) {
                        float r = (beta == 0.0f) ? alpha * acc[i][g * 4 + j] : fmaf(alpha, acc[i][g * 4 + j], beta * cp[j]);
                        if (has_epi) r = apply_epilogue(epi, r, m, n + j);
                        cp[j] = r;
                    }
                }
            }
        }
    }
}

}  // namespace
}  // namespace ugemm
