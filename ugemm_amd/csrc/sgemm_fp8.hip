// FP8-e4m3 SGEMM with one scale per 128 x 128 block (opt-in, reduced precision).
//
// Quantisation (per operand, HBM-bound preparation kernel): op(A) is cut into
// 128 (rows) x 128 (k) blocks, op(B) into 128 (n) x 128 (k) blocks.  Each
// block gets a power-of-two scale s = 2^e, e = ceil(log2(amax / 448)) (448 is
// the largest finite e4m3 value; e = 0 for an all-zero block), and every
// element becomes q = e4m3_rne(x / s) (subnormals kept, |x/s| <= 448 so no
// saturation is needed).  Power-of-two scales are exact and are exactly what
// the gfx950 block-scaled MFMA takes (E8M0 exponents), so
//   C = alpha * sum_kb (sA[mb,kb] sB[nb,kb] * sum_{k in kb} qa qb) + beta * C
// is computed by v_mfma_scale_f32_16x16x128_f8f6f4 with no extra VALU work,
// accumulating in FP32.  Non-finite inputs give NaN (e4m3 has no infinity).
//
// Planes use the k-tile blocked layout of the split-BF16 path with 128-byte
// rows (one K tile = one scale block): byte (r, k) of an operand with rows_pad
// rows is at ((k / 128) * rows_pad + r) * 128 + ((k % 128) / 16 ^ fp8_swz(r)) * 16
// + k % 16.  The XOR makes the fragment reads conflict-free (see fp8_swz)
// and lets the core copy a stage verbatim.
// Fragment layout of the 16x16x128 fp8 MFMA: lane l holds row l%16 and k
// 32*(l/16) .. +31 of A, the same k of column l%16 of B; the per-lane scale
// operand is the E8M0 exponent of the block those elements belong to (one
// value per wave and K tile here).  D holds rows 4*(l/16)+i, column l%16.
#include "internal.h"
#include "kernel_common.h"

namespace ugemm {
namespace {

typedef int int8v __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int SB = 128;   // scale block edge = K tile

// Chunk swizzle of row r: chunk c of a 128-byte plane row is stored at
// c ^ fp8_swz(r).  A fragment read serves 16 lanes at a time as four rows x
// four lane quarters (lanes l..l+3, l+16.., l+32.., l+48..) over 64 banks:
// every quarter reads chunk 2q (+1), so the two even (and the two odd) rows
// of each group of four must XOR with values of opposite parity to land in
// distinct 16-byte bank groups.  r & 7 alone gave rows 4a and 4a + 2 the same
// parity (measured: SQ_LDS_BANK_CONFLICT about 3 cycles per LDS instruction).
__device__ __forceinline__ int fp8_swz(int r) { return (r & 7) ^ ((r >> 1) & 1); }

__device__ __forceinline__ long fp8_blocked_offset(int r, int k, int rows_pad) {
    return ((long)(k >> 7) * rows_pad + r) * 128 + ((((k >> 4) & 7) ^ fp8_swz(r)) << 4);
}

// Power-of-two exponent for a block maximum (see the header comment).
__device__ __forceinline__ int block_exponent(float amax) {
    if (!(amax > 0.0f) || !__builtin_isfinite(amax)) return 0;
    int e;
    const float f = frexpf(amax / 448.0f, &e);   // amax/448 = f * 2^e, f in [0.5, 1)
    if (f == 0.5f) e -= 1;                        // exact power of two: 2^(e-1)
    return e < -127 ? -127 : (e > 127 ? 127 : e);
}

// Packs 4 floats (already divided by the scale) into 4 e4m3 bytes, RNE.
__device__ __forceinline__ int pack4_e4m3(float a, float b, float c, float d) {
    int w = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
    w = __builtin_amdgcn_cvt_pk_fp8_f32(c, d, w, true);
    return w;
}

// One operand to quantise: source element (r, k) is kcont ? X[r*ld+k] :
// X[k*ld+r]; zeros outside rows x K.  rblocks = rows_pad / 128 blocks of
// rows; exps[(r / 128) * kblocks + k / 128] receives each block's exponent.
struct Fp8Operand {
    const float *X;
    long ld;
    int rows, K, vec, kcont;
    unsigned char *out;
    int rows_pad, rblocks;
    int *exps;
};

// LDS of one 128 x 128 block: k-contiguous sources quantise into byte rows
// [r][k] (pitch 144 bytes); row-contiguous ones into words [k][r/4] of four
// r-consecutive bytes (pitch 34 words: the 32 lanes of a word store and the
// two 16-line-apart rows a wave reads back land in different banks), and
// every output chunk is assembled by 4 x 4 byte transposes of 16 words.
constexpr int kF8Pitch = 128 + 16, kF8WPitch = 34;
constexpr int kF8Lds = 128 * (kF8Pitch > 4 * kF8WPitch ? kF8Pitch : 4 * kF8WPitch);

template <bool KCONT>
__device__ __forceinline__ void fp8_quantise_block(const Fp8Operand &op, int rb, int kb, int kblocks,
                                                   unsigned char *smem, float *wmax) {
    const int r0 = rb * SB, k0 = kb * SB;
    const int t = threadIdx.x;
    const float *__restrict__ X = op.X;
    const long ld = op.ld;
    const int rows = op.rows, K = op.K;
    // The block's source is 128 lines (rows for KCONT, k for the other
    // layout) of 128 contiguous floats.  Load q of a thread takes line
    // q*8 + t/32, floats (t%32)*4 .. +3: each load instruction of a wave
    // covers two whole 512-byte lines (coalesced), and the 16 loads are
    // issued before any is used.
    float v[64];
    const int cl = (t & 31) * 4, sl = t >> 5;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int line = q * 8 + sl;
        const int r = KCONT ? r0 + line : r0 + cl, k = KCONT ? k0 + cl : k0 + line;
        const float *p = KCONT ? X + (long)r * ld + k : X + (long)k * ld + r;
        const bool full = KCONT ? (r < rows && k + 4 <= K) : (k < K && r + 4 <= rows);
        if (full && op.vec) {
            const float4 x = *reinterpret_cast<const float4 *>(p);
            v[q * 4 + 0] = x.x; v[q * 4 + 1] = x.y; v[q * 4 + 2] = x.z; v[q * 4 + 3] = x.w;
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool ok = KCONT ? (r < rows && k + j < K) : (k < K && r + j < rows);
                v[q * 4 + j] = ok ? p[j] : 0.0f;
            }
        }
    }
    float amax = 0.0f;
#pragma unroll
    for (int j = 0; j < 64; ++j) amax = fmaxf(amax, fabsf(v[j]));
    // NaN/Inf make the block non-finite: keep it visible in amax.
#pragma unroll
    for (int j = 0; j < 64; ++j)
        if (!__builtin_isfinite(v[j])) amax = __builtin_inff();
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) amax = fmaxf(amax, __shfl_xor(amax, off));
    if ((t & 63) == 0) wmax[t >> 6] = amax;
    __syncthreads();
    const float bmax = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
    const int e = block_exponent(bmax);
    const float inv = ldexpf(1.0f, -e);
    if (t == 0) op.exps[(long)rb * kblocks + kb] = e;
    // Quantise into LDS; each packed word holds 4 consecutive elements of
    // its line (k for KCONT, r otherwise).
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int w = pack4_e4m3(v[q * 4] * inv, v[q * 4 + 1] * inv, v[q * 4 + 2] * inv, v[q * 4 + 3] * inv);
        const int line = q * 8 + sl;
        if (KCONT) *reinterpret_cast<int *>(smem + line * kF8Pitch + cl) = w;
        else reinterpret_cast<int *>(smem)[line * kF8WPitch + (t & 31)] = w;
    }
    __syncthreads();
    if (KCONT) {
        // 128 rows x 8 chunks of 16 bytes = 1024 chunks, 4 per thread.
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int f = t + i * 256, r = f >> 3, c = f & 7;
            const uint4 x = *reinterpret_cast<const uint4 *>(smem + r * kF8Pitch + c * 16);
            *reinterpret_cast<uint4 *>(op.out + fp8_blocked_offset(r0 + r, k0 + c * 16, op.rows_pad)) = x;
        }
    } else {
        // Thread t: rows 4*(t%32) .. +3, k chunk t/32 -- 16 words of 4 rows,
        // transposed 4 x 4 bytes at a time into the 4 rows' 16-byte chunks.
        const int rq = t & 31, kc = t >> 5;
        const int *wt = reinterpret_cast<const int *>(smem);
        unsigned a[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) a[i] = (unsigned)wt[(kc * 16 + i) * kF8WPitch + rq];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            unsigned o[4];
#pragma unroll
            for (int g = 0; g < 4; ++g)
                o[g] = ((a[4 * g] >> (8 * j)) & 0xffu) | (((a[4 * g + 1] >> (8 * j)) & 0xffu) << 8) |
                       (((a[4 * g + 2] >> (8 * j)) & 0xffu) << 16) | (((a[4 * g + 3] >> (8 * j)) & 0xffu) << 24);
            *reinterpret_cast<uint4 *>(op.out + fp8_blocked_offset(r0 + 4 * rq + j, k0 + kc * 16, op.rows_pad)) =
                make_uint4(o[0], o[1], o[2], o[3]);
        }
    }
}

// Both operands in one launch: blocks (x, y) with x < a.rblocks quantise
// block (x, y) of A, the others block (x - a.rblocks, y) of B.
__global__ __launch_bounds__(256) void fp8_prep_operands(Fp8Operand a, Fp8Operand b, int kblocks) {
    __shared__ float wmax[4];
    __shared__ __attribute__((aligned(16))) unsigned char smem[kF8Lds];
    const bool is_b = (int)blockIdx.x >= a.rblocks;
    const Fp8Operand &op = is_b ? b : a;
    const int rb = is_b ? (int)blockIdx.x - a.rblocks : (int)blockIdx.x;
    if (op.kcont) fp8_quantise_block<true>(op, rb, blockIdx.y, kblocks, smem, wmax);
    else fp8_quantise_block<false>(op, rb, blockIdx.y, kblocks, smem, wmax);
}

// ---------------------------------------------------------------------------
// Core: 256 x 256 tiles, 4 waves (2 x 2) of 128 x 128, K tile 128.
// Workgroup tile BM x BN x 128 with WGM x WGN waves; each wave owns a
// WTM x WTN block of C inside one 128 x 128 scale block.  W4: 4 waves of
// 128 x 128 (256 accumulators, one wave per SIMD).  W8: 8 waves of 128 x 64
// (two waves per SIMD, which cover each other's LDS and barrier waits), with
// the A fragments read one row-block ahead of their MFMAs and the stage copy
// interleaved with the last row-block's MFMAs (ILV; see the split core).
template <int BM_, int BN_, int WGM_, int WGN_, bool ILV_>
struct F8Cfg {
    static constexpr int BM = BM_, BN = BN_, BK = 128, WGM = WGM_, WGN = WGN_, NT = 64 * WGM_ * WGN_;
    static constexpr int WTM = BM / WGM, WTN = BN / WGN, FM = WTM / 16, FN = WTN / 16;
    static constexpr bool ILV = ILV_;
    static constexpr int STAGE = (BM + BN) * BK;                  // bytes
    static constexpr int LDS_BYTES = 2 * STAGE;
    static constexpr int PER_T = STAGE / 16 / NT;                 // 16-byte chunks per thread
    static_assert(SB % WTM == 0 || WTM % SB == 0, "a wave's rows lie in one scale block");
    static_assert(WTM <= SB && WTN <= SB, "one scale per wave and K tile");
};
using F8W4 = F8Cfg<256, 256, 2, 2, false>;
using F8W4Small = F8Cfg<128, 128, 2, 2, false>;
using F8W8 = F8Cfg<256, 256, 2, 4, true>;

template <class CF>
__global__ __launch_bounds__(CF::NT) void fp8_scaled_core(int M, int N, int Kp, const unsigned char *__restrict__ Aq,
                                                          int Mp, const int *__restrict__ ea,
                                                          const unsigned char *__restrict__ Bq, int Np,
                                                          const int *__restrict__ eb, float alpha, float beta,
                                                          float *__restrict__ C, long ldc, int tiles_m, int tiles_n) {
    constexpr int BM = CF::BM, BN = CF::BN, BK = CF::BK, FM = CF::FM, FN = CF::FN, PER_T = CF::PER_T;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem8[];
    int tm, tn;
    tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w / CF::WGN, wn = w % CF::WGN;
    const int kblocks = Kp / BK;

    // Stage kt: A rows m0.. (BM * 128 contiguous bytes), then B rows n0...
    constexpr int A_CH = BM * BK / 16 / CF::NT;
    const unsigned char *a_src = Aq + (long)m0 * BK + t * 16;
    const unsigned char *b_src = Bq + (long)n0 * BK + t * 16;
    // A native vector type (not HIP's uint4 struct), so that the array stays in registers.
    u32x4 regs[PER_T];
    auto load_chunk = [&](int kt, int i) {
        if (i < A_CH) regs[i] = *reinterpret_cast<const u32x4 *>(a_src + (long)kt * Mp * BK + i * CF::NT * 16);
        else regs[i] = *reinterpret_cast<const u32x4 *>(b_src + (long)kt * Np * BK + (i - A_CH) * CF::NT * 16);
    };
    auto load_stage = [&](int kt) {
#pragma unroll
        for (int i = 0; i < PER_T; ++i) load_chunk(kt, i);
    };
    auto store_stage = [&](int buf) {
        unsigned char *base = smem8 + buf * CF::STAGE + t * 16;
#pragma unroll
        for (int i = 0; i < PER_T; ++i) *reinterpret_cast<u32x4 *>(base + i * CF::NT * 16) = regs[i];
    };

    floatx4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    load_stage(0);
    store_stage(0);
    if constexpr (CF::ILV) load_stage(min(1, kblocks - 1));
    __syncthreads();
    const int frow = lane & 15, fq = lane >> 4;
    // Scale blocks of this wave: rows (m0 + wm*WTM) / 128, columns likewise.
    const int a_blk = (m0 + wm * CF::WTM) / SB, b_blk = (n0 + wn * CF::WTN) / SB;
    // Fragment of plane row `row` (16 rows x 128 k, 32 bytes per lane).
    auto frag = [&](const unsigned char *base, int row) {
        const u32x4 lo = *reinterpret_cast<const u32x4 *>(base + row * BK + (((2 * fq) ^ fp8_swz(row)) << 4));
        const u32x4 hi = *reinterpret_cast<const u32x4 *>(base + row * BK + (((2 * fq + 1) ^ fp8_swz(row)) << 4));
        return int8v{(int)lo.x, (int)lo.y, (int)lo.z, (int)lo.w, (int)hi.x, (int)hi.y, (int)hi.z, (int)hi.w};
    };
    // The scales of K tile kt + 1 are loaded at the top of step kt, before
    // that step's stage loads: the loads complete in order, so waiting for
    // a scale never waits for a stage load issued after it.  (Loaded in the
    // step that uses them, they made every step wait for all of the stage
    // loads of the step before.)
    const int *__restrict__ ea_w = ea + (long)a_blk * kblocks;
    const int *__restrict__ eb_w = eb + (long)b_blk * kblocks;
    int sa_next = ea_w[0], sb_next = eb_w[0];
    for (int kt = 0; kt < kblocks; ++kt) {
        const int buf = kt & 1;
        const bool more = kt + 1 < kblocks;
        const int sa = sa_next + 127, sb = sb_next + 127;
        sa_next = ea_w[min(kt + 1, kblocks - 1)];
        sb_next = eb_w[min(kt + 1, kblocks - 1)];
        if (!CF::ILV && more) load_stage(kt + 1);
        const unsigned char *as = smem8 + buf * CF::STAGE;
        const unsigned char *bs = as + BM * BK;
        int8v fb[FN];
#pragma unroll
        for (int j = 0; j < FN; ++j) fb[j] = frag(bs, wn * CF::WTN + j * 16 + frow);
        if constexpr (CF::ILV) {
            // Group i: the FN MFMAs of row-block i.  Before its MFMAs a group
            // reads the A fragment of row-block i + 2, and each group of the
            // second half takes its share of the stage copy (stage kt + 1
            // from registers into the idle buffer, then the loads of stage
            // kt + 2; branch-free: the last step stores into the idle buffer
            // and re-loads the last stage).  A scheduling barrier around each
            // group keeps this order.
            constexpr int LA = 2, H = FM / 2;
            int8v fa[FM];
#pragma unroll
            for (int i = 0; i < LA; ++i) fa[i] = frag(as, wm * CF::WTM + i * 16 + frow);
            static_for<FM>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                __builtin_amdgcn_sched_barrier(0);
                if constexpr (i + LA < FM) fa[i + LA] = frag(as, wm * CF::WTM + (i + LA) * 16 + frow);
                if constexpr (i >= H) {
                    static_for<PER_T>([&](auto cc) {
                        constexpr int c = decltype(cc)::value;
                        if constexpr (c * H / PER_T == i - H) {
                            *reinterpret_cast<u32x4 *>(smem8 + (buf ^ 1) * CF::STAGE + t * 16 + c * CF::NT * 16) =
                                regs[c];
                            load_chunk(min(kt + 2, kblocks - 1), c);
                        }
                    });
                }
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(fa[i], fb[j], acc[i][j], 0, 0, 0, sa,
                                                                                 0, sb);
            });
            __builtin_amdgcn_sched_barrier(0);
        } else {
            int8v fa[FM];
#pragma unroll
            for (int i = 0; i < FM; ++i) fa[i] = frag(as, wm * CF::WTM + i * 16 + frow);
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(fa[i], fb[j], acc[i][j], 0, 0, 0, sa,
                                                                                 0, sb);
            if (more) store_stage(buf ^ 1);
        }
        __syncthreads();
    }

    // C is read in batches of IB fragment rows (64 values per lane), all
    // loaded before any is stored (a load cannot pass a possibly aliasing
    // store: one latency per batch, not per element) -- unless 256
    // accumulators per lane leave no room (it spills).
    const bool interior = (m0 + BM <= M) && (n0 + BN <= N);
    constexpr bool BATCH = FM * FN * 4 < 256;
    constexpr int IB = !BATCH ? 1 : (64 / (4 * FN) > FM ? FM : (64 / (4 * FN) < 1 ? 1 : 64 / (4 * FN)));
    static_assert(FM % IB == 0, "whole batches of fragment rows");
#pragma unroll
    for (int i0 = 0; i0 < FM; i0 += IB) {
        float cv[IB][FN][4];
        if (BATCH && beta != 0.0f) {
#pragma unroll
            for (int ii = 0; ii < IB; ++ii) {
                const int mb = m0 + wm * CF::WTM + (i0 + ii) * 16 + (lane >> 4) * 4;
#pragma unroll
                for (int j = 0; j < FN; ++j) {
                    const int n = n0 + wn * CF::WTN + j * 16 + (lane & 15);
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        cv[ii][j][r] = (interior || (mb + r < M && n < N)) ? C[(long)(mb + r) * ldc + n] : 0.0f;
                }
            }
        }
#pragma unroll
        for (int ii = 0; ii < IB; ++ii) {
            const int i = i0 + ii;
            const int mb = m0 + wm * CF::WTM + i * 16 + (lane >> 4) * 4;
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int n = n0 + wn * CF::WTN + j * 16 + (lane & 15);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = mb + r;
                    if (interior || (m < M && n < N)) {
                        float *cp = C + (long)m * ldc + n;
                        if (beta == 0.0f) *cp = alpha * acc[i][j][r];
                        else *cp = fmaf(alpha, acc[i][j][r], beta * (BATCH ? cv[ii][j][r] : *cp));
                    }
                }
            }
        }
    }
}

long round_up(long x, long a) { return (x + a - 1) / a * a; }

Fp8Operand operand(const float *X, long ld, bool kcont, int rows, int K, unsigned char *out, int rows_pad,
                   int *exps) {
    return Fp8Operand{X, ld, rows, K, (ugemm_aligned16(X) && ld % 4 == 0) ? 1 : 0, kcont ? 1 : 0, out, rows_pad,
                      rows_pad / SB, exps};
}

template <class CF>
int run(const Gemm<float> &g, hipStream_t s) {
    const long Mp = round_up(g.M, CF::BM), Np = round_up(g.N, CF::BN), Kp = round_up(g.K, SB);
    const int tiles_m = (int)(Mp / CF::BM), tiles_n = (int)(Np / CF::BN);
    if ((long)tiles_m * tiles_n > 0x7fffffffL) return fail(UGEMM_ERR_ARG, "too many tiles");
    const long kb = Kp / SB;
    if (kb > 65535) return fail(UGEMM_ERR_UNSUPPORTED, "fp8: K = %d exceeds %d", g.K, 65535 * SB);
    const size_t a_bytes = (size_t)Mp * Kp, b_bytes = (size_t)Np * Kp;
    const size_t e_bytes = (size_t)((Mp + Np) / SB) * kb * sizeof(int);
    void *ws = nullptr;
    int rc = stream_alloc(&ws, a_bytes + b_bytes + e_bytes, s);
    if (rc) return rc;
    unsigned char *Aq = (unsigned char *)ws, *Bq = Aq + a_bytes;
    int *ea = (int *)(Bq + b_bytes), *eb = ea + (Mp / SB) * kb;
    {
        const Fp8Operand oa = operand(g.A, g.lda, !g.ta, g.M, g.K, Aq, (int)Mp, ea);
        const Fp8Operand ob = operand(g.B, g.ldb, g.tb, g.N, g.K, Bq, (int)Np, eb);
        hipLaunchKernelGGL(fp8_prep_operands, dim3((unsigned)(oa.rblocks + ob.rblocks), (unsigned)kb), dim3(256), 0, s,
                           oa, ob, (int)kb);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = hip_fail(e, "fp8 quantisation");
    }
    if (!rc) {
        auto kern = fp8_scaled_core<CF>;
        // Once per kernel instantiation (thread-safe static initialisation).
        static const hipError_t attr = hipFuncSetAttribute((const void *)kern,
                                                           hipFuncAttributeMaxDynamicSharedMemorySize, CF::LDS_BYTES);
        if (attr != hipSuccess) rc = hip_fail(attr, "hipFuncSetAttribute");
        if (!rc) {
            timing_begin(s);
            hipLaunchKernelGGL(kern, dim3((unsigned)(tiles_m * tiles_n)), dim3(CF::NT), CF::LDS_BYTES, s, g.M, g.N,
                               (int)Kp, Aq, (int)Mp, ea, Bq, (int)Np, eb, g.alpha, g.beta, g.C, g.ldc, tiles_m,
                               tiles_n);
            timing_end(s);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) rc = hip_fail(e, "fp8_scaled_core launch");
        }
    }
    int rc2 = stream_free(ws, s);
    return rc ? rc : rc2;
}

const char *kNames[] = {"fp8_e4m3_blk128_256x256x128_w2x2", "fp8_e4m3_blk128_128x128x128_w2x2",
                        "fp8_e4m3_blk128_256x256x128_w2x4_ilv"};
constexpr int kNumVariants = 3;

}  // namespace

int fp8_variant_count() { return kNumVariants; }
const char *fp8_variant_name(int v) { return (v >= 0 && v < kNumVariants) ? kNames[v] : nullptr; }

int launch_fp8(const Gemm<float> &g, int variant, hipStream_t s) {
    if (variant < 0) {
        const long t256 = (long)((g.M + 255) / 256) * ((g.N + 255) / 256);
        variant = t256 >= device_cu_count() ? 2 : 1;
    }
    switch (variant) {
        case 0: return run<F8W4>(g, s);
        case 1: return run<F8W4Small>(g, s);
        case 2: return run<F8W8>(g, s);
        default: return fail(UGEMM_ERR_ARG, "fp8: no variant %d", variant);
    }
}

}  // namespace ugemm
