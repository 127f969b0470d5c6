// Split SGEMM on gfx950 MFMA (v_mfma_f32_16x16x32_bf16 / _f16), FP32 accumulate.
//
// Splitting: every FP32 operand x is written as
//   x = h0 + h1 + h2 + r,  h0 = E(x), h1 = E(x - h0), h2 = E(x - h0 - h1)
// in a 16-bit element format E (round-to-nearest-even; the differences are
// exact in FP32).  NP planes are kept and the products
//   NP = 1 (BF16X1): h0*h0
//   NP = 2 (BF16X3): h0*h0 + h0*h1 + h1*h0                  |error| <~ 3*2^-16 |a||b|
//   NP = 3 (BF16X6): + h1*h1 + h0*h2 + h2*h0                |error| <~ 6*2^-24 |a||b|
// are issued as MFMAs into one FP32 accumulator, smallest terms first.
// Products of 16-bit values are exact in FP32, so the remaining error is the
// dropped terms plus FP32 accumulation.
//
// E = bf16 (8 significant bits) has FP32's exponent range and needs no
// scaling.  E = fp16 (11 significant bits: FP16X3, the NP = 2 products, error
// <~ 3*2^-22 |a||b|) has a narrow exponent range, so every row of op(A) and
// every column of op(B) is first scaled by a power of two 2^-e, e =
// exponent(max |x| of that row) - 15, bringing its largest magnitude into
// [2^14, 2^15) (exact: only the exponent changes); the epilogue multiplies
// C(m, n) by 2^(e_m + e_n) with ldexp, which is exact unless the true result
// itself over- or underflows.  Elements more than 2^28 below their row's
// maximum lose low-order bits to fp16 subnormals; their products are at most
// 2^-28 of the row's largest, so the effect on C is below 2^-40 |a||b|.
//
// Two kernels per call:
//  1. split_prep (HBM-bound): reads op(A) / op(B) once in any layout and
//     writes NP bf16 planes in "NT" form -- rows of A (m) and of B (n), k
//     contiguous -- zero-padded to whole tiles, so that the core needs no
//     bounds checks and no layout variants.  Transposed sources go through
//     an LDS tile so both the read and the write are coalesced.
//  2. split_nt_core (MFMA-bound): one workgroup per BM x BN tile of C, 4 waves
//     (2 x 2), each wave a (BM/2) x (BN/2) grid of 16x16 fragments whose
//     FP32 accumulators live in AGPRs; per K step of BK, all NP planes of the
//     A and B tiles are staged global -> registers -> LDS (two stages, the
//     next one in flight during the MFMAs), and every fragment read from LDS
//     serves every product that uses its plane.
// Fragment layout (16x16x32 bf16): lane l holds row l%16, k = 8*(l/16)..+7
// of A and the same k of column l%16 of B; D holds rows 4*(l/16)+i, column
// l%16.  A and B use the same k assignment, so any order of the hardware's
// internal sum pairs the right elements.
#include <string>
#include <type_traits>
#include <utility>

#include "internal.h"
#include "kernel_common.h"

namespace ugemm {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// Element formats of the planes.
struct FmtBf16 {
    typedef __bf16 E;
    typedef bf16x8 V8;
    static constexpr bool SCALED = false;
    static constexpr const char *name = "bf16";
    __device__ static __forceinline__ floatx4 mfma(const V8 &a, const V8 &b, const floatx4 &c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
struct FmtF16 {
    typedef _Float16 E;
    typedef f16x8 V8;
    static constexpr bool SCALED = true;
    static constexpr const char *name = "fp16";
    __device__ static __forceinline__ floatx4 mfma(const V8 &a, const V8 &b, const floatx4 &c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

// Scale exponent of a row from the bit pattern of its largest magnitude:
// x * 2^-row_exp lies in (-2^15, 2^15).  0, inf and NaN give -15 (the
// products are then 0 or non-finite whatever the scale).
__device__ __forceinline__ int row_exp(unsigned amax_bits) {
    int e = 0;
    const float a = __uint_as_float(amax_bits);
    if (a != 0.0f && __builtin_isfinite(a)) frexpf(a, &e);
    return e - 15;
}

// ---------------------------------------------------------------------------
// Splitting.
template <int NP, int W, class E>
__device__ __forceinline__ void split_vec(const float (&x)[W], E (&out)[NP][W]) {
    typedef float fv __attribute__((ext_vector_type(W)));
    typedef E bv __attribute__((ext_vector_type(W)));
    fv v;
#pragma unroll
    for (int i = 0; i < W; ++i) v[i] = x[i];
    bv h = __builtin_convertvector(v, bv);
#pragma unroll
    for (int i = 0; i < W; ++i) out[0][i] = h[i];
    if constexpr (NP > 1) {
        fv hf = __builtin_convertvector(h, fv);
        fv r;
#pragma unroll
        for (int i = 0; i < W; ++i) r[i] = __builtin_isfinite(hf[i]) ? v[i] - hf[i] : 0.0f;
        bv l = __builtin_convertvector(r, bv);
#pragma unroll
        for (int i = 0; i < W; ++i) out[1][i] = l[i];
        if constexpr (NP > 2) {
            fv lf = __builtin_convertvector(l, fv);
            fv r2 = r - lf;
            bv l2 = __builtin_convertvector(r2, bv);
#pragma unroll
            for (int i = 0; i < W; ++i) out[2][i] = l2[i];
        }
    }
}

template <int NP, class FMT>
__device__ __forceinline__ void write_planes(typename FMT::E *dst, long plane_stride, const float (&x)[8]) {
    typename FMT::E o[NP][8];
    split_vec<NP, 8>(x, o);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        typename FMT::V8 w;
#pragma unroll
        for (int j = 0; j < 8; ++j) w[j] = o[p][j];
        *reinterpret_cast<typename FMT::V8 *>(dst + p * plane_stride) = w;
    }
}

// Plane layout written by the preparation kernels ("k-tile blocked"): plane
// p of an operand with rows_pad rows and k_pad (multiple of 32) columns holds
// element (r, k) at
//   p * plane_stride + ((k / 32) * rows_pad + r) * 32 + (chunk ^ sw(r)) * 8 + k % 8,
// chunk = (k % 32) / 8, sw(r) = (r >> 1) & 3.  One K tile of BM rows is thus a
// contiguous BM * 64-byte block whose bytes are already in the (bank-conflict
// free, see swz_chunk) order of the LDS stage: the core copies it verbatim.
__device__ __forceinline__ long blocked_offset(int r, int k, int rows_pad) {
    return ((long)(k >> 5) * rows_pad + r) * 32 + ((((k >> 3) & 3) ^ ((r >> 1) & 3)) << 3);
}

// Source element (r, k) of the operand: KCONT ? X[r*ld + k] : X[k*ld + r];
// zeros outside rows x K.  KCONT version: one thread per 8 k of one row
// (chunk f of the rows_pad x k_pad/8 chunks).
// FMT::SCALED: row r is multiplied by 2^-row_exp(amax[r]) before splitting.
template <int NP, class FMT>
__device__ __forceinline__ void prep_kcont_chunk(const float *__restrict__ X, long ld, int rows, int K, bool vec,
                                                 typename FMT::E *__restrict__ out, long plane_stride, int rows_pad,
                                                 int k_pad, const unsigned *__restrict__ amax, long f) {
    const long chunks_per_row = k_pad / 8;
    const int r = (int)(f / chunks_per_row);
    const int k = (int)(f - (long)r * chunks_per_row) * 8;
    float x[8];
    const float *p = X + (long)r * ld + k;
    if (r < rows && k + 8 <= K && vec) {
        const float4 a = *reinterpret_cast<const float4 *>(p);
        const float4 b = *reinterpret_cast<const float4 *>(p + 4);
        x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w; x[4] = b.x; x[5] = b.y; x[6] = b.z; x[7] = b.w;
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) x[j] = (r < rows && k + j < K) ? p[j] : 0.0f;
    }
    if constexpr (FMT::SCALED) {
        const int e = -row_exp(amax[r]);
#pragma unroll
        for (int j = 0; j < 8; ++j) x[j] = ldexpf(x[j], e);
    }
    write_planes<NP, FMT>(out + blocked_offset(r, k, rows_pad), plane_stride, x);
}

template <int NP, class FMT>
__global__ __launch_bounds__(256) void split_prep_kcont(const float *__restrict__ X, long ld, int rows, int K,
                                                        bool vec, typename FMT::E *__restrict__ out,
                                                        long plane_stride, int rows_pad, int k_pad,
                                                        const unsigned *__restrict__ amax) {
    const long total = (long)rows_pad * (k_pad / 8);
    for (long f = (long)blockIdx.x * blockDim.x + threadIdx.x; f < total; f += (long)gridDim.x * blockDim.x)
        prep_kcont_chunk<NP, FMT>(X, ld, rows, K, vec, out, plane_stride, rows_pad, k_pad, amax, f);
}

// k-contiguous source, 2-D blocks: block (bx, by) covers rows [64 bx, +64)
// and k [128 by, +128) (four 32-wide k tiles).  Chunk f = t + 256 i of a
// thread is row f / 16, k-chunk f % 16: each load instruction of a wave reads
// four 512-byte row segments, and each store writes four 256-byte runs of
// the blocked layout (the rows of one k tile are contiguous there) -- where
// a row-major walk stored sixteen 64-byte pieces per instruction.  All four
// chunks are loaded before any is split.
template <int NP, class FMT>
__device__ __forceinline__ void prep_kcont_tile(const float *__restrict__ X, long ld, int rows, int K, bool vec,
                                                typename FMT::E *__restrict__ out, long plane_stride, int rows_pad,
                                                int k_pad, const unsigned *__restrict__ amax, int r0, int k0) {
    const int t = threadIdx.x;
    // Scaled formats: the row maxima are loaded first, so that their latency
    // overlaps the operand loads instead of following them.
    unsigned am[4] = {0u, 0u, 0u, 0u};
    if constexpr (FMT::SCALED) {
#pragma unroll
        for (int i = 0; i < 4; ++i) am[i] = amax[r0 + ((t + 256 * i) >> 4)];
    }
    float x[4][8];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int f = t + 256 * i, r = r0 + (f >> 4), k = k0 + (f & 15) * 8;
        const float *p = X + (long)r * ld + k;
        if (r < rows && k + 8 <= K && vec) {
            const float4 a = *reinterpret_cast<const float4 *>(p);
            const float4 b = *reinterpret_cast<const float4 *>(p + 4);
            x[i][0] = a.x; x[i][1] = a.y; x[i][2] = a.z; x[i][3] = a.w;
            x[i][4] = b.x; x[i][5] = b.y; x[i][6] = b.z; x[i][7] = b.w;
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) x[i][j] = (r < rows && k + j < K) ? p[j] : 0.0f;
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int f = t + 256 * i, r = r0 + (f >> 4), k = k0 + (f & 15) * 8;
        if (k >= k_pad) continue;
        if constexpr (FMT::SCALED) {
            const int e = -row_exp(am[i]);
#pragma unroll
            for (int j = 0; j < 8; ++j) x[i][j] = ldexpf(x[i][j], e);
        }
        write_planes<NP, FMT>(out + blocked_offset(r, k, rows_pad), plane_stride, x[i]);
    }
}

template <int NP, class FMT>
__global__ __launch_bounds__(256) void split_prep_kcont_tiles(const float *__restrict__ X, long ld, int rows, int K,
                                                              bool vec, typename FMT::E *__restrict__ out,
                                                              long plane_stride, int rows_pad, int k_pad,
                                                              const unsigned *__restrict__ amax) {
    prep_kcont_tile<NP, FMT>(X, ld, rows, K, vec, out, plane_stride, rows_pad, k_pad, amax, blockIdx.x * 64,
                             blockIdx.y * 128);
}

// FP16X3, k-contiguous source with k_pad <= 8192: the row maximum and the
// split in one pass.  Block b reads rows r0 = 2b and r0 + 1 once into
// registers, reduces each row's largest magnitude (amax[r], as the separate
// row-maximum pass computes it: NaN ignored, 0 for padding rows), then
// scales by 2^-row_exp and writes the planes -- the operand is read once
// instead of twice.  Two rows per block, so that their 64-byte segments of
// a K tile -- adjacent in the blocked layout -- leave as one 128-byte line
// (one row per block: 0.172 against 0.138 ms at 8192^3 for the launch below
// that also scans B): thread t takes 8-k chunk t % 4 of row r0 + (t / 4) % 2
// in K tile t / 8 of each 32-tile pass (at most 8 passes, k_pad <= 8192).
// Each row's maximum is reduced over the lanes that hold it (shuffles that
// keep bit 2 of the lane fixed), then over the waves.
template <int NP, class FMT>
__device__ __forceinline__ void prep_kcont_rowscaled_pair(const float *__restrict__ X, long ld, int rows, int K,
                                                          bool vec, typename FMT::E *__restrict__ out,
                                                          long plane_stride, int rows_pad, int k_pad,
                                                          unsigned *__restrict__ amax, int r0) {
    static_assert(FMT::SCALED, "row scales");
    constexpr int PASSES = 8;   // 8 passes x 32 K tiles x 32 = 8192
    const int t = threadIdx.x;
    const int ktl = t >> 3, rr = (t >> 2) & 1, cc = t & 3, r = r0 + rr;
    const int ktiles = k_pad / 32;
    float x[PASSES][8];
    float m = 0.0f;
#pragma unroll
    for (int i = 0; i < PASSES; ++i) {
        const int kt = i * 32 + ktl, k = kt * 32 + cc * 8;
        const float *p = X + (long)r * ld + k;
        if (kt < ktiles && r < rows && k + 8 <= K && vec) {
            const float4 a = *reinterpret_cast<const float4 *>(p);
            const float4 c = *reinterpret_cast<const float4 *>(p + 4);
            x[i][0] = a.x; x[i][1] = a.y; x[i][2] = a.z; x[i][3] = a.w;
            x[i][4] = c.x; x[i][5] = c.y; x[i][6] = c.z; x[i][7] = c.w;
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) x[i][j] = (kt < ktiles && r < rows && k + j < K) ? p[j] : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) m = fmaxf(m, fabsf(x[i][j]));
    }
    m = fmaxf(m, __shfl_xor(m, 1));
    m = fmaxf(m, __shfl_xor(m, 2));
    m = fmaxf(m, __shfl_xor(m, 8));
    m = fmaxf(m, __shfl_xor(m, 16));
    m = fmaxf(m, __shfl_xor(m, 32));
    __shared__ float wmax[4][2];
    if ((t & 63) < 8 && (t & 3) == 0) wmax[t >> 6][rr] = m;
    __syncthreads();
    m = fmaxf(fmaxf(wmax[0][rr], wmax[1][rr]), fmaxf(wmax[2][rr], wmax[3][rr]));
    const unsigned bits = m > 0.0f ? __float_as_uint(m) : 0u;
    if (t < 8 && (t & 3) == 0) amax[r] = bits;
    const int e = -row_exp(bits);
#pragma unroll
    for (int i = 0; i < PASSES; ++i) {
        const int kt = i * 32 + ktl;
        if (kt >= ktiles) break;
#pragma unroll
        for (int j = 0; j < 8; ++j) x[i][j] = ldexpf(x[i][j], e);
        write_planes<NP, FMT>(out + blocked_offset(r, kt * 32 + cc * 8, rows_pad), plane_stride, x[i]);
    }
}

template <int NP, class FMT>
__global__ __launch_bounds__(256) void split_prep_kcont_rowscaled(const float *__restrict__ X, long ld, int rows,
                                                                  int K, bool vec, typename FMT::E *__restrict__ out,
                                                                  long plane_stride, int rows_pad, int k_pad,
                                                                  unsigned *__restrict__ amax) {
    prep_kcont_rowscaled_pair<NP, FMT>(X, ld, rows, K, vec, out, plane_stride, rows_pad, k_pad, amax,
                                       (int)blockIdx.x * 2);
}

// Row-contiguous source (X[k*ld + r]): the 64 x 64 tile (rows r0.., k k0..)
// transposed through LDS (tile[64][65], 256 threads).
template <int NP, class FMT>
__device__ __forceinline__ void prep_rcont_tile(const float *__restrict__ X, long ld, int rows, int K, bool vec,
                                                typename FMT::E *__restrict__ out, long plane_stride, int rows_pad,
                                                int k_pad, const unsigned *__restrict__ amax, int r0, int k0,
                                                float (*tile)[65]) {
    const int t = threadIdx.x;
    // Scaled formats: this thread's two row maxima, loaded before the tile.
    unsigned am[2] = {0u, 0u};
    if constexpr (FMT::SCALED) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int r = r0 + (t + i * 256) / 8;
            am[i] = r < rows_pad ? amax[r] : 0u;
        }
    }
    // Read: 64 k-lines of 64 r each, as 16 float4 per line (coalesced along r).
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int f = t + i * 256;
        const int kk = f / 16, rr = (f % 16) * 4;
        const int k = k0 + kk, r = r0 + rr;
        const float *p = X + (long)k * ld + r;
        float v[4];
        if (k < K && r + 4 <= rows && vec) {
            const float4 a = *reinterpret_cast<const float4 *>(p);
            v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) v[j] = (k < K && r + j < rows) ? p[j] : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) tile[kk][rr + j] = v[j];
    }
    __syncthreads();
    // Write: 64 rows of 8 chunks of 8 k.
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int f = t + i * 256;
        const int rr = f / 8, kk = (f % 8) * 8;
        const int r = r0 + rr, k = k0 + kk;
        if (r >= rows_pad || k >= k_pad) continue;
        float x[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) x[j] = tile[kk + j][rr];
        if constexpr (FMT::SCALED) {
            const int e = -row_exp(am[i]);
#pragma unroll
            for (int j = 0; j < 8; ++j) x[j] = ldexpf(x[j], e);
        }
        write_planes<NP, FMT>(out + blocked_offset(r, k, rows_pad), plane_stride, x);
    }
}

template <int NP, class FMT>
__global__ __launch_bounds__(256) void split_prep_rcont(const float *__restrict__ X, long ld, int rows, int K,
                                                        bool vec, typename FMT::E *__restrict__ out,
                                                        long plane_stride, int rows_pad, int k_pad,
                                                        const unsigned *__restrict__ amax) {
    __shared__ float tile[64][65];
    prep_rcont_tile<NP, FMT>(X, ld, rows, K, vec, out, plane_stride, rows_pad, k_pad, amax, blockIdx.x * 64,
                             blockIdx.y * 64, tile);
}

// block_job[b]: the job of preparation block b (one load, where a search of
// the job table took ~10 dependent loads per block).  A block prepares
// kPrepSpan consecutive k tiles of one 64-row stripe (fewer job lookups and
// block starts per byte): 64-row x 128-k tiles of a k-contiguous job
// (prep_kcont_tile: four 512-byte row segments per load instruction,
// 256-byte runs per store), 64 x 64 transposing tiles of a row-contiguous
// one; nbx blocks along the rows.
template <int NP, class FMT>
__global__ __launch_bounds__(256) void grouped_split_prep(const PrepJob *__restrict__ jobs,
                                                          const int *__restrict__ block_job) {
    __shared__ float tile[64][65];
    const PrepJob j = jobs[block_job[blockIdx.x]];
    const int lb = (int)blockIdx.x - j.block0;
    typename FMT::E *out = static_cast<typename FMT::E *>(j.out);
    if constexpr (FMT::SCALED) {
        if (j.kcont == kPrepRowScaled) {
            // The maxima buffer is the plan's own device memory, written here.
            prep_kcont_rowscaled_pair<NP, FMT>(j.X, j.ld, j.rows, j.K, j.vec != 0, out, j.plane_stride, j.rows_pad,
                                               j.k_pad, const_cast<unsigned *>(j.amax), lb * 2);
            return;
        }
    }
    const int r0 = (lb % j.nbx) * 64, kb = (lb / j.nbx) * kPrepSpan;
    if (j.kcont) {
        for (int s = 0; s < kPrepSpan && (kb + s) * 128 < j.k_pad; ++s)
            prep_kcont_tile<NP, FMT>(j.X, j.ld, j.rows, j.K, j.vec != 0, out, j.plane_stride, j.rows_pad, j.k_pad,
                                     j.amax, r0, (kb + s) * 128);
    } else {
        for (int s = 0; s < kPrepSpan && (kb + s) * 64 < j.k_pad; ++s) {
            if (s > 0) __syncthreads();   // the previous tile's reads of the LDS tile are done
            prep_rcont_tile<NP, FMT>(j.X, j.ld, j.rows, j.K, j.vec != 0, out, j.plane_stride, j.rows_pad, j.k_pad,
                                     j.amax, r0, (kb + s) * 64, tile);
        }
    }
}

// Both operands of one GEMM in one launch (formats without row scales):
// blocks [0, b.block0) prepare a, the rest b, one tile each (64-row x 128-k
// tiles of a k-contiguous operand, 64 x 64 transposing tiles of a
// row-contiguous one; nbx blocks along the rows).  One launch instead of
// two saves a launch and lets b's tiles start while a's last ones drain
// (2048^3 BF16x3: 263 against 256 TF/s per call).
template <int NP, class FMT>
__global__ __launch_bounds__(256) void split_prep_pair(PrepJob a, PrepJob b) {
    static_assert(!FMT::SCALED, "row-scaled formats compute their maxima first");
    __shared__ float tile[64][65];
    const bool second = (int)blockIdx.x >= b.block0;
    const PrepJob &j = second ? b : a;
    const int lb = (int)blockIdx.x - j.block0;
    typename FMT::E *out = static_cast<typename FMT::E *>(j.out);
    const int r0 = (lb % j.nbx) * 64;
    if (j.kcont)
        prep_kcont_tile<NP, FMT>(j.X, j.ld, j.rows, j.K, j.vec != 0, out, j.plane_stride, j.rows_pad, j.k_pad,
                                 nullptr, r0, (lb / j.nbx) * 128);
    else
        prep_rcont_tile<NP, FMT>(j.X, j.ld, j.rows, j.K, j.vec != 0, out, j.plane_stride, j.rows_pad, j.k_pad,
                                 nullptr, r0, (lb / j.nbx) * 64, tile);
}

// Largest magnitude of every plane row (row r of op(A), column r of op(B))
// as IEEE bits, combined with atomicMax on the bit pattern (non-negative
// floats order like their bit patterns).  amax must be zeroed first.
__global__ void fill_zero_u32(unsigned *__restrict__ p, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0u;
}

// k-contiguous source: block b covers row b / kchunks, k in [2048 * (b % kchunks), +2048).
__device__ __forceinline__ void row_amax_kcont_block(const float *__restrict__ X, long ld, int rows, int K, bool vec,
                                                     long kchunks, unsigned *__restrict__ amax, long b) {
    const int r = (int)(b / kchunks);
    const int k = (int)(b % kchunks) * 2048 + threadIdx.x * 8;
    const float *p = X + (long)r * ld + k;
    float m = 0.0f;
    if (k + 8 <= K && vec) {
        const float4 a = *reinterpret_cast<const float4 *>(p);
        const float4 c = *reinterpret_cast<const float4 *>(p + 4);
        m = fmaxf(fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(a.z), fabsf(a.w))),
                  fmaxf(fmaxf(fabsf(c.x), fabsf(c.y)), fmaxf(fabsf(c.z), fabsf(c.w))));
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (k + j < K) m = fmaxf(m, fabsf(p[j]));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    __shared__ float wmax[4];
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
        if (r < rows && m > 0.0f) atomicMax(&amax[r], __float_as_uint(m));
    }
}

__global__ __launch_bounds__(256) void row_amax_kcont(const float *__restrict__ X, long ld, int rows, int K,
                                                      bool vec, long kchunks, unsigned *__restrict__ amax) {
    row_amax_kcont_block(X, ld, rows, K, vec, kchunks, amax, blockIdx.x);
}

// Row-contiguous source (X[k*ld + r]): block b covers columns r in
// [256 * (b % cblocks), +256) and k in [S * (b / cblocks), +S), S =
// kAmaxRcontSpan,
// cblocks = ceil(rows / 256): consecutive blocks read neighbouring 1 KB row
// segments (blocks that shared columns and differed only in k read addresses
// a power-of-two multiple of the row pitch apart, which crowd the same HBM
// channels).
__device__ __forceinline__ void row_amax_rcont_block(const float *__restrict__ X, long ld, int rows, int K, bool vec,
                                                     long kblocks, unsigned *__restrict__ amax, long b) {
    const long cblocks = (rows + 255) / 256;
    constexpr int SPAN = kAmaxRcontSpan;
    const int c0 = (int)(b % cblocks) * 256, k0 = (int)(b / cblocks) * SPAN;
    (void)kblocks;
    const int cg = threadIdx.x & 63, kp = threadIdx.x >> 6;
    const int c = c0 + cg * 4;
    float m[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (vec && c + 4 <= rows && k0 + SPAN <= K) {
        // Whole block: SPAN/4 independent 16-byte loads per thread, issued 16 at
        // a time into registers before any is used (the compiler otherwise
        // waits after every second load, which leaves HBM idle).
        const float *p = X + (long)(k0 + kp) * ld + c;
        for (int kb = 0; kb < SPAN / 4; kb += 16) {
            float4 v[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) v[i] = *reinterpret_cast<const float4 *>(p + (long)(kb + i) * 4 * ld);
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                m[0] = fmaxf(m[0], fabsf(v[i].x)); m[1] = fmaxf(m[1], fabsf(v[i].y));
                m[2] = fmaxf(m[2], fabsf(v[i].z)); m[3] = fmaxf(m[3], fabsf(v[i].w));
            }
        }
    } else {
        for (int kk = kp; kk < SPAN && k0 + kk < K; kk += 4) {
            const float *p = X + (long)(k0 + kk) * ld + c;
            if (vec && c + 4 <= rows) {
                const float4 a = *reinterpret_cast<const float4 *>(p);
                m[0] = fmaxf(m[0], fabsf(a.x)); m[1] = fmaxf(m[1], fabsf(a.y));
                m[2] = fmaxf(m[2], fabsf(a.z)); m[3] = fmaxf(m[3], fabsf(a.w));
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (c + j < rows) m[j] = fmaxf(m[j], fabsf(p[j]));
            }
        }
    }
    __shared__ float sm[4][256];
#pragma unroll
    for (int j = 0; j < 4; ++j) sm[kp][cg * 4 + j] = m[j];
    __syncthreads();
    const int t = threadIdx.x, col = c0 + t;
    const float v = fmaxf(fmaxf(sm[0][t], sm[1][t]), fmaxf(sm[2][t], sm[3][t]));
    if (col < rows && v > 0.0f) atomicMax(&amax[col], __float_as_uint(v));
}

__global__ __launch_bounds__(256) void row_amax_rcont(const float *__restrict__ X, long ld, int rows, int K,
                                                      bool vec, long kblocks, unsigned *__restrict__ amax) {
    row_amax_rcont_block(X, ld, rows, K, vec, kblocks, amax, blockIdx.x);
}

// Row maxima of the operands of a grouped FP16X3 plan: job j (PrepJob with
// out = its maxima, nbx = K chunks per row (k-contiguous: 2048 wide) or per
// 256-column block (row-contiguous: kAmaxRcontSpan deep)), blocks as in the kernels above.
// FP16X3 with k-contiguous op(A) (K <= 8192) and row-contiguous op(B): A's
// fused row maximum + split (one block per two padded rows of A) and B's column
// maxima (row_amax_rcont blocks, combined with atomicMax into zeroed maxima)
// in one launch, so that the two passes over HBM overlap.  Blocks below
// b_blocks scan B; the others split rows 2 (blockIdx.x - b_blocks) and the
// next of A.
template <int NP, class FMT>
__global__ __launch_bounds__(256) void split_prep_rowscaled_a_amax_b(
    const float *__restrict__ A, long lda, int rows_a, int K, bool vec_a, typename FMT::E *__restrict__ out_a,
    long plane_a, int rows_pad_a, int k_pad, unsigned *__restrict__ amax_a, const float *__restrict__ B, long ldb,
    int rows_b, bool vec_b, long kblocks_b, unsigned *__restrict__ amax_b, int b_blocks) {
    if ((int)blockIdx.x < b_blocks)
        row_amax_rcont_block(B, ldb, rows_b, K, vec_b, kblocks_b, amax_b, blockIdx.x);
    else
        prep_kcont_rowscaled_pair<NP, FMT>(A, lda, rows_a, K, vec_a, out_a, plane_a, rows_pad_a, k_pad, amax_a,
                                           ((int)blockIdx.x - b_blocks) * 2);
}

__global__ __launch_bounds__(256) void grouped_row_amax(const PrepJob *__restrict__ jobs,
                                                        const int *__restrict__ block_job) {
    const PrepJob j = jobs[block_job[blockIdx.x]];
    const long lb = (long)blockIdx.x - j.block0;
    unsigned *amax = static_cast<unsigned *>(j.out);
    if (j.kcont) row_amax_kcont_block(j.X, j.ld, j.rows, j.K, j.vec != 0, j.nbx, amax, lb);
    else row_amax_rcont_block(j.X, j.ld, j.rows, j.K, j.vec != 0, j.nbx, amax, lb);
}

// Implicit im2col operand of a convolution (ConvDesc): element (n, k) of the
// N x K operand (opB transposed) is in[c][ho*stride - pad + i][wo*stride - pad + j]
// for n = ho*wcol + wo and k = (c, i, j) = c*kk*kk + i*kk + j, zero outside the
// image and beyond N x K; ktab[k] = {c*H*W + i*W + j, (i << 16) | j}.  The
// gather follows sgemm_gl1.h:166-190 (see oracle_im2col).
__device__ __forceinline__ float im2col_at(const ConvDesc &cd, int n, int k, int N, int K) {
    if (n >= N || k >= K) return 0.0f;
    const int ho = n / cd.wcol, wo = n - ho * cd.wcol;
    const int h0 = ho * cd.stride - cd.pad, w0 = wo * cd.stride - cd.pad;
    const int2 e = cd.ktab[k];
    const int hi = h0 + (e.y >> 16), wi = w0 + (e.y & 0xffff);
    return ((unsigned)hi < (unsigned)cd.H && (unsigned)wi < (unsigned)cd.W) ? cd.in[(long)h0 * cd.W + w0 + e.x] : 0.0f;
}

// Planes of the im2col operand, gathered straight from the image (no im2col
// matrix in HBM): one thread per 8 k of one pixel n, as split_prep_kcont.
template <int NP, class FMT>
__global__ __launch_bounds__(256) void split_prep_im2col(ConvDesc cd, int N, int K, typename FMT::E *__restrict__ out,
                                                         long plane_stride, int rows_pad, int k_pad,
                                                         const unsigned *__restrict__ amax) {
    const long chunks_per_row = k_pad / 8;
    const long total = (long)rows_pad * chunks_per_row;
    for (long f = (long)blockIdx.x * blockDim.x + threadIdx.x; f < total; f += (long)gridDim.x * blockDim.x) {
        const int r = (int)(f / chunks_per_row);
        const int k = (int)(f - (long)r * chunks_per_row) * 8;
        float x[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) x[j] = im2col_at(cd, r, k + j, N, K);
        if constexpr (FMT::SCALED) {
            const int e = -row_exp(amax[r]);
#pragma unroll
            for (int j = 0; j < 8; ++j) x[j] = ldexpf(x[j], e);
        }
        write_planes<NP, FMT>(out + blocked_offset(r, k, rows_pad), plane_stride, x);
    }
}

// Largest magnitude of every pixel column of the im2col operand (scaled
// formats): thread n of block (bx, by) scans k in [by*kspan, +kspan).
__global__ __launch_bounds__(256) void row_amax_im2col(ConvDesc cd, int N, int K, int kspan,
                                                       unsigned *__restrict__ amax) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    const int k0 = blockIdx.y * kspan, k1 = min(K, k0 + kspan);
    float m = 0.0f;
    for (int k = k0; k < k1; ++k) m = fmaxf(m, fabsf(im2col_at(cd, n, k, N, K)));
    if (m > 0.0f) atomicMax(&amax[n], __float_as_uint(m));
}

// ---------------------------------------------------------------------------
// MFMA core over prepared planes.

// 16-byte chunk swizzle inside a plane row of CPR chunks: chunk c of row r is
// stored at c ^ sw(r), sw(r) = (r >> 1) & 3 (CPR 4) or r & 7 (CPR 8).
// Measured with SQ_LDS_BANK_CONFLICT on gfx950: fragment reads (16 lanes =
// 16 consecutive rows, one chunk each) are conflict-free with these.
template <int CPR>
__device__ __forceinline__ int swz_chunk(int row, int c) {
    if constexpr (CPR == 8) return c ^ (row & 7);
    else return c ^ ((row >> 1) & 3);
}

template <int NP>
struct Products;
template <>
struct Products<1> {
    static constexpr int N = 1;
    static constexpr int a[1] = {0}, b[1] = {0};
};
template <>
struct Products<2> {
    static constexpr int N = 3;
    static constexpr int a[3] = {1, 0, 0}, b[3] = {0, 1, 0};
};
template <>
struct Products<3> {
    static constexpr int N = 6;
    static constexpr int a[6] = {2, 0, 1, 1, 0, 0}, b[6] = {0, 2, 1, 0, 1, 0};
};

// First product (in Products order) that uses plane p of A / of B: the
// phased core reads a plane's fragments from LDS right before that product.
template <int NP>
constexpr bool first_use_a(int q) {
    for (int r = 0; r < q; ++r)
        if (Products<NP>::a[r] == Products<NP>::a[q]) return false;
    return true;
}
template <int NP>
constexpr bool first_use_b(int q) {
    for (int r = 0; r < q; ++r)
        if (Products<NP>::b[r] == Products<NP>::b[q]) return false;
    return true;
}

// WGM x WGN waves per workgroup.  PHASED: fragments of each plane are read
// just before the first product that needs them, behind a scheduling
// barrier, so that at most three planes' fragments are live (2 waves per
// SIMD fit in 256 registers); the other wave on the SIMD covers the LDS
// latency.  Otherwise all fragments of a K step are read first.
//
// ILV (phased variants with two LDS stages): stage kt+1 waits in registers
// during step kt.  Its LDS stores, and the global loads of stage kt+2 into
// the same registers, are interleaved with the MFMAs of the step's last
// product instead of running between the MFMAs and the barrier: the loads
// get a full step of latency and no wave stalls on a burst of stores in
// front of the barrier.  It needs the registers of 2 waves per SIMD.
// STREAM (required by ILV): each A fragment is read two MFMA groups (of FN
// MFMAs) ahead of its use, and the next product's B fragments during the
// current product's last groups, in an order that scheduling barriers fix.
// (Left to itself the compiler issued a product's reads in one burst and
// then waited on them.)  Every variant issues the same MFMAs per accumulator
// in the same order: results are bitwise identical across them.
// EARLY (with STREAM, for NP >= 2, whose last product reuses planes held in
// registers): the block barrier moves from the end of a K tile to the start
// of its last product.  The stage copy is dealt over the products before the
// barrier, so after it the last product's MFMAs cover the reads of the next
// K tile's first fragments (from the other buffer) instead of the tile
// starting on a burst of reads.
template <int BM_, int BN_, int BK_, int NP_, class FMT_ = FmtBf16, int WGM_ = 2, int WGN_ = 2, bool PHASED_ = false,
          bool ILV_ = false, bool STREAM_ = false, bool EARLY_ = false>
struct CoreCfg {
    using FMT = FMT_;
    static constexpr int BM = BM_, BN = BN_, BK = BK_, NP = NP_;
    static constexpr bool PHASED = PHASED_, ILV = ILV_, STREAM = STREAM_, EARLY = EARLY_;
    static constexpr int WGM = WGM_, WGN = WGN_, NT = 64 * WGM_ * WGN_;
    static constexpr int WTM = BM / WGM, WTN = BN / WGN;
    static constexpr int FM = WTM / 16, FN = WTN / 16;
    static constexpr int KSTEPS = BK / 32, CPR = BK / 8;
    static constexpr int A_PLANE = BM * BK, B_PLANE = BN * BK;           // bf16 elements
    static constexpr int STAGE = NP * (A_PLANE + B_PLANE);
    static constexpr int STAGES = (2 * STAGE * 2 <= 160 * 1024) ? 2 : 1;
    static constexpr int LDS_BYTES = STAGES * STAGE * 2;
    static constexpr int CHUNKS = NP * (BM + BN) * CPR;                 // 16-byte chunks per stage
    // 8-wave variants can also split a k-contiguous FP32 operand while
    // staging it (no preparation pass for that operand); 16 waves have no
    // registers to spare for it.
    static constexpr bool FUSABLE = 64 * WGM_ * WGN_ == 512;
    static constexpr int PER_T = CHUNKS / NT;
    static_assert(CHUNKS % NT == 0, "stage must split evenly over the threads");
    static_assert(WTM % 16 == 0 && WTN % 16 == 0 && BK % 32 == 0, "tile shape");
    static_assert(!ILV || (PHASED && STAGES == 2 && KSTEPS == 1), "interleaved staging needs phased reads, two stages");
    static_assert(ILV == STREAM && (!STREAM || FM >= 4), "the interleaved staging runs with streamed reads");
    static_assert(!EARLY || (STREAM && NP >= 2 && FN + 2 <= FM), "early reads: streamed, last product resident");
};

enum { FUSE_NONE = 0, FUSE_KCONT = 1, FUSE_RCONT = 2 };

// amax_a / amax_b: per-row largest magnitudes of the scaled formats
// (FMT::SCALED), unused otherwise.
// FA / FB (FUSE_*): where the core takes that operand's stage from.
//   FUSE_NONE:  prepared planes, copied verbatim.
//   FUSE_KCONT: split while staging from the k-contiguous FP32 source
//     X[r*ld + k].  Thread t handles the chunks f = t + NT*i of the stage
//     (row r = f / 4, stored chunk f % 4, which holds k-chunk (f % 4) ^ sw(r)).
//   FUSE_RCONT: split while staging from the row-contiguous FP32 source
//     X[k*ld + r], transposed in registers: thread t loads 8 k-lines of the
//     RPT adjacent rows rp*RPT .. rp*RPT + RPT - 1 (RPT = rows*4/NT, 1 or 2;
//     rp = t % (rows/RPT), k-chunk t / (rows/RPT)), coalesced along r, and
//     splits one 8-k chunk per row.
// Either way each split chunk lands where the plane copy would have put it
// (zeros beyond M, N, K).
// GR (grouped plans): block b computes tile gtiles[b] of problem
// gprobs[gtiles[b].problem], whose planes, sizes, alpha, beta and C (and,
// with FB, B's FP32 source) replace the corresponding arguments (no split K,
// no epilogue, A always from planes).
template <class CF, int FA, int FB, bool GR = false>
__global__ __launch_bounds__(CF::NT) void split_nt_core(int M_, int N_, int Kp_, const typename CF::FMT::E *__restrict__ Ap_,
                                                       long a_plane_, int Mp_, const typename CF::FMT::E *__restrict__ Bp_,
                                                       long b_plane_, int Np_, float alpha_, float beta_,
                                                       float *__restrict__ C_, long ldc_, int tiles_m, int tiles_n,
                                                       const unsigned *__restrict__ amax_a_,
                                                       const unsigned *__restrict__ amax_b_, Epilogue epi,
                                                       const float *__restrict__ Af, long lda,
                                                       const float *__restrict__ Bf_, long ldb_, int K_,
                                                       int vec_ab_, int kt_per, float *__restrict__ ws,
                                                       const GroupedSplitProblem *__restrict__ gprobs,
                                                       const GroupedSplitTile *__restrict__ gtiles) {
    using FMT = typename CF::FMT;
    using E = typename FMT::E;
    using V8 = typename FMT::V8;
    static_assert(!GR || FA == 0, "grouped plans stream A's prepared planes");
    const GroupedSplitTile gt = GR ? gtiles[blockIdx.x] : GroupedSplitTile{0, 0, 0, 0};
    const GroupedSplitProblem *gp = GR ? gprobs + gt.problem : nullptr;
    const int M = GR ? gp->M : M_, N = GR ? gp->N : N_, Kp = GR ? gp->Kp : Kp_;
    const int Mp = GR ? gp->Mp : Mp_, Np = GR ? gp->Np : Np_;
    const E *__restrict__ Ap = GR ? static_cast<const E *>(gp->Ap) : Ap_;
    const E *__restrict__ Bp = GR ? static_cast<const E *>(gp->Bp) : Bp_;
    const long a_plane = GR ? gp->a_plane : a_plane_, b_plane = GR ? gp->b_plane : b_plane_;
    const float alpha = GR ? gp->alpha : alpha_, beta = GR ? gp->beta : beta_;
    float *__restrict__ C = GR ? gp->C : C_;
    const long ldc = GR ? gp->ldc : ldc_;
    const unsigned *__restrict__ amax_a = GR ? gp->amax_a : amax_a_;
    const unsigned *__restrict__ amax_b = GR ? gp->amax_b : amax_b_;
    // B's FP32 source when the core splits it (grouped: per problem).
    const float *__restrict__ Bf = GR ? gp->Bf : Bf_;
    const long ldb = GR ? gp->ldb : ldb_;
    const int K = GR ? gp->K : K_;
    const int vec_ab = GR ? (gp->vec_b ? 2 : 0) : vec_ab_;
    constexpr int BM = CF::BM, BN = CF::BN, BK = CF::BK, NP = CF::NP, FM = CF::FM, FN = CF::FN;
    constexpr int CPR = CF::CPR, STAGE = CF::STAGE, STAGES = CF::STAGES, PER_T = CF::PER_T;
    extern __shared__ __attribute__((aligned(16))) char smem_raw[];
    E *lds = reinterpret_cast<E *>(smem_raw);

    int tm = 0, tn = 0;
    if (!GR) tile_coords(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const int m0 = GR ? gt.m0 : tm * BM, n0 = GR ? gt.n0 : tn * BN;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w / CF::WGN, wn = w % CF::WGN;

    // Stage copy.  In the k-tile blocked plane layout the A tile of K tile kt
    // and plane p is the contiguous block Ap + p*a_plane + (kt*Mp + m0)*BK of
    // BM*BK elements, already in LDS order; likewise B.  Chunk i of thread t
    // is f = t + NT*i (A chunks first), copied to the same position in LDS.
    static_assert(BK == 32, "blocked plane layout uses 32-wide K tiles");
    constexpr int A_CH = NP * BM * CPR / CF::NT;         // 16-byte chunks per thread from A
    static_assert((NP * BM * CPR) % CF::NT == 0 && (BM * CPR) % CF::NT == 0 && (BN * CPR) % CF::NT == 0,
                  "stage split");
    const E *a_src = Ap + (long)m0 * BK + t * 8;
    const E *b_src = Bp + (long)n0 * BK + t * 8;
    constexpr int B_CH = NP * BN * CPR / CF::NT;          // 16-byte chunks per thread from B planes
    constexpr int FA_CH = BM * CPR / CF::NT, FB_CH = BN * CPR / CF::NT;   // fused chunk rows per thread
    static_assert(CPR == 4, "fused staging assumes 32-wide K tiles");
    static_assert((FA != FUSE_RCONT || FA_CH == 1 || FA_CH == 2) && (FB != FUSE_RCONT || FB_CH == 1 || FB_CH == 2),
                  "row-contiguous fusion: one or two rows per thread");
    V8 pa[FA ? 1 : A_CH], pb[FB ? 1 : B_CH];
    float fa[FA ? FA_CH : 1][8], fb[FB ? FB_CH : 1][8];
    // Row-contiguous source: rows [r, r+RW) of k-lines k .. k+7 into v[i][j]
    // (row r+i, k+j); zeros outside rows x K.  RW is 1 or 2 (stage rows per
    // thread); vec: 8-byte loads allowed (RW == 2).  full: the whole stage
    // lies inside rows x K and vector loads are allowed (uniform over the
    // block): unguarded loads, no per-element branches.
    auto load_rc = [&](const float *X, long ld, int rows, int r, int k, bool vec, auto &v, bool full) {
        constexpr int RW = std::extent<std::remove_reference_t<decltype(v)>>::value;
        if (full) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                if constexpr (RW == 2) {
                    const float2 x = *reinterpret_cast<const float2 *>(X + (long)(k + j) * ld + r);
                    v[0][j] = x.x;
                    v[1][j] = x.y;
                } else {
                    v[0][j] = X[(long)(k + j) * ld + r];
                }
            }
            return;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const float *p = X + (long)(k + j) * ld + r;
            if constexpr (RW == 2) {
                if (k + j < K && r + 2 <= rows && vec) {
                    const float2 x = *reinterpret_cast<const float2 *>(p);
                    v[0][j] = x.x;
                    v[1][j] = x.y;
                } else {
                    v[0][j] = (k + j < K && r < rows) ? p[0] : 0.0f;
                    v[1][j] = (k + j < K && r + 1 < rows) ? p[1] : 0.0f;
                }
            } else {
                v[0][j] = (k + j < K && r < rows) ? p[0] : 0.0f;
            }
        }
    };
    // 8 floats of row `row`, k in [k, k+8), zeros outside rows x K.
    auto load8 = [&](const float *X, long ld, int rows, int row, int k, bool vec, float (&v)[8], bool full) {
        const float *p = X + (long)row * ld + k;
        if (full || (row < rows && k + 8 <= K && vec)) {
            const float4 x0 = *reinterpret_cast<const float4 *>(p);
            const float4 x1 = *reinterpret_cast<const float4 *>(p + 4);
            v[0] = x0.x; v[1] = x0.y; v[2] = x0.z; v[3] = x0.w; v[4] = x1.x; v[5] = x1.y; v[6] = x1.z; v[7] = x1.w;
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = (row < rows && k + j < K) ? p[j] : 0.0f;
        }
    };
    // Plane copy, chunk c (A chunks first): global -> registers, and
    // registers -> LDS buffer buf.
    auto load_chunk = [&](int kt, int c) {
        if (c < A_CH) {
            const int e = c * CF::NT * 8, p = e / (BM * BK), off = e % (BM * BK);
            pa[c] = *reinterpret_cast<const V8 *>(a_src + p * a_plane + (long)kt * Mp * BK + off);
        } else {
            const int e = (c - A_CH) * CF::NT * 8, p = e / (BN * BK), off = e % (BN * BK);
            pb[c - A_CH] = *reinterpret_cast<const V8 *>(b_src + p * b_plane + (long)kt * Np * BK + off);
        }
    };
    auto store_chunk = [&](int buf, int c) {
        E *a_reg = lds + buf * STAGE, *b_reg = a_reg + NP * CF::A_PLANE;
        if (c < A_CH)
            *reinterpret_cast<V8 *>(a_reg + t * 8 + c * CF::NT * 8) = pa[c];
        else
            *reinterpret_cast<V8 *>(b_reg + t * 8 + (c - A_CH) * CF::NT * 8) = pb[c - A_CH];
    };
    auto load_stage = [&](int kt) {
        const int k0 = kt * BK;
        const bool kin = k0 + BK <= K;
        if constexpr (FA == FUSE_RCONT) {
            load_rc(Af, lda, M, m0 + (t % (BM / FA_CH)) * FA_CH, k0 + (t / (BM / FA_CH)) * 8, vec_ab & 1, fa,
                    kin && (FA_CH == 1 || (vec_ab & 1)) && m0 + BM <= M);
        } else if constexpr (FA == FUSE_KCONT) {
            const bool full = kin && (vec_ab & 1) && m0 + BM <= M;
#pragma unroll
            for (int i = 0; i < FA_CH; ++i) {
                const int f = t + CF::NT * i, r = f >> 2, c = (f & 3) ^ ((r >> 1) & 3);
                load8(Af, lda, M, m0 + r, k0 + c * 8, vec_ab & 1, fa[i], full);
            }
        } else {
#pragma unroll
            for (int i = 0; i < A_CH; ++i) load_chunk(kt, i);
        }
        if constexpr (FB == FUSE_RCONT) {
            load_rc(Bf, ldb, N, n0 + (t % (BN / FB_CH)) * FB_CH, k0 + (t / (BN / FB_CH)) * 8, (vec_ab & 2) != 0, fb,
                    kin && (FB_CH == 1 || (vec_ab & 2)) && n0 + BN <= N);
        } else if constexpr (FB == FUSE_KCONT) {
            const bool full = kin && (vec_ab & 2) && n0 + BN <= N;
#pragma unroll
            for (int i = 0; i < FB_CH; ++i) {
                const int f = t + CF::NT * i, r = f >> 2, c = (f & 3) ^ ((r >> 1) & 3);
                load8(Bf, ldb, N, n0 + r, k0 + c * 8, (vec_ab & 2) != 0, fb[i], full);
            }
        } else {
#pragma unroll
            for (int i = 0; i < B_CH; ++i) load_chunk(kt, A_CH + i);
        }
    };
    // Split one fused chunk (scaled formats: times 2^-e of its row) into its
    // NP plane chunks at plane offset f * 8 of the operand's LDS region.
    auto store_fused = [&](E *region, int plane, const unsigned *amax, int row, int f, float (&v)[8]) {
        if constexpr (FMT::SCALED) {
            const int e = -row_exp(amax[row]);
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = ldexpf(v[j], e);
        }
        E o[NP][8];
        split_vec<NP, 8>(v, o);
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            V8 w;
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = o[p][j];
            *reinterpret_cast<V8 *>(region + p * plane + f * 8) = w;
        }
    };
    auto store_stage = [&](int buf) {
        E *a_reg = lds + buf * STAGE, *b_reg = a_reg + NP * CF::A_PLANE;
        if constexpr (FA == FUSE_RCONT) {
#pragma unroll
            for (int i = 0; i < FA_CH; ++i) {
                const int r = (t % (BM / FA_CH)) * FA_CH + i, f = r * 4 + ((t / (BM / FA_CH)) ^ ((r >> 1) & 3));
                store_fused(a_reg, CF::A_PLANE, amax_a, m0 + r, f, fa[i]);
            }
        } else if constexpr (FA == FUSE_KCONT) {
#pragma unroll
            for (int i = 0; i < FA_CH; ++i) {
                const int f = t + CF::NT * i;
                store_fused(a_reg, CF::A_PLANE, amax_a, m0 + (f >> 2), f, fa[i]);
            }
        } else {
#pragma unroll
            for (int i = 0; i < A_CH; ++i) store_chunk(buf, i);
        }
        if constexpr (FB == FUSE_RCONT) {
#pragma unroll
            for (int i = 0; i < FB_CH; ++i) {
                const int r = (t % (BN / FB_CH)) * FB_CH + i, f = r * 4 + ((t / (BN / FB_CH)) ^ ((r >> 1) & 3));
                store_fused(b_reg, CF::B_PLANE, amax_b, n0 + r, f, fb[i]);
            }
        } else if constexpr (FB == FUSE_KCONT) {
#pragma unroll
            for (int i = 0; i < FB_CH; ++i) {
                const int f = t + CF::NT * i;
                store_fused(b_reg, CF::B_PLANE, amax_b, n0 + (f >> 2), f, fb[i]);
            }
        } else {
#pragma unroll
            for (int i = 0; i < B_CH; ++i) store_chunk(buf, A_CH + i);
        }
    };

    floatx4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    // K tiles [kt0, ktiles) of this block: all of them, or (split-K,
    // gridDim.y > 1) the kt_per tiles of split blockIdx.y, whose alpha = 1,
    // beta = 0 product goes to the partial slice ws + z*M*N (row length N).
    const bool part = gridDim.y > 1;
    const int kt0 = part ? (int)blockIdx.y * kt_per : 0;
    const int ktiles = part ? min(Kp / BK, kt0 + kt_per) : Kp / BK;
    float *const Co = part ? ws + (long)blockIdx.y * M * N : C;
    const long ldo = part ? (long)N : ldc;
    const float al = part ? 1.0f : alpha, be = part ? 0.0f : beta;
    constexpr bool ILV = CF::ILV && !FA && !FB;
    // Fused operands on two stages (ESTORE): stage kt+1, loaded during step
    // kt-1, is split into the idle buffer at the top of step kt and stage
    // kt+2 is loaded right after, so the FP32 loads get a whole step of
    // latency and the split of one block overlaps the MFMAs of others
    // (split at the end of the step, every wave waited on its loads and then
    // split in front of the barrier).  Branch-free like ILV: the last steps
    // store into the idle buffer and re-load the last stage.
    constexpr bool ESTORE = !ILV && (FA || FB) && STAGES == 2;
    load_stage(kt0);
    store_stage(0);
    if constexpr (ILV || ESTORE) load_stage(min(kt0 + 1, ktiles - 1));
    __syncthreads();

    const int frow = lane & 15, fc = lane >> 4;
    if constexpr (CF::EARLY && ILV) {
        // Group (q, i): the FN MFMAs of product q with A fragment i.  In the
        // products before the last, a group reads the A fragment of group
        // (q, i + LA) and (last FN groups) a B fragment of product q + 1, and
        // takes its share of the stage copy (K tile kt + 1 into the idle
        // buffer, K tile kt + 2 into registers).  The barrier then opens the
        // last product, whose planes are held in registers; its groups read
        // K tile kt + 1's first B and A fragments from the other buffer.
        // Every wave has finished reading a buffer (the products before the
        // barrier) before any wave writes it again (after the next barrier).
        constexpr int NQ = Products<NP>::N, LA = 2, NCH = A_CH + B_CH, NGC = (NQ - 1) * FM;
        constexpr int a0 = Products<NP>::a[0], b0 = Products<NP>::b[0];
        static_assert(!first_use_a<NP>(NQ - 1) && !first_use_b<NP>(NQ - 1), "last product from registers");
        V8 fa[NP][FM], fb[NP][FN], na[LA], nb[FN];
        auto frag = [&](const E *base, int row) {
            return *reinterpret_cast<const V8 *>(base + row * BK + swz_chunk<CPR>(row, fc) * 8);
        };
        auto a_row = [&](int i) { return wm * CF::WTM + i * 16 + frow; };
        auto b_row = [&](int j) { return wn * CF::WTN + j * 16 + frow; };
#pragma unroll
        for (int j = 0; j < FN; ++j) fb[b0][j] = frag(lds + NP * CF::A_PLANE + b0 * CF::B_PLANE, b_row(j));
#pragma unroll
        for (int i = 0; i < LA; ++i) fa[a0][i] = frag(lds + a0 * CF::A_PLANE, a_row(i));
        for (int kt = kt0; kt < ktiles; ++kt) {
            const int buf = (kt - kt0) & 1;
            const E *as = lds + buf * STAGE, *bs = as + NP * CF::A_PLANE;
            const E *an = lds + (buf ^ 1) * STAGE, *bn = an + NP * CF::A_PLANE;
            static_for<NQ>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                constexpr int qa = Products<NP>::a[q], qb = Products<NP>::b[q];
                constexpr bool last = q == NQ - 1;
                constexpr int qn = last ? q : q + 1;
                static_for<FM>([&](auto ic) {
                    constexpr int i = decltype(ic)::value;
                    __builtin_amdgcn_sched_barrier(0);
                    if constexpr (!last) {
                        if constexpr (i + LA < FM) {
                            if constexpr (first_use_a<NP>(q)) fa[qa][i + LA] = frag(as + qa * CF::A_PLANE, a_row(i + LA));
                        } else if constexpr (first_use_a<NP>(qn)) {
                            constexpr int pn = Products<NP>::a[qn];
                            fa[pn][i + LA - FM] = frag(as + pn * CF::A_PLANE, a_row(i + LA - FM));
                        }
                        if constexpr (first_use_b<NP>(qn) && i >= FM - FN) {
                            constexpr int pn = Products<NP>::b[qn];
                            fb[pn][i - (FM - FN)] = frag(bs + pn * CF::B_PLANE, b_row(i - (FM - FN)));
                        }
                        static_for<NCH>([&](auto cc) {
                            constexpr int c = decltype(cc)::value;
                            if constexpr (c * NGC / NCH == q * FM + i) {
                                store_chunk(buf ^ 1, c);
                                load_chunk(min(kt + 2, ktiles - 1), c);
                            }
                        });
                    } else {
                        if constexpr (i == 0) {
                            __syncthreads();
                            __builtin_amdgcn_sched_barrier(0);
                        }
                        if constexpr (i < FN) nb[i] = frag(bn + b0 * CF::B_PLANE, b_row(i));
                        else if constexpr (i - FN < LA) na[i - FN] = frag(an + a0 * CF::A_PLANE, a_row(i - FN));
                    }
#pragma unroll
                    for (int j = 0; j < FN; ++j) acc[i][j] = FMT::mfma(fa[qa][i], fb[qb][j], acc[i][j]);
                });
            });
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < FN; ++j) fb[b0][j] = nb[j];
#pragma unroll
            for (int i = 0; i < LA; ++i) fa[a0][i] = na[i];
        }
    } else
    for (int kt = kt0; kt < ktiles; ++kt) {
        const int buf = (STAGES == 2) ? ((kt - kt0) & 1) : 0;
        const bool more = kt + 1 < ktiles;
        if constexpr (ESTORE) {
            store_stage(buf ^ 1);
            load_stage(min(kt + 2, ktiles - 1));
        } else if (!ILV && more) {
            load_stage(kt + 1);
        }
        const E *as = lds + buf * STAGE;
        const E *bs = as + NP * CF::A_PLANE;
#pragma unroll
        for (int s = 0; s < CF::KSTEPS; ++s) {
            V8 fa[NP][FM], fb[NP][FN];
            if constexpr (CF::STREAM && ILV) {
                // Group (q, i): the FN MFMAs of product q with A fragment i.
                // Before its MFMAs a group reads the A fragment that group
                // (q, i + LA) needs (product q + 1's first ones near the end
                // of q), and each of the last FN groups of product q one B
                // fragment of product q + 1.  Planes used again later stay in
                // registers.  A scheduling barrier around every group keeps
                // the reads in this order; the last product deals the stage
                // copy out one chunk per group.
                constexpr int NQ = Products<NP>::N, LA = 2, NCH = A_CH + B_CH;
                auto read_a = [&](int p, int i) {
                    const int row = wm * CF::WTM + i * 16 + frow;
                    fa[p][i] = *reinterpret_cast<const V8 *>(as + p * CF::A_PLANE + row * BK +
                                                             swz_chunk<CPR>(row, s * 4 + fc) * 8);
                };
                auto read_b = [&](int p, int j) {
                    const int row = wn * CF::WTN + j * 16 + frow;
                    fb[p][j] = *reinterpret_cast<const V8 *>(bs + p * CF::B_PLANE + row * BK +
                                                             swz_chunk<CPR>(row, s * 4 + fc) * 8);
                };
#pragma unroll
                for (int j = 0; j < FN; ++j) read_b(Products<NP>::b[0], j);
#pragma unroll
                for (int i = 0; i < LA; ++i) read_a(Products<NP>::a[0], i);
                static_for<NQ>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    constexpr int qa = Products<NP>::a[q], qb = Products<NP>::b[q];
                    constexpr bool last = q == NQ - 1;
                    constexpr int qn = last ? q : q + 1;
                    static_for<FM>([&](auto ic) {
                        constexpr int i = decltype(ic)::value;
                        __builtin_amdgcn_sched_barrier(0);
                        if constexpr (i + LA < FM) {
                            if constexpr (first_use_a<NP>(q)) read_a(qa, i + LA);
                        } else if constexpr (!last && first_use_a<NP>(qn)) {
                            read_a(Products<NP>::a[qn], i + LA - FM);
                        }
                        if constexpr (!last && first_use_b<NP>(qn) && i >= FM - FN)
                            read_b(Products<NP>::b[qn], i - (FM - FN));
                        if constexpr (last) {
                            // Chunks c with c * FM / NCH == i: store into the
                            // idle buffer, then load stage kt + 2 (the last
                            // steps re-load the last stage: addresses in range).
                            static_for<NCH>([&](auto cc) {
                                constexpr int c = decltype(cc)::value;
                                if constexpr (c * FM / NCH == i) {
                                    store_chunk(buf ^ 1, c);
                                    load_chunk(min(kt + 2, ktiles - 1), c);
                                }
                            });
                        }
#pragma unroll
                        for (int j = 0; j < FN; ++j) acc[i][j] = FMT::mfma(fa[qa][i], fb[qb][j], acc[i][j]);
                    });
                });
                __builtin_amdgcn_sched_barrier(0);
                continue;
            }
            if constexpr (CF::PHASED) {
                static_for<Products<NP>::N>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    constexpr int pa = Products<NP>::a[q], pb = Products<NP>::b[q];
                    if constexpr (q > 0 && (first_use_a<NP>(q) || first_use_b<NP>(q)))
                        __builtin_amdgcn_sched_barrier(0);
                    if constexpr (first_use_a<NP>(q)) {
#pragma unroll
                        for (int i = 0; i < FM; ++i) {
                            const int row = wm * CF::WTM + i * 16 + frow;
                            fa[pa][i] = *reinterpret_cast<const V8 *>(as + pa * CF::A_PLANE + row * BK +
                                                                      swz_chunk<CPR>(row, s * 4 + fc) * 8);
                        }
                    }
                    if constexpr (first_use_b<NP>(q)) {
#pragma unroll
                        for (int j = 0; j < FN; ++j) {
                            const int row = wn * CF::WTN + j * 16 + frow;
                            fb[pb][j] = *reinterpret_cast<const V8 *>(bs + pb * CF::B_PLANE + row * BK +
                                                                      swz_chunk<CPR>(row, s * 4 + fc) * 8);
                        }
                    }
#pragma unroll
                    for (int i = 0; i < FM; ++i)
#pragma unroll
                        for (int j = 0; j < FN; ++j) acc[i][j] = FMT::mfma(fa[pa][i], fb[pb][j], acc[i][j]);
                });
                continue;
            }
#pragma unroll
            for (int p = 0; p < NP; ++p) {
#pragma unroll
                for (int i = 0; i < FM; ++i) {
                    const int row = wm * CF::WTM + i * 16 + frow;
                    fa[p][i] = *reinterpret_cast<const V8 *>(as + p * CF::A_PLANE + row * BK +
                                                                 swz_chunk<CPR>(row, s * 4 + fc) * 8);
                }
#pragma unroll
                for (int j = 0; j < FN; ++j) {
                    const int row = wn * CF::WTN + j * 16 + frow;
                    fb[p][j] = *reinterpret_cast<const V8 *>(bs + p * CF::B_PLANE + row * BK +
                                                                 swz_chunk<CPR>(row, s * 4 + fc) * 8);
                }
            }
#pragma unroll
            for (int q = 0; q < Products<NP>::N; ++q)
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int j = 0; j < FN; ++j)
                        acc[i][j] = FMT::mfma(fa[Products<NP>::a[q]][i], fb[Products<NP>::b[q]][j], acc[i][j]);
        }
        if (STAGES == 1) __syncthreads();
        if (!ILV && !ESTORE && more) store_stage(STAGES == 2 ? (buf ^ 1) : 0);
        __syncthreads();
    }

    // Epilogue: C = alpha * acc (+ beta * C), rows 4*(l/16)+r, column l%16 of each fragment;
    // scaled formats first undo the row and column scales: acc * 2^(e_m + e_n).
    // C is read in batches of IB x JB fragments (64 values per lane), all
    // loaded before any is stored: a load cannot move above a store to C
    // that might alias it, so every batch exposes one memory latency.  With
    // 256 accumulators per lane there is no room for a batch (it spills),
    // nor with 16 waves: those variants interleave.
    const bool interior = (m0 + BM <= M) && (n0 + BN <= N);
    const bool has_epi = !GR && !part && epi.active();
    constexpr bool BATCH = FM * FN * 4 < 256;
    constexpr int JB = !BATCH ? 1 : (CF::NT >= 1024 ? 1 : FN);
    constexpr int IB = (JB < FN) ? 1 : (64 / (4 * FN) < 1 ? 1 : (64 / (4 * FN) > FM ? FM : 64 / (4 * FN)));
    static_assert(FM % IB == 0, "whole batches of fragment rows");
#pragma unroll
    for (int i0 = 0; i0 < FM; i0 += IB)
#pragma unroll
    for (int j0 = 0; j0 < FN; j0 += JB) {
        float cv[IB][JB][4];
        if (BATCH && be != 0.0f) {
#pragma unroll
            for (int ii = 0; ii < IB; ++ii) {
                const int mb = m0 + wm * CF::WTM + (i0 + ii) * 16 + (lane >> 4) * 4;
#pragma unroll
                for (int jj = 0; jj < JB; ++jj) {
                    const int n = n0 + wn * CF::WTN + (j0 + jj) * 16 + (lane & 15);
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        cv[ii][jj][r] = (interior || (mb + r < M && n < N)) ? Co[(long)(mb + r) * ldo + n] : 0.0f;
                }
            }
        }
#pragma unroll
        for (int ii = 0; ii < IB; ++ii) {
            const int i = i0 + ii;
            const int mb = m0 + wm * CF::WTM + i * 16 + (lane >> 4) * 4;
#pragma unroll
            for (int jj = 0; jj < JB; ++jj) {
                const int j = j0 + jj;
                const int n = n0 + wn * CF::WTN + j * 16 + (lane & 15);
                int en = 0;
                if constexpr (FMT::SCALED) en = row_exp(amax_b[n]);   // n < Np: inside the array
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = mb + r;
                    if (interior || (m < M && n < N)) {
                        float a = acc[i][j][r];
                        if constexpr (FMT::SCALED) a = ldexpf(a, row_exp(amax_a[m]) + en);
                        float *cp = Co + (long)m * ldo + n;
                        float v = (be == 0.0f) ? al * a : fmaf(al, a, be * (BATCH ? cv[ii][jj][r] : *cp));
                        if (has_epi) v = apply_epilogue(epi, v, m, n);
                        *cp = v;
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Variants: (tile, BK) per form.
template <int NP, class FMT = FmtBf16>
struct Table;
// V0-V3: 4 waves (2 x 2), one per SIMD.  V4, V5: 8 or 16 waves, 2 or 4 per
// SIMD, which hide each other's LDS and global latency (measured at 8192^3:
// 15-20% faster core than V0).  V6: 8 waves with streamed reads, the
// interleaved stage copy and (NP >= 2) the early barrier; 6-10% faster than
// V5 at 8192^3.  NP = 1, 2: 256x256x32 tiles (32 / 64 KiB per stage); NP = 3:
// 256x128x32 (72 KiB per stage).  V7: V6's schedule on 128x128 tiles (8
// waves of 64x32) for grids too small for V6.  Smaller tiles serve small
// grids.
template <>
struct Table<1> {
    using V0 = CoreCfg<256, 256, 32, 1>;
    using V1 = CoreCfg<128, 128, 32, 1>;
    using V2 = CoreCfg<256, 128, 32, 1>;
    using V3 = CoreCfg<64, 64, 32, 1>;
    using V4 = CoreCfg<256, 256, 32, 1, FmtBf16, 2, 4, false>;
    using V5 = CoreCfg<256, 256, 32, 1, FmtBf16, 4, 4, true>;
    using V6 = CoreCfg<256, 256, 32, 1, FmtBf16, 2, 4, true, true, true>;
    using V7 = CoreCfg<128, 128, 32, 1, FmtBf16, 2, 4, true, true, true>;
};
template <>
struct Table<2> {
    using V0 = CoreCfg<256, 256, 32, 2>;
    using V1 = CoreCfg<128, 128, 32, 2>;
    using V2 = CoreCfg<256, 128, 32, 2>;
    using V3 = CoreCfg<64, 64, 32, 2>;
    using V4 = CoreCfg<256, 256, 32, 2, FmtBf16, 2, 4, false>;
    using V5 = CoreCfg<256, 256, 32, 2, FmtBf16, 4, 4, true>;
    using V6 = CoreCfg<256, 256, 32, 2, FmtBf16, 2, 4, true, true, true, true>;
    using V7 = CoreCfg<128, 128, 32, 2, FmtBf16, 2, 4, true, true, true, true>;
};
template <>
struct Table<3> {
    using V0 = CoreCfg<256, 128, 32, 3>;
    using V1 = CoreCfg<128, 128, 32, 3>;
    using V2 = CoreCfg<128, 256, 32, 3>;
    using V3 = CoreCfg<64, 64, 32, 3>;
    using V4 = CoreCfg<256, 128, 32, 3, FmtBf16, 2, 4, true>;
    using V5 = CoreCfg<256, 128, 32, 3, FmtBf16, 4, 2, true>;
    using V6 = CoreCfg<256, 128, 32, 3, FmtBf16, 2, 4, true, true, true, true>;
    using V7 = CoreCfg<128, 128, 32, 3, FmtBf16, 2, 4, true, true, true, true>;
};
template <>
struct Table<2, FmtF16> {
    using V0 = CoreCfg<256, 256, 32, 2, FmtF16>;
    using V1 = CoreCfg<128, 128, 32, 2, FmtF16>;
    using V2 = CoreCfg<256, 128, 32, 2, FmtF16>;
    using V3 = CoreCfg<64, 64, 32, 2, FmtF16>;
    using V4 = CoreCfg<256, 256, 32, 2, FmtF16, 2, 4, false>;
    using V5 = CoreCfg<256, 256, 32, 2, FmtF16, 4, 4, true>;
    using V6 = CoreCfg<256, 256, 32, 2, FmtF16, 2, 4, true, true, true, true>;
    using V7 = CoreCfg<128, 128, 32, 2, FmtF16, 2, 4, true, true, true, true>;
};
constexpr int kNumVariants = 8;

template <class CF>
const char *cfg_name() {
    // Formatted once per instantiation; function-local static initialisation is thread-safe.
    static const std::string name = [] {
        char buf[48];
        snprintf(buf, sizeof buf, "%sx%d_nt_%dx%dx%d_w%dx%d%s", CF::FMT::name,
                 CF::NP == 1 ? 1 : (CF::NP == 2 ? 3 : 6), CF::BM, CF::BN, CF::BK, CF::WGM, CF::WGN,
                 CF::EARLY ? "_st_early" : (CF::STREAM ? "_st" : (CF::PHASED ? "_ph" : "")));
        return std::string(buf);
    }();
    return name.c_str();
}

int planes_of(int mode) {
    switch (mode) {
        case UGEMM_MODE_BF16X1: return 1;
        case UGEMM_MODE_BF16X3: return 2;
        case UGEMM_MODE_BF16X6: return 3;
        case UGEMM_MODE_FP16X3: return 2;
        default: return 0;
    }
}

static long round_up(long x, long a) { return (x + a - 1) / a * a; }

// Scaled formats: amax (rows_pad entries, zeroed by the caller) receives the
// per-row largest magnitudes before the split reads them.
// split == false: only the row maxima (scaled formats); the core splits the
// operand itself while staging it.
template <int NP, class FMT>
int prep_operand(const float *X, long ld, bool kcont, int rows, int K, typename FMT::E *out, long plane_stride,
                 int rows_pad, int k_pad, unsigned *amax, hipStream_t s, bool split = true, bool amax_done = false) {
    const bool vec = ugemm_aligned16(X) && (ld % 4 == 0);
    if constexpr (FMT::SCALED) {
        if (amax_done) goto planes;   // the row maxima were computed by an earlier launch
        if (split && kcont && rows > 0 && K > 0 && k_pad <= 8192 && rows_pad % 2 == 0) {
            // Row maxima and planes in one pass (one block per two padded rows).
            hipLaunchKernelGGL((split_prep_kcont_rowscaled<NP, FMT>), dim3((unsigned)(rows_pad / 2)), dim3(256), 0, s, X,
                               ld, rows, K, vec, out, plane_stride, rows_pad, k_pad, amax);
            UGEMM_HIP_TRY(hipGetLastError());
            return UGEMM_OK;
        }
        if (rows > 0 && K > 0) {
            if (kcont) {
                const long kchunks = (K + 2047) / 2048;
                hipLaunchKernelGGL(row_amax_kcont, dim3((unsigned)(rows * kchunks)), dim3(256), 0, s, X, ld, rows, K,
                                   vec, kchunks, amax);
            } else {
                const long kblocks = (K + kAmaxRcontSpan - 1) / kAmaxRcontSpan, cblocks = (rows + 255) / 256;
                hipLaunchKernelGGL(row_amax_rcont, dim3((unsigned)(cblocks * kblocks)), dim3(256), 0, s, X, ld, rows,
                                   K, vec, kblocks, amax);
            }
        }
    }
    if (!split) {
        UGEMM_HIP_TRY(hipGetLastError());
        return UGEMM_OK;
    }
planes:
    if (kcont && rows_pad % 64 == 0 && (k_pad + 127) / 128 < 65536) {
        dim3 grid((unsigned)(rows_pad / 64), (unsigned)((k_pad + 127) / 128));
        hipLaunchKernelGGL((split_prep_kcont_tiles<NP, FMT>), grid, dim3(256), 0, s, X, ld, rows, K, vec, out,
                           plane_stride, rows_pad, k_pad, (const unsigned *)amax);
    } else if (kcont) {
        const long total = (long)rows_pad * (k_pad / 8);
        // One chunk per thread (the grid-stride loop only matters beyond 2^28
        // chunks): a grid capped at 8192 blocks, four chunks per thread, ran
        // 13% slower at 8192^3 (fewer loads in flight).
        long blocks = (total + 255) / 256;
        if (blocks > (1L << 20)) blocks = 1L << 20;
        hipLaunchKernelGGL((split_prep_kcont<NP, FMT>), dim3((unsigned)blocks), dim3(256), 0, s, X, ld, rows, K, vec,
                           out, plane_stride, rows_pad, k_pad, (const unsigned *)amax);
    } else {
        dim3 grid((unsigned)((rows_pad + 63) / 64), (unsigned)((k_pad + 63) / 64));
        hipLaunchKernelGGL((split_prep_rcont<NP, FMT>), grid, dim3(256), 0, s, X, ld, rows, K, vec, out, plane_stride,
                           rows_pad, k_pad, (const unsigned *)amax);
    }
    UGEMM_HIP_TRY(hipGetLastError());
    return UGEMM_OK;
}

// Split-K count of the core's grid, with *kc the K chunk (whole 32-wide K tiles).
template <class CF>
int core_splits(const Gemm<float> &g, int tiles_m, int tiles_n, int Kp, int *kc) {
    *kc = Kp;
    return split_count(g.splits, (long)tiles_m * tiles_n, Kp, CF::BK, kc, CF::NT >= 512 ? 1 : 2);
}

// The core launch (and, split along K, the reduction of the partial products
// in ws, splits * M * N floats that the caller allocated with the planes).
template <class CF, int FA, int FB>
int launch_core(const Gemm<float> &g, const typename CF::FMT::E *Ap, long a_plane, int Mp, const typename CF::FMT::E *Bp,
                long b_plane, int Np, int Kp, int tiles_m, int tiles_n, const unsigned *amax_a,
                const unsigned *amax_b, int splits, int kc, float *ws, hipStream_t s) {
    auto kern = split_nt_core<CF, FA, FB>;
    // Once per kernel instantiation (thread-safe static initialisation).
    static const hipError_t attr = hipFuncSetAttribute((const void *)kern,
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, CF::LDS_BYTES);
    if (attr != hipSuccess) return hip_fail(attr, "hipFuncSetAttribute");
    // Vector loads of a fused operand: 16 bytes (k-contiguous) or 8 bytes (row-contiguous).
    auto vec_ok = [](const float *X, long ld, int fuse) {
        return fuse == FUSE_RCONT ? ((reinterpret_cast<uintptr_t>(X) & 7) == 0 && ld % 2 == 0)
                                  : (ugemm_aligned16(X) && ld % 4 == 0);
    };
    const int vec_ab = (vec_ok(g.A, g.lda, FA) ? 1 : 0) | (vec_ok(g.B, g.ldb, FB) ? 2 : 0);
    timing_begin(s);
    hipLaunchKernelGGL(kern, dim3((unsigned)(tiles_m * tiles_n), (unsigned)splits), dim3(CF::NT), CF::LDS_BYTES, s,
                       g.M, g.N, Kp, Ap, a_plane, Mp, Bp, b_plane, Np, g.alpha, g.beta, g.C, g.ldc, tiles_m, tiles_n,
                       amax_a, amax_b, g.epi, g.A, g.lda, g.B, g.ldb, g.K, vec_ab, kc / CF::BK, ws,
                       (const GroupedSplitProblem *)nullptr, (const GroupedSplitTile *)nullptr);
    timing_end(s);
    hipError_t e = hipGetLastError();
    int rc = e == hipSuccess ? UGEMM_OK : hip_fail(e, "split core launch");
    if (!rc && splits > 1) rc = launch_splitk_reduce_f32(g.M, g.N, splits, g.alpha, ws, g.beta, g.C, g.ldc, g.epi, s);
    return rc;
}

// One workspace per call, allocated before the first launch: the planes and
// row maxima (base_bytes, then 256 bytes of slack), then the split-K partial
// products at the next 256-byte boundary.
static size_t splitk_ws_offset(size_t base_bytes) { return (base_bytes + 256 + 255) / 256 * 256; }
static size_t splitk_ws_bytes(int splits, const Gemm<float> &g) {
    return splits > 1 ? (size_t)splits * g.M * g.N * sizeof(float) : 0;
}

// Workspace: NP planes of A (Mp x Kp) and of B (Np x Kp) -- none for an
// operand the core splits itself -- then (scaled formats) Mp + Np row maxima.
template <class CF>
int run_cfg(const Gemm<float> &g, hipStream_t s) {
    using FMT = typename CF::FMT;
    using E = typename FMT::E;
    constexpr int NP = CF::NP;
    const long Mp = round_up(g.M, CF::BM), Np = round_up(g.N, CF::BN), Kp = round_up(g.K, 64);
    // A k-contiguous FP32 source (A not transposed, B transposed) read by at
    // most two tiles along the other dimension is split in the core (8-wave
    // variants): its preparation pass would cost as much as the GEMM reads.
    // Operands re-read by many tiles stay on prepared planes, whose blocked
    // layout the core streams contiguously (measured: splitting a re-read
    // operand in the core is ~50% slower at 8192^3).
    // A row-contiguous source (A transposed, B not) is transposed in
    // registers on the way; that needs one or two rows per thread.  At most one
    // operand is fused per call (the larger, when both qualify).
    constexpr int RW_A = CF::BM * CF::CPR / CF::NT, RW_B = CF::BN * CF::CPR / CF::NT;
    constexpr bool RC_A = RW_A == 1 || RW_A == 2, RC_B = RW_B == 1 || RW_B == 2;
    bool fa = CF::FUSABLE && Np / CF::BN <= 2 && (!g.ta || RC_A);
    bool fb = CF::FUSABLE && Mp / CF::BM <= 2 && (g.tb || RC_B);
    if (fa && fb) (g.M >= g.N ? fb : fa) = false;
    const int fa_mode = !fa ? FUSE_NONE : (g.ta ? FUSE_RCONT : FUSE_KCONT);
    const int fb_mode = !fb ? FUSE_NONE : (g.tb ? FUSE_KCONT : FUSE_RCONT);
    const long a_plane = Mp * Kp, b_plane = Np * Kp;
    const size_t a_bytes = fa ? 0 : (size_t)NP * a_plane * sizeof(E), b_bytes = fb ? 0 : (size_t)NP * b_plane * sizeof(E);
    const size_t amax_bytes = FMT::SCALED ? (size_t)(Mp + Np) * sizeof(unsigned) : 0;
    const int tiles_m = (int)(Mp / CF::BM), tiles_n = (int)(Np / CF::BN);
    if ((long)tiles_m * tiles_n > 0x7fffffffL) return fail(UGEMM_ERR_ARG, "too many tiles");
    if (FMT::SCALED && (long)g.M * ((g.K + 2047) / 2048) > 0x7fffffffL)
        return fail(UGEMM_ERR_ARG, "too many row blocks");
    int kc;
    const int splits = core_splits<CF>(g, tiles_m, tiles_n, (int)Kp, &kc);
    const size_t ws_off = splitk_ws_offset(a_bytes + b_bytes + amax_bytes);
    void *ws = nullptr;
    int rc = stream_alloc(&ws, ws_off + splitk_ws_bytes(splits, g), s);
    if (rc) return rc;
    E *Ap = (E *)ws, *Bp = (E *)((char *)ws + a_bytes);
    unsigned *amax_a = FMT::SCALED ? (unsigned *)((char *)ws + a_bytes + b_bytes) : nullptr;
    unsigned *amax_b = FMT::SCALED ? amax_a + Mp : nullptr;
    float *part = splits > 1 ? (float *)((char *)ws + ws_off) : nullptr;
    if (FMT::SCALED) {
        // A fill kernel rather than a memset: callers may capture this call in a graph.
        hipLaunchKernelGGL(fill_zero_u32, dim3((unsigned)std::min<long>((Mp + Np + 255) / 256, 1024)), dim3(256), 0,
                           s, amax_a, Mp + Np);
    }
    // FP16X3 NN (A k-contiguous with K <= 8192, B row-contiguous, both on
    // planes): A's fused row maximum + split and B's column maxima in one
    // launch, then B's split; otherwise one preparation per operand.
    bool merged = false;
    if constexpr (FMT::SCALED) {
        if (!fa && !fb && !g.ta && !g.tb && g.M > 0 && g.N > 0 && g.K > 0 && Kp <= 8192) {
            const long kblocks = (g.K + kAmaxRcontSpan - 1) / kAmaxRcontSpan, cblocks = (g.N + 255) / 256;
            const bool vec_a = ugemm_aligned16(g.A) && g.lda % 4 == 0, vec_b = ugemm_aligned16(g.B) && g.ldb % 4 == 0;
            hipLaunchKernelGGL((split_prep_rowscaled_a_amax_b<NP, FMT>), dim3((unsigned)(cblocks * kblocks + Mp / 2)),
                               dim3(256), 0, s, g.A, g.lda, g.M, g.K, vec_a, Ap, a_plane, (int)Mp, (int)Kp, amax_a,
                               g.B, g.ldb, g.N, vec_b, kblocks, amax_b, (int)(cblocks * kblocks));
            const hipError_t e = hipGetLastError();
            rc = e == hipSuccess ? UGEMM_OK : hip_fail(e, "split preparation");
            if (!rc) rc = prep_operand<NP, FMT>(g.B, g.ldb, false, g.N, g.K, Bp, b_plane, (int)Np, (int)Kp, amax_b, s,
                                                true, true);
            merged = true;
        }
    }
    if constexpr (!FMT::SCALED) {
        if (!fa && !fb) {
            // Both operands on planes: one preparation launch.
            auto job = [&](const float *X, long ld, bool kcont, int rows, E *out, long plane, long rows_pad,
                           int block0) {
                PrepJob j{};
                j.X = X;
                j.ld = ld;
                j.out = out;
                j.plane_stride = plane;
                j.rows = rows;
                j.K = g.K;
                j.rows_pad = (int)rows_pad;
                j.k_pad = (int)Kp;
                j.kcont = kcont;
                j.vec = ugemm_aligned16(X) && ld % 4 == 0;
                j.block0 = block0;
                j.nbx = (int)(rows_pad / 64);
                j.amax = nullptr;
                return j;
            };
            const long ka = !g.ta ? (Kp + 127) / 128 : (Kp + 63) / 64, kb = g.tb ? (Kp + 127) / 128 : (Kp + 63) / 64;
            const long blocks_a = Mp / 64 * ka, blocks = blocks_a + Np / 64 * kb;
            if (Mp % 64 == 0 && Np % 64 == 0 && blocks <= 0x7fffffffL) {
                const PrepJob ja = job(g.A, g.lda, !g.ta, g.M, Ap, a_plane, Mp, 0);
                const PrepJob jb = job(g.B, g.ldb, g.tb, g.N, Bp, b_plane, Np, (int)blocks_a);
                hipLaunchKernelGGL((split_prep_pair<NP, FMT>), dim3((unsigned)blocks), dim3(256), 0, s, ja, jb);
                const hipError_t e = hipGetLastError();
                rc = e == hipSuccess ? UGEMM_OK : hip_fail(e, "split preparation");
                merged = true;
            }
        }
    }
    if (!merged) {
        rc = prep_operand<NP, FMT>(g.A, g.lda, !g.ta, g.M, g.K, Ap, a_plane, (int)Mp, (int)Kp, amax_a, s, !fa);
        if (!rc) rc = prep_operand<NP, FMT>(g.B, g.ldb, g.tb, g.N, g.K, Bp, b_plane, (int)Np, (int)Kp, amax_b, s, !fb);
    }
    if (!rc) {
        const unsigned *ma = amax_a, *mb = amax_b;
#define UGEMM_CORE(FA_, FB_) \
    launch_core<CF, FA_, FB_>(g, Ap, a_plane, (int)Mp, Bp, b_plane, (int)Np, (int)Kp, tiles_m, tiles_n, ma, mb, splits, \
                              kc, part, s)
        if constexpr (CF::FUSABLE) {
            if (fa_mode == FUSE_KCONT) rc = UGEMM_CORE(FUSE_KCONT, FUSE_NONE);
            else if (fb_mode == FUSE_KCONT) rc = UGEMM_CORE(FUSE_NONE, FUSE_KCONT);
            else if (fa_mode == FUSE_RCONT) {
                if constexpr (RC_A) rc = UGEMM_CORE(FUSE_RCONT, FUSE_NONE);
            } else if (fb_mode == FUSE_RCONT) {
                if constexpr (RC_B) rc = UGEMM_CORE(FUSE_NONE, FUSE_RCONT);
            } else rc = UGEMM_CORE(FUSE_NONE, FUSE_NONE);
        } else {
            rc = UGEMM_CORE(FUSE_NONE, FUSE_NONE);
        }
#undef UGEMM_CORE
    }
    int rc2 = stream_free(ws, s);
    return rc ? rc : rc2;
}

template <int NP, class FMT = FmtBf16>
int launch_np(const Gemm<float> &g, int variant, hipStream_t s) {
    using T = Table<NP, FMT>;
    switch (variant) {
        case 0: return run_cfg<typename T::V0>(g, s);
        case 1: return run_cfg<typename T::V1>(g, s);
        case 2: return run_cfg<typename T::V2>(g, s);
        case 3: return run_cfg<typename T::V3>(g, s);
        case 4: return run_cfg<typename T::V4>(g, s);
        case 5: return run_cfg<typename T::V5>(g, s);
        case 6: return run_cfg<typename T::V6>(g, s);
        case 7: return run_cfg<typename T::V7>(g, s);
        default: return fail(UGEMM_ERR_ARG, "split form: no variant %d", variant);
    }
}

// Convolution as a split-form GEMM: C (channels x pixels) = weights (M x K,
// k-contiguous) times the im2col operand, whose planes are gathered from the
// image by split_prep_im2col; then the core with the epilogue.
template <class CF>
int run_cfg_conv(const Gemm<float> &g, const ConvDesc &cd, hipStream_t s) {
    using FMT = typename CF::FMT;
    using E = typename FMT::E;
    constexpr int NP = CF::NP;
    const long Mp = round_up(g.M, CF::BM), Np = round_up(g.N, CF::BN), Kp = round_up(g.K, 64);
    const long a_plane = Mp * Kp, b_plane = Np * Kp;
    const size_t a_bytes = (size_t)NP * a_plane * sizeof(E), b_bytes = (size_t)NP * b_plane * sizeof(E);
    const size_t amax_bytes = FMT::SCALED ? (size_t)(Mp + Np) * sizeof(unsigned) : 0;
    const int tiles_m = (int)(Mp / CF::BM), tiles_n = (int)(Np / CF::BN);
    if ((long)tiles_m * tiles_n > 0x7fffffffL) return fail(UGEMM_ERR_ARG, "too many tiles");
    if (FMT::SCALED && (long)g.M * ((g.K + 2047) / 2048) > 0x7fffffffL)
        return fail(UGEMM_ERR_ARG, "too many row blocks");
    int kc;
    const int splits = core_splits<CF>(g, tiles_m, tiles_n, (int)Kp, &kc);
    const size_t ws_off = splitk_ws_offset(a_bytes + b_bytes + amax_bytes);
    void *ws = nullptr;
    int rc = stream_alloc(&ws, ws_off + splitk_ws_bytes(splits, g), s);
    if (rc) return rc;
    E *Ap = (E *)ws, *Bp = (E *)((char *)ws + a_bytes);
    unsigned *amax_a = FMT::SCALED ? (unsigned *)((char *)ws + a_bytes + b_bytes) : nullptr;
    unsigned *amax_b = FMT::SCALED ? amax_a + Mp : nullptr;
    float *part = splits > 1 ? (float *)((char *)ws + ws_off) : nullptr;
    if (FMT::SCALED)
        hipLaunchKernelGGL(fill_zero_u32, dim3((unsigned)std::min<long>((Mp + Np + 255) / 256, 1024)), dim3(256), 0,
                           s, amax_a, Mp + Np);
    rc = prep_operand<NP, FMT>(g.A, g.lda, true, g.M, g.K, Ap, a_plane, (int)Mp, (int)Kp, amax_a, s);
    if (!rc) {
        constexpr int KSPAN = 32;   // short spans: enough blocks for small images
        if (FMT::SCALED)
            hipLaunchKernelGGL(row_amax_im2col, dim3((unsigned)((g.N + 255) / 256), (unsigned)((g.K + KSPAN - 1) / KSPAN)),
                               dim3(256), 0, s, cd, g.N, g.K, KSPAN, amax_b);
        const long total = Np * (Kp / 8);
        const long blocks = std::max<long>(1, std::min<long>((total + 255) / 256, 64L * device_cu_count()));
        hipLaunchKernelGGL((split_prep_im2col<NP, FMT>), dim3((unsigned)blocks), dim3(256), 0, s, cd, g.N, g.K, Bp,
                           b_plane, (int)Np, (int)Kp, (const unsigned *)amax_b);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = hip_fail(e, "conv: operand gather");
    }
    if (!rc)
        rc = launch_core<CF, FUSE_NONE, FUSE_NONE>(g, Ap, a_plane, (int)Mp, Bp, b_plane, (int)Np, (int)Kp, tiles_m,
                                                   tiles_n, amax_a, amax_b, splits, kc, part, s);
    const int rc2 = stream_free(ws, s);
    return rc ? rc : rc2;
}

template <int NP, class FMT = FmtBf16>
int launch_np_conv(const Gemm<float> &g, const ConvDesc &cd, int variant, hipStream_t s) {
    using T = Table<NP, FMT>;
    switch (variant) {
        case 0: return run_cfg_conv<typename T::V0>(g, cd, s);
        case 1: return run_cfg_conv<typename T::V1>(g, cd, s);
        case 2: return run_cfg_conv<typename T::V2>(g, cd, s);
        case 3: return run_cfg_conv<typename T::V3>(g, cd, s);
        case 4: return run_cfg_conv<typename T::V4>(g, cd, s);
        case 5: return run_cfg_conv<typename T::V5>(g, cd, s);
        case 6: return run_cfg_conv<typename T::V6>(g, cd, s);
        case 7: return run_cfg_conv<typename T::V7>(g, cd, s);
        default: return fail(UGEMM_ERR_ARG, "split form: no variant %d", variant);
    }
}

template <int NP, class FMT = FmtBf16>
const char *name_np(int v) {
    using T = Table<NP, FMT>;
    switch (v) {
        case 0: return cfg_name<typename T::V0>();
        case 1: return cfg_name<typename T::V1>();
        case 2: return cfg_name<typename T::V2>();
        case 3: return cfg_name<typename T::V3>();
        case 4: return cfg_name<typename T::V4>();
        case 5: return cfg_name<typename T::V5>();
        case 6: return cfg_name<typename T::V6>();
        case 7: return cfg_name<typename T::V7>();
        default: return nullptr;
    }
}

}  // namespace

int mfma_variant_count(int mode) { return planes_of(mode) ? kNumVariants : 0; }

const char *mfma_variant_name(int mode, int v) {
    if (v < 0 || v >= kNumVariants) return nullptr;
    switch (mode) {
        case UGEMM_MODE_BF16X1: return name_np<1>(v);
        case UGEMM_MODE_BF16X3: return name_np<2>(v);
        case UGEMM_MODE_BF16X6: return name_np<3>(v);
        case UGEMM_MODE_FP16X3: return name_np<2, FmtF16>(v);
        default: return nullptr;
    }
}

int mfma_pick_variant(const Gemm<float> &g, int mode) {
    const long cus = device_cu_count();
    const int np = planes_of(mode);
    // Largest tile whose grid still gives every CU work; there, the
    // multi-wave variant measured fastest for the form.
    const long big_m = 256, big_n = (np == 3) ? 128 : 256;
    const long tiles_big = ((g.M + big_m - 1) / big_m) * ((g.N + big_n - 1) / big_n);
    if (tiles_big >= cus) {
        // A skinny side whose operand (either layout) the 8-wave variant
        // splits in the core (no preparation pass): V4.  Otherwise two and
        // three planes run fastest on the streamed V6, one plane on the
        // 16-wave V5.
        const long tm = (g.M + big_m - 1) / big_m, tn = (g.N + big_n - 1) / big_n;
        const bool fuse = np <= 2 && (tn <= 2 || tm <= 2);
        return fuse ? 4 : (np >= 2 ? 6 : 5);
    }
    // 128 x 128 tiles: the 8-wave streamed V7 (2048^3: 12-24% faster per
    // call than the 4-wave V1, whose one wave per SIMD hides no latency).
    const long t128 = (long)((g.M + 127) / 128) * ((g.N + 127) / 128);
    if (t128 >= cus / 2) return 7;
    return 3;
}

// ---------------------------------------------------------------------------
// Grouped split-form plans (grouped.hip): any variant (no operand is split
// in the core here); the FP16 form adds the row-maximum pass.
namespace {
template <class CF, int FB = FUSE_NONE>
int grouped_core(const PrepJob *jobs, const int *block_job, int prep_blocks, const GroupedSplitProblem *probs,
                 const GroupedSplitTile *tiles, int ntiles, hipStream_t s, const GroupedScales *sc) {
    using FMT = typename CF::FMT;
    if (FMT::SCALED) {
        // Row maxima first (a fill kernel, not a memset: the launches are captured in graphs).
        hipLaunchKernelGGL(fill_zero_u32, dim3((unsigned)std::min<long>((sc->count + 255) / 256, 4096)), dim3(256), 0,
                           s, sc->amax, sc->count);
        if (sc->blocks > 0)
            hipLaunchKernelGGL(grouped_row_amax, dim3((unsigned)sc->blocks), dim3(256), 0, s, sc->jobs,
                               sc->block_job);
    }
    if (prep_blocks > 0)
        hipLaunchKernelGGL((grouped_split_prep<CF::NP, FMT>), dim3((unsigned)prep_blocks), dim3(256), 0, s, jobs,
                           block_job);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "grouped split-form preparation");
    auto kern = split_nt_core<CF, FUSE_NONE, FB, true>;
    static const hipError_t attr = hipFuncSetAttribute((const void *)kern,
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, CF::LDS_BYTES);
    if (attr != hipSuccess) return hip_fail(attr, "hipFuncSetAttribute");
    timing_begin(s);
    hipLaunchKernelGGL(kern, dim3((unsigned)ntiles), dim3(CF::NT), CF::LDS_BYTES, s, 0, 0, 0,
                       (const typename FMT::E *)nullptr, 0L, 0, (const typename FMT::E *)nullptr, 0L, 0, 0.0f, 0.0f,
                       (float *)nullptr, 0L, 0, 0, (const unsigned *)nullptr, (const unsigned *)nullptr, Epilogue(),
                       (const float *)nullptr, 0L, (const float *)nullptr, 0L, 0, 0, 0, (float *)nullptr, probs, tiles);
    timing_end(s);
    e = hipGetLastError();
    return e == hipSuccess ? UGEMM_OK : hip_fail(e, "grouped split-form core");
}

// Grouped-only variants (kNumVariants + i): base variant split_b_base(np, i)
// with B split in the core -- the 4-wave tiles with 128 B rows (each thread
// stages two rows of a row-contiguous B): 128 x 128, and 256 x 128; then
// the 8-wave streamed 128 x 128 (one row per thread).
constexpr int kNumSplitB = 3;
constexpr int split_b_base(int np, int i) { return i == 0 ? 1 : i == 2 ? 7 : (np == 3 ? 0 : 2); }

template <int NP, class FMT = FmtBf16>
int grouped_np(int variant, int b_kcont, const PrepJob *jobs, const int *block_job, int prep_blocks,
               const GroupedSplitProblem *probs, const GroupedSplitTile *tiles, int ntiles, hipStream_t s,
               const GroupedScales *sc) {
    using T = Table<NP, FMT>;
#define UGEMM_SPLIT_B(CFG)                                                                                    \
    (b_kcont ? grouped_core<CFG, FUSE_KCONT>(jobs, block_job, prep_blocks, probs, tiles, ntiles, s, sc)           \
             : grouped_core<CFG, FUSE_RCONT>(jobs, block_job, prep_blocks, probs, tiles, ntiles, s, sc))
    static_assert(split_b_base(NP, 0) == 1 && split_b_base(NP, 1) == (NP == 3 ? 0 : 2) && split_b_base(NP, 2) == 7,
                  "split-B table");
    switch (variant) {
        case kNumVariants + 0: return UGEMM_SPLIT_B(typename T::V1);
        case kNumVariants + 1:
            if constexpr (NP == 3) return UGEMM_SPLIT_B(typename T::V0);
            else return UGEMM_SPLIT_B(typename T::V2);
        case kNumVariants + 2: return UGEMM_SPLIT_B(typename T::V7);
        default: break;
    }
#undef UGEMM_SPLIT_B
    switch (variant) {
        case 0: return grouped_core<typename T::V0>(jobs, block_job, prep_blocks, probs, tiles, ntiles, s, sc);
        case 1: return grouped_core<typename T::V1>(jobs, block_job, prep_blocks, probs, tiles, ntiles, s, sc);
        case 2: return grouped_core<typename T::V2>(jobs, block_job, prep_blocks, probs, tiles, ntiles, s, sc);
        case 3: return grouped_core<typename T::V3>(jobs, block_job, prep_blocks, probs, tiles, ntiles, s, sc);
        case 4: return grouped_core<typename T::V4>(jobs, block_job, prep_blocks, probs, tiles, ntiles, s, sc);
        case 5: return grouped_core<typename T::V5>(jobs, block_job, prep_blocks, probs, tiles, ntiles, s, sc);
        case 6: return grouped_core<typename T::V6>(jobs, block_job, prep_blocks, probs, tiles, ntiles, s, sc);
        case 7: return grouped_core<typename T::V7>(jobs, block_job, prep_blocks, probs, tiles, ntiles, s, sc);
        default: return fail(UGEMM_ERR_ARG, "grouped split form: no variant %d", variant);
    }
}

template <int NP, class FMT = FmtBf16>
void tile_of(int variant, int *bm, int *bn) {
    using T = Table<NP, FMT>;
    switch (variant) {
        case 0: *bm = T::V0::BM; *bn = T::V0::BN; break;
        case 1: *bm = T::V1::BM; *bn = T::V1::BN; break;
        case 2: *bm = T::V2::BM; *bn = T::V2::BN; break;
        case 3: *bm = T::V3::BM; *bn = T::V3::BN; break;
        case 4: *bm = T::V4::BM; *bn = T::V4::BN; break;
        case 5: *bm = T::V5::BM; *bn = T::V5::BN; break;
        case 6: *bm = T::V6::BM; *bn = T::V6::BN; break;
        default: *bm = T::V7::BM; *bn = T::V7::BN; break;
    }
}
}  // namespace

int mfma_grouped_variant_count(int mode) { return planes_of(mode) ? kNumVariants + kNumSplitB : 0; }

int mfma_grouped_splits_b(int mode, int variant) {
    return planes_of(mode) && variant >= kNumVariants && variant < kNumVariants + kNumSplitB ? 1 : 0;
}

const char *mfma_grouped_variant_name(int mode, int variant) {
    if (!planes_of(mode) || variant < 0 || variant >= kNumVariants + kNumSplitB) return nullptr;
    if (variant < kNumVariants) return mfma_variant_name(mode, variant);
    // Formatted once per (mode, variant); function-local static initialisation is thread-safe.
    static const std::string names[UGEMM_MODE_COUNT][kNumSplitB] = {
#define UGEMM_N(m, i) (std::string(mfma_variant_name(m, split_b_base(planes_of(m), i)) \
                                    ? mfma_variant_name(m, split_b_base(planes_of(m), i)) : "") + "_split_b")
        {UGEMM_N(0, 0), UGEMM_N(0, 1), UGEMM_N(0, 2)}, {UGEMM_N(1, 0), UGEMM_N(1, 1), UGEMM_N(1, 2)},
        {UGEMM_N(2, 0), UGEMM_N(2, 1), UGEMM_N(2, 2)}, {UGEMM_N(3, 0), UGEMM_N(3, 1), UGEMM_N(3, 2)},
        {UGEMM_N(4, 0), UGEMM_N(4, 1), UGEMM_N(4, 2)}, {UGEMM_N(5, 0), UGEMM_N(5, 1), UGEMM_N(5, 2)},
        {UGEMM_N(6, 0), UGEMM_N(6, 1), UGEMM_N(6, 2)}
#undef UGEMM_N
    };
    return names[mode][variant - kNumVariants].c_str();
}

int mfma_grouped_tile(int mode, int variant, int *bm, int *bn) {
    if (variant >= kNumVariants && variant < kNumVariants + kNumSplitB)
        variant = split_b_base(planes_of(mode), variant - kNumVariants);
    if (variant < 0 || variant >= kNumVariants) return fail(UGEMM_ERR_ARG, "grouped split form: no variant %d", variant);
    switch (mode) {
        case UGEMM_MODE_BF16X1: tile_of<1>(variant, bm, bn); return UGEMM_OK;
        case UGEMM_MODE_BF16X3: tile_of<2>(variant, bm, bn); return UGEMM_OK;
        case UGEMM_MODE_BF16X6: tile_of<3>(variant, bm, bn); return UGEMM_OK;
        case UGEMM_MODE_FP16X3: tile_of<2, FmtF16>(variant, bm, bn); return UGEMM_OK;
        default: return fail(UGEMM_ERR_UNSUPPORTED, "grouped: mode %d not provided", mode);
    }
}

int launch_mfma_grouped(int mode, int variant, int b_kcont, const PrepJob *jobs, const int *block_job,
                        int prep_blocks,
                        const GroupedSplitProblem *probs, const GroupedSplitTile *tiles, int ntiles, hipStream_t s,
                        const GroupedScales *sc) {
    if (ntiles == 0) return UGEMM_OK;
    switch (mode) {
        case UGEMM_MODE_BF16X1:
            return grouped_np<1>(variant, b_kcont, jobs, block_job, prep_blocks, probs, tiles, ntiles, s, sc);
        case UGEMM_MODE_BF16X3:
            return grouped_np<2>(variant, b_kcont, jobs, block_job, prep_blocks, probs, tiles, ntiles, s, sc);
        case UGEMM_MODE_BF16X6:
            return grouped_np<3>(variant, b_kcont, jobs, block_job, prep_blocks, probs, tiles, ntiles, s, sc);
        case UGEMM_MODE_FP16X3:
            if (!sc) return fail(UGEMM_ERR_ARG, "grouped FP16X3: no row-scale jobs");
            return grouped_np<2, FmtF16>(variant, b_kcont, jobs, block_job, prep_blocks, probs, tiles, ntiles, s, sc);
        default: return fail(UGEMM_ERR_UNSUPPORTED, "grouped: mode %d not provided", mode);
    }
}

int launch_mfma_conv(const Gemm<float> &g, const ConvDesc &cd, int mode, int variant, hipStream_t s) {
    // No operand is split in the core here (the im2col operand is gathered
    // into planes), so the picker's fusing variant gives way to the
    // plane-streaming one.
    int v = variant >= 0 ? variant : mfma_pick_variant(g, mode);
    if (variant < 0 && v == 4) v = planes_of(mode) == 2 ? 6 : 5;
    switch (mode) {
        case UGEMM_MODE_BF16X1: return launch_np_conv<1>(g, cd, v, s);
        case UGEMM_MODE_BF16X3: return launch_np_conv<2>(g, cd, v, s);
        case UGEMM_MODE_BF16X6: return launch_np_conv<3>(g, cd, v, s);
        case UGEMM_MODE_FP16X3: return launch_np_conv<2, FmtF16>(g, cd, v, s);
        default: return fail(UGEMM_ERR_UNSUPPORTED, "conv: mode %d is not provided", mode);
    }
}

int launch_mfma(const Gemm<float> &g, int mode, int variant, hipStream_t s) {
    switch (mode) {
        case UGEMM_MODE_BF16X1: return launch_np<1>(g, variant, s);
        case UGEMM_MODE_BF16X3: return launch_np<2>(g, variant, s);
        case UGEMM_MODE_BF16X6: return launch_np<3>(g, variant, s);
        case UGEMM_MODE_FP16X3: return launch_np<2, FmtF16>(g, variant, s);
        default: return fail(UGEMM_ERR_UNSUPPORTED, "mode %d is not a split form", mode);
    }
}

}  // namespace ugemm
