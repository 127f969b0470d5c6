// Grouped SGEMM: hundreds of independent, uneven row-major problems in one
// persistent launch (optionally replayed as a hipGraph).
//
// Plan (host, once): every problem is cut into BM x BN tiles of one kernel
// variant, listed in grouped order (8 tile-rows walked column by column, as
// in tile_coords); each tile costs ~ceil(K/BK) + 2 k-steps.  Tiles are sorted
// by cost, longest first (stably, so a problem's tiles stay together), and
// dealt in chunks of W = G/8 consecutive tiles to 8 queues, chunk c to queue
// c % 8 (G = CUs x resident blocks per CU, the persistent grid).
// Run: a one-block kernel zeroes the 8 queue heads, then the persistent
// kernel starts.  Workgroup b takes tiles from queue b % 8 -- the hardware's
// round-robin dispatch puts block b on XCD b % 8, so one XCD's workgroups
// take consecutive tiles of the same problems and share their A and B panels
// in its L2 -- by an atomic increment of the queue head; when that queue is
// empty it moves to the next one, and it exits after finding all 8 empty
// (a queue never refills within a launch, so every workgroup reaches the
// exit).  Taking tiles on demand balances the load whatever each tile really
// costs: measured tile times vary by +-20% around the k-step model (edge
// tiles, co-resident workgroups), which a static schedule cannot absorb.
// Each tile is computed whole by one workgroup, so results do not depend on
// which workgroup took it: replays are bitwise identical.  The heads are
// cleared by a kernel, not a memset, so the pair can be captured in a graph.
// Kernel body: the plain FP32 tile (fp32_tile.h), dispatched on the
// problem's layout; 128-bit operand loads are used for the problems whose A
// and B allow them.
// A plan's launches must not overlap in time (the queue heads belong to the
// plan): run one plan from one stream, or order the streams.
// Used for convolution-as-GEMM batches (the conv -> GEMM shape mapping of
// sgemm_ocl1.h:271-300: M = output channels, N = output pixels, K = k*k*Cin).
#include "internal.h"
#include "fp32_tile.h"
#include "mfma32_tile.h"

#include <algorithm>
#include <type_traits>
#include <utility>
#include <vector>

namespace ugemm {
namespace {

struct DevProblem {
    const float *A, *B;
    float *C;
    long lda, ldb, ldc;
    int M, N, K;          // K is 0 when alpha == 0 (A and B are then not read)
    float alpha, beta;
    int layout;           // bit 0: A transposed, bit 1: B transposed, bit 2: A and B allow 128-bit
                          // loads, bit 3: B is the implicit im2col operand cd (convolution)
    Epilogue epi;         // bias + activation (convolutions)
    ConvDesc cd;
};

struct DevTile {
    int problem, m0, n0, shape;   // shape: which tile of a mixed plan (0: the larger, 1: the smaller)
};

constexpr int kQueues = 8;   // one per XCD

__global__ void grouped_reset_heads(int *heads) {
    if (threadIdx.x < kQueues) heads[threadIdx.x] = 0;
}

// One tile of problem p at (m0, n0), dispatched on the problem's layout.
template <class CF, bool CONV>
__device__ __forceinline__ void grouped_tile(const DevProblem &p, int m0, int n0, float *lds) {
    const bool aligned = (p.layout & 4) != 0;
    if constexpr (IsM32<CF>::value && CONV) {   // FP32-MFMA convolution plans
        m32_tile<CF, false, false, 4, true>(p.M, p.N, p.K, p.alpha, p.A, p.lda, nullptr, 0, p.beta, p.C, p.ldc, m0,
                                            n0, lds, aligned, p.epi, p.cd);
    } else if constexpr (IsM32<CF>::value) {   // FP32-MFMA GEMM plans
        switch (p.layout & 3) {
            case 0:
                m32_tile<CF, false, false, 4>(p.M, p.N, p.K, p.alpha, p.A, p.lda, p.B, p.ldb, p.beta, p.C, p.ldc, m0,
                                              n0, lds, aligned);
                break;
            case 1:
                m32_tile<CF, true, false, 4>(p.M, p.N, p.K, p.alpha, p.A, p.lda, p.B, p.ldb, p.beta, p.C, p.ldc, m0,
                                             n0, lds, aligned);
                break;
            case 2:
                m32_tile<CF, false, true, 4>(p.M, p.N, p.K, p.alpha, p.A, p.lda, p.B, p.ldb, p.beta, p.C, p.ldc, m0,
                                             n0, lds, aligned);
                break;
            default:
                m32_tile<CF, true, true, 4>(p.M, p.N, p.K, p.alpha, p.A, p.lda, p.B, p.ldb, p.beta, p.C, p.ldc, m0,
                                            n0, lds, aligned);
                break;
        }
    } else if constexpr (CONV) {
        fp32_tile<CF, false, false, 4, true>(p.M, p.N, p.K, p.alpha, p.A, p.lda, nullptr, 0, p.beta, p.C, p.ldc, m0,
                                             n0, lds, aligned, p.epi, p.cd);
    } else {
        switch (p.layout & 3) {
            case 0:
                fp32_tile<CF, false, false, 4>(p.M, p.N, p.K, p.alpha, p.A, p.lda, p.B, p.ldb, p.beta, p.C, p.ldc, m0,
                                               n0, lds, aligned);
                break;
            case 1:
                fp32_tile<CF, true, false, 4>(p.M, p.N, p.K, p.alpha, p.A, p.lda, p.B, p.ldb, p.beta, p.C, p.ldc, m0,
                                              n0, lds, aligned);
                break;
            case 2:
                fp32_tile<CF, false, true, 4>(p.M, p.N, p.K, p.alpha, p.A, p.lda, p.B, p.ldb, p.beta, p.C, p.ldc, m0,
                                              n0, lds, aligned);
                break;
            default:
                fp32_tile<CF, true, true, 4>(p.M, p.N, p.K, p.alpha, p.A, p.lda, p.B, p.ldb, p.beta, p.C, p.ldc, m0,
                                             n0, lds, aligned);
                break;
        }
    }
}

template <class CF>
constexpr int lds_floats() {
    if constexpr (IsM32<CF>::value) return CF::kLdsFloats;
    else return 2 * CF::BK * (CF::LDA + CF::LDB);
}

// Waves per SIMD the persistent kernel is compiled for (register budget).
template <class CF>
constexpr int grouped_waves() {
    return IsM32<CF>::value ? 2 : 3;
}

// At least 3 waves per SIMD for the FP32 vector tiles (<= 168 VGPRs): the
// layout switch would otherwise push the 128 x 128 variant just past that
// limit; 2 for the FP32-MFMA tiles.
// CONV: a plan of convolutions (every problem has layout bit 3), else of GEMMs.
// CS: the tile of a mixed plan whose DevTile::shape is 1 (CF for shape 0);
// the same as CF for a single-shape plan.  Both shapes have CF::NT threads.
template <class CF, bool CONV, class CS = CF>
__global__ __launch_bounds__(CF::NT) __attribute__((amdgpu_waves_per_eu(grouped_waves<CF>()))) void grouped_fp32_kernel(
    const DevProblem *__restrict__ probs, const DevTile *__restrict__ tiles, const int *__restrict__ q_begin,
    int *heads) {
    static_assert(CS::NT == CF::NT, "the shapes of a mixed plan share the workgroup size");
    constexpr int LDS_F = lds_floats<CF>() > lds_floats<CS>() ? lds_floats<CF>() : lds_floats<CS>();
    __shared__ __attribute__((aligned(16))) float lds[LDS_F];
    __shared__ int s_tile;
    int q = blockIdx.x % kQueues;
    for (int empty = 0; empty < kQueues;) {
        if (threadIdx.x == 0) {
            const int base = q_begin[q], size = q_begin[q + 1] - base;
            const int i = size > 0 ? atomicAdd(&heads[q], 1) : 0;
            s_tile = i < size ? base + i : -1;
        }
        __syncthreads();
        const int i = s_tile;
        __syncthreads();   // every thread has read s_tile before it is written again
        if (i < 0) {
            q = (q + 1) % kQueues;
            ++empty;
            continue;
        }
        const DevTile tl = tiles[i];
        const DevProblem p = probs[tl.problem];
        if (std::is_same<CF, CS>::value || tl.shape == 0) grouped_tile<CF, CONV>(p, tl.m0, tl.n0, lds);
        else grouped_tile<CS, CONV>(p, tl.m0, tl.n0, lds);
        __syncthreads();   // the next tile's first LDS stage overwrites this one's
    }
}

using G0 = Fp32Cfg<128, 128, 16, 8, 8>;
using G1 = Fp32Cfg<64, 64, 16, 4, 4>;
using G2 = Fp32Cfg<128, 64, 16, 8, 4>;
using G3 = Fp32Cfg<64, 128, 16, 4, 8>;
// Variant 4 ("mixed"): every problem gets the 128 x 128 tile (G0) or the
// 64 x 128 tile (G3), whichever covers its rows with less padding (the
// larger, more efficient tile when it pads at most 10% more).
constexpr int kNumVariants = 5, kMixed = 4;
// The same five tile choices on the FP32 matrix cores (UGEMM_MODE_FP32_MFMA
// plans), 4 waves each; variant 4 mixes GM0 and GM3 per problem.
using GM0 = M32Cfg<128, 128, 16, 2, 2>;
using GM1 = M32Cfg<64, 64, 16, 2, 2>;
using GM2 = M32Cfg<128, 64, 16, 2, 2>;
using GM3 = M32Cfg<64, 128, 16, 2, 2>;

struct VariantOps {
    int bm, bn, bk;          // shape 0
    int bm1, bn1;            // shape 1 (mixed plans; equal to shape 0 otherwise)
    const void *kern, *kern_conv;
    int nt;
};

template <class CF, class CS = CF>
VariantOps ops_of() {
    return VariantOps{CF::BM, CF::BN, CF::BK, CS::BM, CS::BN, (const void *)grouped_fp32_kernel<CF, false, CS>,
                      (const void *)grouped_fp32_kernel<CF, true, CS>, CF::NT};
}

VariantOps variant_ops(int mode, int v) {
    if (mode == UGEMM_MODE_FP32_MFMA) {
        switch (v) {
            case 0: return ops_of<GM0>();
            case 1: return ops_of<GM1>();
            case 2: return ops_of<GM2>();
            case 3: return ops_of<GM3>();
            default: return ops_of<GM0, GM3>();
        }
    }
    switch (v) {
        case 0: return ops_of<G0>();
        case 1: return ops_of<G1>();
        case 2: return ops_of<G2>();
        case 3: return ops_of<G3>();
        default: return ops_of<G0, G3>();
    }
}

const char *kGroupedNames[] = {"fp32_128x128x16_t8x8", "fp32_64x64x16_t4x4", "fp32_128x64x16_t8x4",
                               "fp32_64x128x16_t4x8", "fp32_mixed_128x128_or_64x128_per_problem"};
const char *kGroupedM32Names[] = {"f32_mfma_128x128x16_w2x2", "f32_mfma_64x64x16_w2x2", "f32_mfma_128x64x16_w2x2",
                                  "f32_mfma_64x128x16_w2x2", "f32_mfma_mixed_128x128_or_64x128_per_problem"};

// Tile shape of problem rows M in a mixed plan: 0 (BM0 rows) unless its
// padded rows exceed those of shape 1 (BM1 rows) by more than 10%.
int mixed_shape(int M, int bm0, int bm1) {
    const long r0 = (long)(M + bm0 - 1) / bm0 * bm0, r1 = (long)(M + bm1 - 1) / bm1 * bm1;
    return (r0 * 10 <= r1 * 11) ? 0 : 1;
}

}  // namespace

int grouped_variant_count(int mode) {
    if (mode == UGEMM_MODE_FP32 || mode == UGEMM_MODE_FP32_MFMA) return kNumVariants;
    if (mode == UGEMM_MODE_BF16X1 || mode == UGEMM_MODE_BF16X3 || mode == UGEMM_MODE_BF16X6 ||
        mode == UGEMM_MODE_FP16X3)
        return mfma_grouped_variant_count(mode);
    return 0;
}

}  // namespace ugemm

using namespace ugemm;

struct ugemm_grouped_plan {
    std::vector<DevProblem> probs;
    double flops = 0;
    int mode = UGEMM_MODE_FP32;
    int variant = 0;
    int grid = 0;
    int ntiles = 0;
    DevProblem *d_probs = nullptr;
    DevTile *d_tiles = nullptr;
    int *d_queues = nullptr;   // kQueues + 1 queue starts
    int *d_heads = nullptr;    // kQueues queue heads (zeroed before every launch)
    bool capture = false;
    bool conv = false;          // a plan of convolutions (grouped_fp32_kernel<CF, true>)
    int2 *d_ktab = nullptr;     // the convolutions' im2col offset tables, one block per problem
    // Split-form plans (BF16 forms): every problem's planes, prepared by one
    // job-table kernel, then one split-core block per tile.
    bool split = false;
    std::vector<char> ta, tb;   // layouts (the FP32 DevProblem keeps them in `layout`)
    void *d_planes = nullptr;
    PrepJob *d_jobs = nullptr;
    GroupedSplitProblem *d_sprobs = nullptr;
    GroupedSplitTile *d_stiles = nullptr;
    int njobs = 0, prep_blocks = 0;
    int *d_block_job = nullptr;    // preparation block -> job
    int *d_ablock_job = nullptr;   // row-maximum block -> job (FP16X3)
    int b_kcont = 0;            // variants that split B in the core: every B is k-contiguous ('T')
    // FP16X3: every operand row's maximum (the row scales), and its jobs.
    unsigned *d_amax = nullptr;
    PrepJob *d_ajobs = nullptr;
    long amax_count = 0, amax_blocks = 0;
    int najobs = 0;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void release_schedule(ugemm_grouped_plan *p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
    (void)hipFree(p->d_tiles);
    (void)hipFree(p->d_queues);
    (void)hipFree(p->d_heads);
    (void)hipFree(p->d_planes);
    (void)hipFree(p->d_jobs);
    (void)hipFree(p->d_sprobs);
    (void)hipFree(p->d_stiles);
    (void)hipFree(p->d_amax);
    (void)hipFree(p->d_ajobs);
    (void)hipFree(p->d_block_job);
    (void)hipFree(p->d_ablock_job);
    p->d_amax = nullptr;
    p->d_ajobs = nullptr;
    p->d_block_job = nullptr;
    p->d_ablock_job = nullptr;
    p->d_tiles = nullptr;
    p->d_queues = nullptr;
    p->d_heads = nullptr;
    p->d_planes = nullptr;
    p->d_jobs = nullptr;
    p->d_sprobs = nullptr;
    p->d_stiles = nullptr;
}

int launch_plan_kernel(ugemm_grouped_plan *p, hipStream_t s) {
    if (p->grid == 0) return UGEMM_OK;
    if (p->split) {
        const GroupedScales sc{p->d_amax, p->amax_count, p->d_ajobs, p->d_ablock_job, p->amax_blocks};
        return launch_mfma_grouped(p->mode, p->variant, p->b_kcont, p->d_jobs, p->d_block_job, p->prep_blocks,
                                   p->d_sprobs, p->d_stiles, p->ntiles, s, p->d_amax ? &sc : nullptr);
    }
    const VariantOps ops = variant_ops(p->mode, p->variant);
    hipLaunchKernelGGL(grouped_reset_heads, dim3(1), dim3(64), 0, s, p->d_heads);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "grouped queue reset");
    void *args[] = {&p->d_probs, &p->d_tiles, &p->d_queues, &p->d_heads};
    timing_begin(s);
    e = hipLaunchKernel(p->conv ? ops.kern_conv : ops.kern, dim3(p->grid), dim3(ops.nt), args, 0, s);
    timing_end(s);
    if (e != hipSuccess) return hip_fail(e, "grouped launch");
    return UGEMM_OK;
}

int capture_plan(ugemm_grouped_plan *p);   // records the plan's launches in its graph

// Builds the tile table and the per-XCD tile queues for variant v,
// uploads it, and (if requested) captures the launch in a graph.
int build_schedule(ugemm_grouped_plan *p, int v) {
    release_schedule(p);
    p->variant = v;
    const VariantOps ops = variant_ops(p->mode, v);
    struct HostTile {
        DevTile t;
        long cost;
    };
    std::vector<HostTile> tiles;
    for (int i = 0; i < (int)p->probs.size(); ++i) {
        const DevProblem &q = p->probs[i];
        if (q.M == 0 || q.N == 0) continue;
        if (q.K == 0 && q.beta == 1.0f) continue;   // nothing to do
        // Shape of this problem's tiles; cost ~ k-steps x tile rows.
        const int shape = (v == kMixed) ? mixed_shape(q.M, ops.bm, ops.bm1) : 0;
        const int bm = shape ? ops.bm1 : ops.bm, bn = shape ? ops.bn1 : ops.bn;
        const long cost = ((q.K + ops.bk - 1) / ops.bk + 2) * bm;
        // Grouped order: GROUP_M tile-rows at a time, walked column by column.
        constexpr int GROUP_M = 8;
        const int tiles_m = (q.M + bm - 1) / bm, tiles_n = (q.N + bn - 1) / bn;
        for (int g0 = 0; g0 < tiles_m; g0 += GROUP_M) {
            const int g1 = std::min(tiles_m, g0 + GROUP_M);
            for (int tn = 0; tn < tiles_n; ++tn)
                for (int tm = g0; tm < g1; ++tm)
                    tiles.push_back(HostTile{DevTile{i, tm * bm, tn * bn, shape}, cost});
        }
    }
    std::stable_sort(tiles.begin(), tiles.end(), [](const HostTile &a, const HostTile &b) { return a.cost > b.cost; });
    int per_cu = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, p->conv ? ops.kern_conv : ops.kern, ops.nt, 0) !=
            hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    const long slots = (long)device_cu_count() * per_cu;
    const int G = (int)std::min<long>((long)tiles.size(), slots);
    p->grid = G;
    p->ntiles = (int)tiles.size();
    if (G == 0) return UGEMM_OK;
    // Chunks of W consecutive (sorted) tiles, chunk c to queue c % kQueues.
    const long ntiles = (long)tiles.size();
    const long W = std::max<long>(1, G / kQueues);
    std::vector<DevTile> flat;
    std::vector<int> qbegin(kQueues + 1, 0);
    flat.reserve(tiles.size());
    for (int q = 0; q < kQueues; ++q) {
        qbegin[q] = (int)flat.size();
        for (long c0 = q * W; c0 < ntiles; c0 += kQueues * W)
            for (long j = c0; j < std::min(ntiles, c0 + W); ++j) flat.push_back(tiles[j].t);
    }
    qbegin[kQueues] = (int)flat.size();
    UGEMM_HIP_TRY(hipMalloc(&p->d_tiles, flat.size() * sizeof(DevTile)));
    UGEMM_HIP_TRY(hipMalloc(&p->d_queues, qbegin.size() * sizeof(int)));
    UGEMM_HIP_TRY(hipMalloc(&p->d_heads, kQueues * sizeof(int)));
    UGEMM_HIP_TRY(hipMemcpy(p->d_tiles, flat.data(), flat.size() * sizeof(DevTile), hipMemcpyHostToDevice));
    UGEMM_HIP_TRY(hipMemcpy(p->d_queues, qbegin.data(), qbegin.size() * sizeof(int), hipMemcpyHostToDevice));
    if (p->capture) return capture_plan(p);
    return UGEMM_OK;
}

int capture_plan(ugemm_grouped_plan *p) {
    UGEMM_HIP_TRY(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_plan_kernel(p, p->cap_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(p->cap_stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    p->graph = g;
    UGEMM_HIP_TRY(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    return UGEMM_OK;
}

// Split-form plan for variant v of the form's split-core table: planes
// of every problem (K padded to 64; a problem with K == 0 or alpha == 0 gets
// zero planes, so the core computes beta * C), the preparation jobs, and the
// tiles in grouped order sorted by K, longest first (the hardware dispatches
// blocks in order, which balances them as they finish).
// The job of every block of a job grid (jobs in order of their first block).
int upload_block_jobs(const std::vector<PrepJob> &jobs, long nblocks, int **d_map) {
    *d_map = nullptr;
    if (nblocks <= 0) return UGEMM_OK;
    std::vector<int> map((size_t)nblocks);
    for (size_t i = 0; i < jobs.size(); ++i) {
        const long b1 = i + 1 < jobs.size() ? jobs[i + 1].block0 : nblocks;
        for (long b = jobs[i].block0; b < b1; ++b) map[(size_t)b] = (int)i;
    }
    UGEMM_HIP_TRY(hipMalloc(d_map, map.size() * sizeof(int)));
    UGEMM_HIP_TRY(hipMemcpy(*d_map, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice));
    return UGEMM_OK;
}

int build_split_schedule(ugemm_grouped_plan *p, int v) {
    release_schedule(p);
    p->variant = v;
    int bm = 0, bn = 0;
    int rc = mfma_grouped_tile(p->mode, v, &bm, &bn);
    if (rc) return rc;
    const int np = p->mode == UGEMM_MODE_BF16X1 ? 1 : (p->mode == UGEMM_MODE_BF16X6 ? 3 : 2);
    const bool scaled = p->mode == UGEMM_MODE_FP16X3;
    // Variants that split B in the core read every B in one layout.
    const bool split_b = mfma_grouped_splits_b(p->mode, v) != 0;
    p->b_kcont = 0;
    if (split_b) {
        int n_t = 0, n_n = 0;
        for (const DevProblem &q : p->probs) {
            if (q.M == 0 || q.N == 0) continue;
            ((q.layout & 2) ? n_t : n_n)++;
        }
        if (n_t && n_n)
            return fail(UGEMM_ERR_UNSUPPORTED, "grouped variant %d splits B in the core: B layouts differ", v);
        p->b_kcont = n_t ? 1 : 0;
    }
    std::vector<GroupedSplitProblem> sp;
    std::vector<PrepJob> jobs, ajobs;
    long acount = 0, ablocks = 0;   // row maxima, row-maximum blocks (FP16X3)
    // Row-maximum job of an operand (FP16X3): its maxima at element offset
    // `at` of the maxima buffer (rebased below).
    auto add_amax_job = [&](const float *X, long ld, bool kcont, int rows, int K, long at) {
        PrepJob j{};
        j.X = X;
        j.ld = ld;
        j.out = (void *)(intptr_t)at;
        j.rows = rows;
        j.K = K;
        j.kcont = kcont ? 1 : 0;
        j.vec = (K > 0 && ugemm_aligned16(X) && ld % 4 == 0) ? 1 : 0;
        j.block0 = (int)ablocks;
        j.nbx = kcont ? (K + 2047) / 2048 : (K + kAmaxRcontSpan - 1) / kAmaxRcontSpan;
        ablocks += kcont ? (long)rows * j.nbx : (long)((rows + 255) / 256) * j.nbx;
        if (K > 0) ajobs.push_back(j);
        else ablocks = j.block0;   // no rows to scan
    };
    struct HostTile {
        GroupedSplitTile t;
        long cost;
    };
    std::vector<HostTile> tiles;
    std::vector<long> a_off, b_off;
    long total = 0, blocks = 0;   // plane elements, preparation blocks
    auto add_job = [&](const float *X, long ld, bool kcont, int rows, int K, int rows_pad, int k_pad, long off,
                       long plane, long amax_at, bool row_scaled = false) {
        PrepJob j{};
        j.amax = scaled ? (const unsigned *)(intptr_t)amax_at : nullptr;
        j.X = X;
        j.ld = ld;
        j.out = (void *)(intptr_t)off;   // element offset; rebased on the planes' address below
        j.plane_stride = plane;
        j.rows = rows;
        j.K = K;
        j.rows_pad = rows_pad;
        j.k_pad = k_pad;
        j.kcont = row_scaled ? kPrepRowScaled : (kcont ? 1 : 0);
        j.vec = (K > 0 && ugemm_aligned16(X) && ld % 4 == 0) ? 1 : 0;
        j.block0 = (int)blocks;
        j.nbx = rows_pad / 64;
        const long ktiles = kcont ? (k_pad + 127) / 128 : k_pad / 64;
        blocks += row_scaled ? rows_pad / 2 : (long)(rows_pad / 64) * ((ktiles + kPrepSpan - 1) / kPrepSpan);
        jobs.push_back(j);
    };
    for (int i = 0; i < (int)p->probs.size(); ++i) {
        const DevProblem &q = p->probs[i];
        if (q.M == 0 || q.N == 0) continue;
        if (q.K == 0 && q.beta == 1.0f) continue;   // nothing to do
        const long Mp = (q.M + bm - 1) / bm * (long)bm, Np = (q.N + bn - 1) / bn * (long)bn;
        const long Kp = std::max<long>(64, (q.K + 63) / 64 * 64);
        const long a_plane = Mp * Kp, b_plane = Np * Kp;
        const long ao = total, bo = total + np * a_plane;
        total = bo + (split_b ? 0 : np * b_plane);
        const bool ta = (q.layout & 1) != 0, tb = (q.layout & 2) != 0;
        const long am_a = acount, am_b = acount + Mp;
        // FP16X3 with a k-contiguous A of at most 8192 columns: A's maxima come
        // with its split (two rows per block), not from the maximum pass.
        const bool a_rows = scaled && !ta && Kp <= 8192 && Mp % 2 == 0;
        if (scaled) {
            acount += Mp + Np;
            if (!a_rows) add_amax_job(q.A, q.lda, !ta, q.M, q.K, am_a);
            add_amax_job(q.B, q.ldb, tb, q.N, q.K, am_b);
        }
        add_job(q.A, q.lda, !ta, q.M, q.K, (int)Mp, (int)Kp, ao, a_plane, am_a, a_rows);
        if (!split_b) add_job(q.B, q.ldb, tb, q.N, q.K, (int)Np, (int)Kp, bo, b_plane, am_b);
        GroupedSplitProblem g{};
        if (split_b) {
            // FP32 B read by the core; K == 0 (alpha == 0) reads nothing.
            g.Bf = q.B;
            g.ldb = q.ldb;
            g.K = q.K;
            g.vec_b = tb ? (q.K > 0 && ugemm_aligned16(q.B) && q.ldb % 4 == 0)
                         : ((reinterpret_cast<uintptr_t>(q.B) & 7) == 0 && q.ldb % 2 == 0);
        }
        g.amax_a = scaled ? (const unsigned *)(intptr_t)am_a : nullptr;
        g.amax_b = scaled ? (const unsigned *)(intptr_t)am_b : nullptr;
        g.Ap = (const void *)(intptr_t)ao;
        g.Bp = (const void *)(intptr_t)bo;
        g.a_plane = a_plane;
        g.b_plane = b_plane;
        g.M = q.M;
        g.N = q.N;
        g.Kp = (int)Kp;
        g.Mp = (int)Mp;
        g.Np = (int)Np;
        g.alpha = q.alpha;
        g.beta = q.beta;
        g.C = q.C;
        g.ldc = q.ldc;
        const int idx = (int)sp.size();
        sp.push_back(g);
        constexpr int GROUP_M = 8;
        const int tiles_m = (int)(Mp / bm), tiles_n = (int)(Np / bn);
        for (int g0 = 0; g0 < tiles_m; g0 += GROUP_M) {
            const int g1 = std::min(tiles_m, g0 + GROUP_M);
            for (int tn = 0; tn < tiles_n; ++tn)
                for (int tm = g0; tm < g1; ++tm) tiles.push_back(HostTile{GroupedSplitTile{idx, tm * bm, tn * bn, 0}, Kp});
        }
    }
    if (blocks > 0x7fffffffL || ablocks > 0x7fffffffL || (long)tiles.size() > 0x7fffffffL)
        return fail(UGEMM_ERR_ARG, "grouped: plan too large");
    std::stable_sort(tiles.begin(), tiles.end(), [](const HostTile &a, const HostTile &b) { return a.cost > b.cost; });
    p->grid = tiles.empty() ? 0 : 1;
    p->ntiles = (int)tiles.size();
    p->njobs = (int)jobs.size();
    p->prep_blocks = (int)blocks;
    if (tiles.empty()) return UGEMM_OK;
    // Planes (16-bit elements), then rebase the offsets on their address.
    UGEMM_HIP_TRY(hipMalloc(&p->d_planes, (size_t)total * 2));
    char *base = (char *)p->d_planes;
    for (PrepJob &j : jobs) j.out = base + (intptr_t)j.out * 2;
    for (GroupedSplitProblem &g : sp) {
        g.Ap = base + (intptr_t)g.Ap * 2;
        g.Bp = split_b ? nullptr : base + (intptr_t)g.Bp * 2;
    }
    if (scaled) {
        // Row maxima: rebase the offsets on the buffer's address.
        UGEMM_HIP_TRY(hipMalloc(&p->d_amax, (size_t)std::max<long>(1, acount) * sizeof(unsigned)));
        unsigned *am = p->d_amax;
        for (PrepJob &j : jobs) j.amax = am + (intptr_t)j.amax;
        for (PrepJob &j : ajobs) j.out = am + (intptr_t)j.out;
        for (GroupedSplitProblem &g : sp) {
            g.amax_a = am + (intptr_t)g.amax_a;
            g.amax_b = am + (intptr_t)g.amax_b;
        }
        p->amax_count = acount;
        p->amax_blocks = ablocks;
        p->najobs = (int)ajobs.size();
        if (!ajobs.empty()) {
            UGEMM_HIP_TRY(hipMalloc(&p->d_ajobs, ajobs.size() * sizeof(PrepJob)));
            UGEMM_HIP_TRY(hipMemcpy(p->d_ajobs, ajobs.data(), ajobs.size() * sizeof(PrepJob), hipMemcpyHostToDevice));
            rc = upload_block_jobs(ajobs, ablocks, &p->d_ablock_job);
            if (rc) return rc;
        }
    }
    rc = upload_block_jobs(jobs, blocks, &p->d_block_job);
    if (rc) return rc;
    std::vector<GroupedSplitTile> flat(tiles.size());
    for (size_t i = 0; i < tiles.size(); ++i) flat[i] = tiles[i].t;
    UGEMM_HIP_TRY(hipMalloc(&p->d_jobs, jobs.size() * sizeof(PrepJob)));
    UGEMM_HIP_TRY(hipMalloc(&p->d_sprobs, sp.size() * sizeof(GroupedSplitProblem)));
    UGEMM_HIP_TRY(hipMalloc(&p->d_stiles, flat.size() * sizeof(GroupedSplitTile)));
    UGEMM_HIP_TRY(hipMemcpy(p->d_jobs, jobs.data(), jobs.size() * sizeof(PrepJob), hipMemcpyHostToDevice));
    UGEMM_HIP_TRY(hipMemcpy(p->d_sprobs, sp.data(), sp.size() * sizeof(GroupedSplitProblem), hipMemcpyHostToDevice));
    UGEMM_HIP_TRY(hipMemcpy(p->d_stiles, flat.data(), flat.size() * sizeof(GroupedSplitTile), hipMemcpyHostToDevice));
    if (p->capture) return capture_plan(p);
    return UGEMM_OK;
}

int default_split_variant(const ugemm_grouped_plan *p) {
    // Most of the work in problems of 128+ rows and columns: 128 x 128 tiles; otherwise 64 x 64.
    // With 128 x 128 tiles, every B in one layout and few tile rows per
    // problem (B split again by each, on average at most 4), the 8-wave
    // streamed variant that splits B in the core: B's planes would cost more
    // than re-splitting (512 im2col-shaped problems, M = 32..512,
    // N = 49..3136, K = 144..4608: 4.1 ms, against 4.7 for the 4-wave
    // split-B variant and 5.8 with planes, for BF16x3).
    double big = 0, all = 0, rows_w = 0;
    int n_t = 0, n_n = 0;
    for (const DevProblem &q : p->probs) {
        const double f = 2.0 * q.M * q.N * q.K;
        all += f;
        if (q.M >= 128 && q.N >= 128) big += f;
        rows_w += f * ((q.M + 127) / 128);
        if (q.M > 0 && q.N > 0) ((q.layout & 2) ? n_t : n_n)++;
    }
    if (all == 0 || big < 0.5 * all) return 3;
    const int split_b = mfma_variant_count(p->mode) + 2;   // split-B on the 8-wave streamed 128 x 128
    if (!(n_t && n_n) && rows_w <= 4.0 * all && mfma_grouped_splits_b(p->mode, split_b)) return split_b;
    return 1;
}

bool parse_t(char c, bool *t) {
    if (c == 'N' || c == 'n') { *t = false; return true; }
    if (c == 'T' || c == 't' || c == 'C' || c == 'c') { *t = true; return true; }
    return false;
}

int default_variant(const ugemm_grouped_plan *p) {
    // Most of the work in large-N problems: 128 x 128 tiles; otherwise 64 x 64.
    double big = 0, all = 0;
    for (const DevProblem &q : p->probs) {
        const double f = 2.0 * q.M * q.N * q.K;
        all += f;
        if (q.M >= 128 && q.N >= 128) big += f;
    }
    return (all == 0 || big >= 0.5 * all) ? 0 : 1;
}

// Uploads the problem table, picks the tile variant (forced, tuned, or by
// default), builds the queues and, if asked, captures the launch.  Takes
// ownership of p: on failure it is destroyed.
int finish_plan(ugemm_grouped_plan *p, int count, int capture, ugemm_grouped_plan_t **plan) {
    p->capture = capture != 0;
    hipError_t e;
    if ((e = hipMalloc(&p->d_probs, std::max<size_t>(1, p->probs.size()) * sizeof(DevProblem))) != hipSuccess ||
        (count > 0 && (e = hipMemcpy(p->d_probs, p->probs.data(), p->probs.size() * sizeof(DevProblem),
                                     hipMemcpyHostToDevice)) != hipSuccess)) {
        ugemm_hip_grouped_plan_destroy(p);
        return hip_fail(e, "grouped: descriptor upload");
    }
    if (p->capture) {
        // Capture needs a non-null stream; an owned one is used for capture
        // only (replays go to the stream passed to ugemm_hip_grouped_run).
        if ((e = hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking)) != hipSuccess) {
            ugemm_hip_grouped_plan_destroy(p);
            return hip_fail(e, "grouped: capture stream");
        }
    }
    // A forced single-GEMM variant applies when the grouped kernel has it too
    // (split forms: every split-core variant).
    const int nv = grouped_variant_count(p->mode);
    int v = forced_variant(p->mode);
    if (v >= nv) v = -1;
    if (v < 0)
        v = tuned_variant(TUNE_KEY_GROUPED, p->conv, p->mode, count, (int)std::min(p->flops / 1e6, 2e9), 0);
    if (v < 0 || v >= nv) v = p->split ? default_split_variant(p) : default_variant(p);
    const int rc = p->split ? build_split_schedule(p, v) : build_schedule(p, v);
    if (rc) {
        ugemm_hip_grouped_plan_destroy(p);
        return rc;
    }
    *plan = p;
    return UGEMM_OK;
}

}  // namespace

extern "C" {

int ugemm_hip_grouped_plan_create(ugemm_grouped_plan_t **plan, int count, const ugemm_problem_t *problems, int mode,
                                  int capture, void *stream) {
    if (!plan) return fail(UGEMM_ERR_ARG, "grouped: NULL plan pointer");
    *plan = nullptr;
    if (count < 0 || (count > 0 && !problems)) return fail(UGEMM_ERR_ARG, "grouped: bad problem list");
    const bool split = mode == UGEMM_MODE_BF16X1 || mode == UGEMM_MODE_BF16X3 || mode == UGEMM_MODE_BF16X6 ||
                       mode == UGEMM_MODE_FP16X3;
    if (mode != UGEMM_MODE_FP32 && mode != UGEMM_MODE_FP32_MFMA && !split)
        return fail(UGEMM_ERR_UNSUPPORTED, "grouped: mode %d not provided (FP32, FP32-MFMA and the split forms)",
                    mode);
    int rc = ensure_device();
    if (rc) return rc;
    ugemm_grouped_plan *p = new ugemm_grouped_plan();
    p->mode = mode;
    p->split = split;
    for (int i = 0; i < count; ++i) {
        const ugemm_problem_t &q = problems[i];
        bool ta, tb;
        if (!parse_t(q.transa, &ta) || !parse_t(q.transb, &tb)) {
            delete p;
            return fail(UGEMM_ERR_ARG, "grouped: problem %d: bad trans flag", i);
        }
        if (q.M < 0 || q.N < 0 || q.K < 0) {
            delete p;
            return fail(UGEMM_ERR_ARG, "grouped: problem %d: negative size", i);
        }
        const int la = ta ? q.M : q.K, lb = tb ? q.K : q.N;
        if (q.lda < std::max(la, 1) || q.ldb < std::max(lb, 1) || q.ldc < std::max(q.N, 1)) {
            delete p;
            return fail(UGEMM_ERR_ARG, "grouped: problem %d: leading dimension too small", i);
        }
        const bool reads_ab = q.alpha != 0.0f && q.K > 0;
        if (q.M > 0 && q.N > 0 && (!q.C || (reads_ab && (!q.A || !q.B)))) {
            delete p;
            return fail(UGEMM_ERR_ARG, "grouped: problem %d: NULL operand", i);
        }
        const bool vec = ugemm_aligned16(q.A) && ugemm_aligned16(q.B) && q.lda % 4 == 0 && q.ldb % 4 == 0;
        DevProblem d{q.A, q.B, q.C, q.lda, q.ldb, q.ldc, q.M, q.N, reads_ab ? q.K : 0, q.alpha, q.beta,
                     (ta ? 1 : 0) | (tb ? 2 : 0) | (vec ? 4 : 0)};
        p->probs.push_back(d);
        p->flops += 2.0 * q.M * q.N * q.K;
    }
    return finish_plan(p, count, capture, plan);
}

int ugemm_hip_conv_grouped_plan_create(ugemm_grouped_plan_t **plan, int count, const ugemm_conv_problem_t *convs,
                                       int capture, void *stream) {
    return ugemm_hip_conv_grouped_plan_create_mode(plan, count, convs, UGEMM_MODE_FP32, capture, stream);
}

int ugemm_hip_conv_grouped_plan_create_mode(ugemm_grouped_plan_t **plan, int count, const ugemm_conv_problem_t *convs,
                                            int mode, int capture, void *stream) {
    if (!plan) return fail(UGEMM_ERR_ARG, "grouped conv: NULL plan pointer");
    *plan = nullptr;
    if (mode != UGEMM_MODE_FP32 && mode != UGEMM_MODE_FP32_MFMA)
        return fail(UGEMM_ERR_UNSUPPORTED, "grouped conv: mode %d not provided (FP32 and FP32-MFMA)", mode);
    if (count < 0 || (count > 0 && !convs)) return fail(UGEMM_ERR_ARG, "grouped conv: bad problem list");
    int rc = ensure_device();
    if (rc) return rc;
    std::vector<int> hcol(count), wcol(count);
    long ktab_total = 0;
    for (int i = 0; i < count; ++i) {
        const ugemm_conv_problem_t &c = convs[i];
        rc = conv_shape(c.ich, c.w, c.h, c.k, c.pad, c.stride, c.ch, &hcol[i], &wcol[i]);
        if (rc) return rc;
        if (!c.in || !c.weights || !c.out) return fail(UGEMM_ERR_ARG, "grouped conv: problem %d: NULL buffer", i);
        if (c.act != UGEMM_ACT_NONE && c.act != UGEMM_ACT_LEAKY_RELU)
            return fail(UGEMM_ERR_ARG, "grouped conv: problem %d: bad activation %d", i, c.act);
        ktab_total += (long)c.ich * c.k * c.k;
    }
    // im2col offset tables: k = (ch*kh + i)*kw + j -> {ch*H*W + i*W + j, (i << 16) | j}.
    std::vector<int2> ktab((size_t)std::max<long>(1, ktab_total));
    std::vector<long> kofs(count);
    long at = 0;
    for (int i = 0; i < count; ++i) {
        const ugemm_conv_problem_t &c = convs[i];
        kofs[i] = at;
        for (int q = 0; q < c.ich * c.k * c.k; ++q) {
            const int ci = q / (c.k * c.k), r = q - ci * c.k * c.k, ki = r / c.k, kj = r - ki * c.k;
            ktab[at++] = make_int2(ci * c.h * c.w + ki * c.w + kj, (ki << 16) | kj);
        }
    }
    ugemm_grouped_plan *p = new ugemm_grouped_plan();
    p->mode = mode;
    p->conv = true;
    hipError_t e;
    if ((e = hipMalloc(&p->d_ktab, ktab.size() * sizeof(int2))) != hipSuccess ||
        (e = hipMemcpy(p->d_ktab, ktab.data(), ktab.size() * sizeof(int2), hipMemcpyHostToDevice)) != hipSuccess) {
        ugemm_hip_grouped_plan_destroy(p);
        return hip_fail(e, "grouped conv: offset tables");
    }
    for (int i = 0; i < count; ++i) {
        const ugemm_conv_problem_t &c = convs[i];
        const int K = c.ich * c.k * c.k, N = hcol[i] * wcol[i];
        DevProblem d{};
        d.A = c.weights;
        d.lda = K;
        d.C = c.out;
        d.ldc = N;
        d.M = c.ch;
        d.N = N;
        d.K = K;
        d.alpha = 1.0f;
        d.beta = 0.0f;
        d.layout = 8 | ((ugemm_aligned16(c.weights) && K % 4 == 0) ? 4 : 0);
        d.epi.bias = c.bias;
        d.epi.bias_rows = 1;
        d.epi.act = c.act;
        d.epi.slope = c.slope;
        d.cd = ConvDesc{c.in, p->d_ktab + kofs[i], c.h, c.w, wcol[i], c.stride, c.pad};
        p->probs.push_back(d);
        p->flops += 2.0 * c.ch * N * K;
    }
    (void)stream;
    return finish_plan(p, count, capture, plan);
}

int ugemm_hip_grouped_run(ugemm_grouped_plan_t *p, void *stream) {
    if (!p) return fail(UGEMM_ERR_ARG, "grouped: NULL plan");
    if (p->grid == 0) return UGEMM_OK;
    if (p->exec) {
        UGEMM_HIP_TRY(hipGraphLaunch(p->exec, (hipStream_t)stream));
        return UGEMM_OK;
    }
    return launch_plan_kernel(p, (hipStream_t)stream);
}

void ugemm_hip_grouped_plan_destroy(ugemm_grouped_plan_t *p) {
    if (!p) return;
    if (p->cap_stream) (void)hipStreamSynchronize(p->cap_stream);
    release_schedule(p);
    (void)hipFree(p->d_probs);
    (void)hipFree(p->d_ktab);
    if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
    delete p;
}

int ugemm_hip_sgemm_grouped(int count, const ugemm_problem_t *problems, int mode, void *stream) {
    ugemm_grouped_plan_t *p = nullptr;
    int rc = ugemm_hip_grouped_plan_create(&p, count, problems, mode, 0, stream);
    if (rc) return rc;
    rc = ugemm_hip_grouped_run(p, stream);
    // The plan's device tables must outlive the launch.
    hipError_t e = hipStreamSynchronize((hipStream_t)stream);
    ugemm_hip_grouped_plan_destroy(p);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "grouped: synchronize");
    return UGEMM_OK;
}

int ugemm_hip_grouped_variant_count(void) { return kNumVariants; }

const char *ugemm_hip_grouped_variant_name(int v) { return (v >= 0 && v < kNumVariants) ? kGroupedNames[v] : nullptr; }

int ugemm_hip_grouped_mode_variant_count(int mode) {
    const int n = grouped_variant_count(mode);
    return n > 0 ? n : fail(UGEMM_ERR_ARG, "grouped: mode %d not provided", mode);
}

const char *ugemm_hip_grouped_mode_variant_name(int mode, int v) {
    if (v < 0 || v >= grouped_variant_count(mode)) return nullptr;
    if (mode == UGEMM_MODE_FP32) return kGroupedNames[v];
    if (mode == UGEMM_MODE_FP32_MFMA) return kGroupedM32Names[v];
    return mfma_grouped_variant_name(mode, v);
}

// Times each tile variant on the plan's own problems (C is overwritten, as by
// a run) and keeps the fastest; the choice is remembered for plans with the
// same problem count, work bucket and form.  Returns the chosen variant
// (split-form plans: the split-core variant).
int ugemm_hip_tune_grouped(ugemm_grouped_plan_t *p) {
    if (!p) return fail(UGEMM_ERR_ARG, "grouped: NULL plan");
    if (p->grid == 0) return p->variant;
    // The caller's pending work on any stream (uploads of A, B, C) completes
    // before the timing runs read the operands.
    UGEMM_HIP_TRY(hipDeviceSynchronize());
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    UGEMM_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    int best = -1, rc = UGEMM_OK;
    float best_ms = 0;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = fail(UGEMM_ERR_HIP, "events");
    // Every timing run overwrites each problem's C: keep a copy of every C
    // (packed, M x N) and put it back afterwards, so that tuning leaves C as
    // the caller left it.
    std::vector<size_t> c_off(p->probs.size());
    size_t c_bytes = 0;
    for (size_t i = 0; i < p->probs.size(); ++i) {
        c_off[i] = c_bytes;
        c_bytes += ((size_t)p->probs[i].M * p->probs[i].N * sizeof(float) + 255) / 256 * 256;
    }
    char *c_save = nullptr;
    if (!rc && c_bytes) {
        const hipError_t e = hipMalloc((void **)&c_save, c_bytes);
        if (e != hipSuccess) rc = hip_fail(e, "grouped tune: copy of C");
    }
    auto copy_c = [&](bool save) {
        for (size_t i = 0; i < p->probs.size() && !rc; ++i) {
            const DevProblem &q = p->probs[i];
            if (q.M == 0 || q.N == 0) continue;
            const size_t row = (size_t)q.N * sizeof(float), ld = (size_t)q.ldc * sizeof(float);
            const hipError_t e = save ? hipMemcpy2DAsync(c_save + c_off[i], row, q.C, ld, row, q.M,
                                                         hipMemcpyDeviceToDevice, s)
                                      : hipMemcpy2DAsync(q.C, ld, c_save + c_off[i], row, row, q.M,
                                                         hipMemcpyDeviceToDevice, s);
            if (e != hipSuccess) rc = hip_fail(e, "grouped tune: copy of C");
        }
    };
    if (c_save) copy_c(true);
    const bool cap = p->capture;
    p->capture = false;
    const int nv = grouped_variant_count(p->mode);
    for (int v = 0; v < nv && !rc; ++v) {
        rc = p->split ? build_split_schedule(p, v) : build_schedule(p, v);
        if (rc == UGEMM_ERR_UNSUPPORTED) {   // a variant this plan's layouts do not allow
            rc = UGEMM_OK;
            continue;
        }
        if (rc) break;
        launch_plan_kernel(p, s);   // warm-up
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < 3; ++r) launch_plan_kernel(p, s);
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess) {
            rc = fail(UGEMM_ERR_HIP, "grouped tune: timing");
            break;
        }
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (best < 0 || ms < best_ms) {
            best = v;
            best_ms = ms;
        }
    }
    p->capture = cap;
    if (c_save) {
        const int rc_run = rc;
        rc = UGEMM_OK;
        copy_c(false);
        if (rc_run) rc = rc_run;
    }
    if (!rc) rc = p->split ? build_split_schedule(p, best) : build_schedule(p, best);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    if (c_save) (void)hipFree(c_save);
    if (rc) return rc;
    tune_record(TUNE_KEY_GROUPED, p->conv, p->mode, (int)p->probs.size(),
                (int)std::min(p->flops / 1e6, 2e9), 0, best, best_ms / 3);
    return best;
}

}  // extern "C"
