// Grouped SGEMM: hundreds of independent, uneven row-major problems in one
// persistent launch (optionally replayed as a hipGraph).
//
// Plan (host, once): every problem is cut into BM x BN tiles of one kernel
// variant; each tile costs ~ceil(K/BK) + 2 k-steps.  Tiles are sorted by cost
// and assigned longest-first to the least-loaded of G workgroups (G = CUs x
// resident blocks per CU), giving each workgroup a static list.  Problem
// descriptors, the tile table and the per-workgroup ranges are uploaded once;
// the launch reads nothing else, so the captured graph replays identically
// (no counters to reset, no atomics, deterministic results).
// Kernel: workgroup b walks its tile range and runs the plain FP32 tile body
// (fp32_tile.h) on each tile, dispatching on the tile's layout.
// Used for convolution-as-GEMM batches (the conv -> GEMM shape mapping of
// sgemm_ocl1.h:271-300: M = output channels, N = output pixels, K = k*k*Cin).
#include "internal.h"
#include "fp32_tile.h"

#include <algorithm>
#include <functional>
#include <queue>
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
    int layout;           // bit 0: A transposed, bit 1: B transposed
};

struct DevTile {
    int problem, m0, n0, pad;
};

template <class CF, int VEC>
__global__ __launch_bounds__(CF::NT) void grouped_fp32_kernel(const DevProblem *__restrict__ probs,
                                                              const DevTile *__restrict__ tiles,
                                                              const int *__restrict__ wg_begin) {
    __shared__ __attribute__((aligned(16))) float lds[2 * CF::BK * (CF::LDA + CF::LDB)];
    const int t0 = wg_begin[blockIdx.x], t1 = wg_begin[blockIdx.x + 1];
    for (int i = t0; i < t1; ++i) {
        const DevTile tl = tiles[i];
        const DevProblem p = probs[tl.problem];
        switch (p.layout) {
            case 0:
                fp32_tile<CF, false, false, VEC>(p.M, p.N, p.K, p.alpha, p.A, p.lda, p.B, p.ldb, p.beta, p.C, p.ldc,
                                                 tl.m0, tl.n0, lds);
                break;
            case 1:
                fp32_tile<CF, true, false, VEC>(p.M, p.N, p.K, p.alpha, p.A, p.lda, p.B, p.ldb, p.beta, p.C, p.ldc,
                                                tl.m0, tl.n0, lds);
                break;
            case 2:
                fp32_tile<CF, false, true, VEC>(p.M, p.N, p.K, p.alpha, p.A, p.lda, p.B, p.ldb, p.beta, p.C, p.ldc,
                                                tl.m0, tl.n0, lds);
                break;
            default:
                fp32_tile<CF, true, true, VEC>(p.M, p.N, p.K, p.alpha, p.A, p.lda, p.B, p.ldb, p.beta, p.C, p.ldc,
                                               tl.m0, tl.n0, lds);
                break;
        }
        __syncthreads();   // the next tile's first LDS stage overwrites this one's
    }
}

using G0 = Fp32Cfg<128, 128, 16, 8, 8>;
using G1 = Fp32Cfg<64, 64, 16, 4, 4>;
using G2 = Fp32Cfg<128, 64, 16, 8, 4>;
using G3 = Fp32Cfg<64, 128, 16, 4, 8>;
constexpr int kNumVariants = 4;

struct VariantOps {
    int bm, bn, bk;
    const void *kern_vec4, *kern_vec1;
    int nt;
};

template <class CF>
VariantOps ops_of() {
    return VariantOps{CF::BM, CF::BN, CF::BK, (const void *)grouped_fp32_kernel<CF, 4>,
                      (const void *)grouped_fp32_kernel<CF, 1>, CF::NT};
}

VariantOps variant_ops(int v) {
    switch (v) {
        case 0: return ops_of<G0>();
        case 1: return ops_of<G1>();
        case 2: return ops_of<G2>();
        default: return ops_of<G3>();
    }
}

}  // namespace
}  // namespace ugemm

using namespace ugemm;

struct ugemm_grouped_plan {
    std::vector<DevProblem> probs;
    double flops = 0;
    int mode = UGEMM_MODE_FP32;
    int variant = 0;
    int vec = 4;
    int grid = 0;
    int ntiles = 0;
    DevProblem *d_probs = nullptr;
    DevTile *d_tiles = nullptr;
    int *d_wg = nullptr;
    bool capture = false;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void release_schedule(ugemm_grouped_plan *p) {
    if (p->exec) hipGraphExecDestroy(p->exec);
    if (p->graph) hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
    hipFree(p->d_tiles);
    hipFree(p->d_wg);
    p->d_tiles = nullptr;
    p->d_wg = nullptr;
}

int launch_plan_kernel(ugemm_grouped_plan *p, hipStream_t s) {
    if (p->grid == 0) return UGEMM_OK;
    const VariantOps ops = variant_ops(p->variant);
    void *args[] = {&p->d_probs, &p->d_tiles, &p->d_wg};
    timing_begin(s);
    hipError_t e = hipLaunchKernel(p->vec == 4 ? ops.kern_vec4 : ops.kern_vec1, dim3(p->grid), dim3(ops.nt), args, 0, s);
    timing_end(s);
    if (e != hipSuccess) return hip_fail(e, "grouped launch");
    return UGEMM_OK;
}

// Builds the tile table and the static longest-first schedule for variant v,
// uploads it, and (if requested) captures the launch in a graph.
int build_schedule(ugemm_grouped_plan *p, int v) {
    release_schedule(p);
    p->variant = v;
    const VariantOps ops = variant_ops(v);
    struct HostTile {
        DevTile t;
        long cost;
    };
    std::vector<HostTile> tiles;
    for (int i = 0; i < (int)p->probs.size(); ++i) {
        const DevProblem &q = p->probs[i];
        if (q.M == 0 || q.N == 0) continue;
        if (q.K == 0 && q.beta == 1.0f) continue;   // nothing to do
        const long cost = (q.K + ops.bk - 1) / ops.bk + 2;
        for (int m0 = 0; m0 < q.M; m0 += ops.bm)
            for (int n0 = 0; n0 < q.N; n0 += ops.bn) tiles.push_back(HostTile{DevTile{i, m0, n0, 0}, cost});
    }
    std::stable_sort(tiles.begin(), tiles.end(), [](const HostTile &a, const HostTile &b) { return a.cost > b.cost; });
    int per_cu = 1;
    const void *kern = p->vec == 4 ? ops.kern_vec4 : ops.kern_vec1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, ops.nt, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    const long slots = (long)device_cu_count() * per_cu;
    const int G = (int)std::min<long>((long)tiles.size(), slots);
    p->grid = G;
    p->ntiles = (int)tiles.size();
    if (G == 0) return UGEMM_OK;
    // Greedy longest-processing-time assignment.
    std::vector<std::vector<int>> lists(G);
    typedef std::pair<long, int> Load;
    std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
    for (int g = 0; g < G; ++g) heap.push(Load(0, g));
    for (int i = 0; i < (int)tiles.size(); ++i) {
        Load l = heap.top();
        heap.pop();
        lists[l.second].push_back(i);
        heap.push(Load(l.first + tiles[i].cost, l.second));
    }
    std::vector<DevTile> flat;
    std::vector<int> begin(G + 1, 0);
    flat.reserve(tiles.size());
    for (int g = 0; g < G; ++g) {
        begin[g] = (int)flat.size();
        for (int i : lists[g]) flat.push_back(tiles[i].t);
    }
    begin[G] = (int)flat.size();
    UGEMM_HIP_TRY(hipMalloc(&p->d_tiles, flat.size() * sizeof(DevTile)));
    UGEMM_HIP_TRY(hipMalloc(&p->d_wg, begin.size() * sizeof(int)));
    UGEMM_HIP_TRY(hipMemcpy(p->d_tiles, flat.data(), flat.size() * sizeof(DevTile), hipMemcpyHostToDevice));
    UGEMM_HIP_TRY(hipMemcpy(p->d_wg, begin.data(), begin.size() * sizeof(int), hipMemcpyHostToDevice));
    if (p->capture) {
        UGEMM_HIP_TRY(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
        int rc = launch_plan_kernel(p, p->cap_stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(p->cap_stream, &g);
        if (rc) {
            if (g) hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        p->graph = g;
        UGEMM_HIP_TRY(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    }
    return UGEMM_OK;
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

}  // namespace

extern "C" {

int ugemm_hip_grouped_plan_create(ugemm_grouped_plan_t **plan, int count, const ugemm_problem_t *problems, int mode,
                                  int capture, void *stream) {
    if (!plan) return fail(UGEMM_ERR_ARG, "grouped: NULL plan pointer");
    *plan = nullptr;
    if (count < 0 || (count > 0 && !problems)) return fail(UGEMM_ERR_ARG, "grouped: bad problem list");
    if (mode != UGEMM_MODE_FP32)
        return fail(UGEMM_ERR_UNSUPPORTED, "grouped: mode %d not provided (FP32 only)", mode);
    int rc = ensure_device();
    if (rc) return rc;
    ugemm_grouped_plan *p = new ugemm_grouped_plan();
    p->mode = mode;
    bool vec = true;
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
        DevProblem d{q.A, q.B, q.C, q.lda, q.ldb, q.ldc, q.M, q.N, reads_ab ? q.K : 0, q.alpha, q.beta,
                     (ta ? 1 : 0) | (tb ? 2 : 0)};
        if (reads_ab)
            vec = vec && ugemm_aligned16(q.A) && ugemm_aligned16(q.B) && q.lda % 4 == 0 && q.ldb % 4 == 0;
        p->probs.push_back(d);
        p->flops += 2.0 * q.M * q.N * q.K;
    }
    p->vec = vec ? 4 : 1;
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
    (void)stream;
    int v = forced_variant(UGEMM_MODE_FP32);
    if (v < 0) v = tuned_variant(TUNE_KEY_GROUPED, false, false, count, (int)std::min(p->flops / 1e6, 2e9), 0);
    if (v < 0) v = default_variant(p);
    rc = build_schedule(p, v);
    if (rc) {
        ugemm_hip_grouped_plan_destroy(p);
        return rc;
    }
    *plan = p;
    return UGEMM_OK;
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
    if (p->cap_stream) hipStreamSynchronize(p->cap_stream);
    release_schedule(p);
    hipFree(p->d_probs);
    if (p->cap_stream) hipStreamDestroy(p->cap_stream);
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

// Times each tile variant on the plan's own problems (C is overwritten, as by
// a run) and keeps the fastest; the choice is remembered for plans with the
// same problem count and work bucket.  Returns the chosen variant.
int ugemm_hip_tune_grouped(ugemm_grouped_plan_t *p) {
    if (!p) return fail(UGEMM_ERR_ARG, "grouped: NULL plan");
    if (p->grid == 0) return p->variant;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    UGEMM_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    int best = -1, rc = UGEMM_OK;
    float best_ms = 0;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = fail(UGEMM_ERR_HIP, "events");
    const bool cap = p->capture;
    p->capture = false;
    for (int v = 0; v < kNumVariants && !rc; ++v) {
        rc = build_schedule(p, v);
        if (rc) break;
        launch_plan_kernel(p, s);   // warm-up
        hipEventRecord(e0, s);
        for (int r = 0; r < 3; ++r) launch_plan_kernel(p, s);
        hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess) {
            rc = fail(UGEMM_ERR_HIP, "grouped tune: timing");
            break;
        }
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        if (best < 0 || ms < best_ms) {
            best = v;
            best_ms = ms;
        }
    }
    p->capture = cap;
    if (!rc) rc = build_schedule(p, best);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipStreamSynchronize(s);
    hipStreamDestroy(s);
    if (rc) return rc;
    tune_record(TUNE_KEY_GROUPED, false, false, (int)p->probs.size(), (int)std::min(p->flops / 1e6, 2e9), 0, best,
                best_ms / 3);
    return best;
}

}  // extern "C"
