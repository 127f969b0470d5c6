// Autotuner: times every compiled kernel variant of a mode on a problem --
// and, when the tile grid is too small to fill the GPU, split-K counts -- and
// remembers the fastest per shape bucket; the table persists as a small text
// file (safe to delete: lookups then fall back to the built-in heuristics).
//
// Bucket key: (mode, transa, transb, ceil(log2 M), ceil(log2 N), ceil(log2 K));
// grouped plans use key TUNE_KEY_GROUPED with (count, MFLOP, 0) in place of
// (M, N, K).
// File format (version 3; version-1 files, which have no <splits> column,
// are read with splits = 0, the built-in choice; version 3 added a split-form
// core variant, which renumbered the grouped split-B variants, so the
// split-form entries of version-1 and -2 files are skipped):
//   ugemm-hip-tune 3 <gcnArch> <CUs>
//   <mode> <ta> <tb> <lgM> <lgN> <lgK> <variant> <ms> <splits>
// Entries from a different architecture or CU count are ignored on load.
#include "internal.h"

#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

namespace ugemm {

int run_sgemm(const Gemm<float> &g, int mode, int variant, hipStream_t s);

namespace {

struct Entry {
    int variant;
    float ms;
    int splits;   // 0: the launcher's choice
};
using Key = std::tuple<int, int, int, int, int, int>;
std::mutex g_mu;
std::map<Key, Entry> g_table;
// Forced variant per mode (slot UGEMM_MODE_COUNT: DGEMM); stored + 1 so that
// zero-initialisation means "none".
thread_local int t_forced_p1[UGEMM_MODE_COUNT + 1] = {};
thread_local int t_forced_splits[UGEMM_MODE_COUNT + 1] = {};   // 0: none

int lg(int x) {
    int b = 0;
    while ((1L << b) < (long)x) ++b;
    return b;
}
Key key(int mode, int ta, int tb, int M, int N, int K) { return Key(mode, ta, tb, lg(M), lg(N), lg(K)); }
int slot(int mode) { return mode < 0 ? UGEMM_MODE_COUNT : mode; }

__global__ void fill_uniform(float *p, long n, unsigned seed) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        unsigned x = (unsigned)i * 2654435761u ^ seed;
        x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
        p[i] = (float)(x >> 8) * (2.0f / 16777216.0f) - 1.0f;
    }
}

void arch_info(char *arch, size_t n, int *cus) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
        snprintf(arch, n, "%s", prop.gcnArchName);
        for (char *c = arch; *c; ++c)
            if (*c == ':' || *c == ' ') *c = '_';
        *cus = prop.multiProcessorCount;
    } else {
        snprintf(arch, n, "unknown");
        *cus = 0;
    }
}

}  // namespace

int tuned_variant(int mode, int ta, int tb, int M, int N, int K, int *splits) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_table.find(key(mode, ta, tb, M, N, K));
    if (splits) *splits = it == g_table.end() ? 0 : it->second.splits;
    return it == g_table.end() ? -1 : it->second.variant;
}

int forced_variant(int mode) { return t_forced_p1[slot(mode)] - 1; }
int forced_splits(int mode) { return t_forced_splits[slot(mode)]; }

void tune_record(int mode, int ta, int tb, int M, int N, int K, int variant, float ms, int splits) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_table[key(mode, ta, tb, M, N, K)] = Entry{variant, ms, splits};
}

}  // namespace ugemm

using namespace ugemm;

extern "C" {

int ugemm_hip_force_variant(int mode, int variant) {
    if (mode < -1 || mode >= UGEMM_MODE_COUNT) return fail(UGEMM_ERR_ARG, "unknown mode %d", mode);
    // Single-GEMM variants, or the (possibly longer) table of the mode's grouped plans.
    const int nv = std::max(ugemm_hip_variant_count(mode), mode >= 0 ? grouped_variant_count(mode) : 0);
    if (variant >= nv) return fail(UGEMM_ERR_ARG, "mode %d has no variant %d", mode, variant);
    t_forced_p1[slot(mode)] = variant < 0 ? 0 : variant + 1;
    return UGEMM_OK;
}

int ugemm_hip_force_splits(int mode, int splits) {
    if (mode < -1 || mode >= UGEMM_MODE_COUNT) return fail(UGEMM_ERR_ARG, "unknown mode %d", mode);
    if (splits < 0 || splits > 64) return fail(UGEMM_ERR_ARG, "splits must be in [0, 64] (got %d)", splits);
    t_forced_splits[slot(mode)] = splits;
    return UGEMM_OK;
}

int ugemm_hip_tune_lookup(char transa, char transb, int M, int N, int K, int mode) {
    const bool ta = !(transa == 'N' || transa == 'n'), tb = !(transb == 'N' || transb == 'n');
    return tuned_variant(mode, ta, tb, M, N, K);
}

int ugemm_hip_tune_sgemm(char transa, char transb, int M, int N, int K, int mode) {
    if (mode < 0 || mode >= UGEMM_MODE_COUNT) return fail(UGEMM_ERR_ARG, "unknown mode %d", mode);
    if (M <= 0 || N <= 0 || K <= 0) return fail(UGEMM_ERR_ARG, "tune: sizes must be positive");
    int rc = ensure_device();
    if (rc) return rc;
    const bool ta = !(transa == 'N' || transa == 'n'), tb = !(transb == 'N' || transb == 'n');
    const int nv = ugemm_hip_variant_count(mode);
    if (nv <= 0) return fail(UGEMM_ERR_UNSUPPORTED, "mode %d has no variants", mode);
    const long lda = ta ? M : K, ldb = tb ? K : N, ldc = N;
    const long na = (long)M * K, nb = (long)K * N, nc = (long)M * N;
    float *A = nullptr, *B = nullptr, *C = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int best = -1, best_splits = 0;
    float best_ms = 0;
    rc = UGEMM_OK;
    // Split counts worth timing: only when 128x128 tiles leave CUs idle
    // (0 = the built-in choice, always timed; FP8 does not split).
    const long tiles128 = (long)((M + 127) / 128) * ((N + 127) / 128);
    const int split_set[] = {0, 1, 2, 4, 8, 16};
    const int nsplit = (tiles128 < device_cu_count() && mode != UGEMM_MODE_FP8_BLOCK) ? 6 : 1;
    do {
        hipError_t e;
        if ((e = hipMalloc(&A, na * 4)) || (e = hipMalloc(&B, nb * 4)) || (e = hipMalloc(&C, nc * 4)) ||
            (e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) || (e = hipEventCreate(&e0)) ||
            (e = hipEventCreate(&e1))) {
            rc = hip_fail(e, "tune: set-up");
            break;
        }
        hipLaunchKernelGGL(fill_uniform, dim3(1024), dim3(256), 0, s, A, na, 1u);
        hipLaunchKernelGGL(fill_uniform, dim3(1024), dim3(256), 0, s, B, nb, 2u);
        hipLaunchKernelGGL(fill_uniform, dim3(1024), dim3(256), 0, s, C, nc, 3u);
        Gemm<float> g{M, N, K, 1.0f, 0.0f, A, lda, B, ldb, C, ldc, ta, tb};
        const double flop = 2.0 * M * N * K;
        const int reps = flop > 1e11 ? 2 : (flop > 1e9 ? 5 : 20);
        for (int v = 0; v < nv && !rc; ++v) {
            for (int si = 0; si < nsplit; ++si) {
                g.splits = split_set[si];
                if (run_sgemm(g, mode, v, s)) continue;   // variant not applicable here
                (void)hipEventRecord(e0, s);
                for (int r = 0; r < reps; ++r) run_sgemm(g, mode, v, s);
                (void)hipEventRecord(e1, s);
                if ((e = hipEventSynchronize(e1))) {
                    rc = hip_fail(e, "tune: timing");
                    break;
                }
                float ms = 0;
                (void)hipEventElapsedTime(&ms, e0, e1);
                ms /= reps;
                if (best < 0 || ms < best_ms) {
                    best = v;
                    best_splits = g.splits;
                    best_ms = ms;
                }
            }
        }
    } while (0);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    (void)hipFree(A);
    (void)hipFree(B);
    (void)hipFree(C);
    if (rc) return rc;
    if (best < 0) return fail(UGEMM_ERR_UNSUPPORTED, "tune: no variant ran");
    tune_record(mode, ta, tb, M, N, K, best, best_ms, best_splits);
    return best;
}

int ugemm_hip_tune_save(const char *path) {
    if (!path) return fail(UGEMM_ERR_ARG, "tune_save: NULL path");
    int rc = ensure_device();
    if (rc) return rc;
    FILE *f = fopen(path, "w");
    if (!f) return fail(UGEMM_ERR_TUNE_FILE, "cannot write %s", path);
    char arch[64];
    int cus;
    arch_info(arch, sizeof arch, &cus);
    fprintf(f, "ugemm-hip-tune 3 %s %d\n", arch, cus);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        for (auto &kv : g_table) {
            const Key &k = kv.first;
            fprintf(f, "%d %d %d %d %d %d %d %.6f %d\n", std::get<0>(k), std::get<1>(k), std::get<2>(k),
                    std::get<3>(k), std::get<4>(k), std::get<5>(k), kv.second.variant, kv.second.ms, kv.second.splits);
        }
    }
    if (fclose(f)) return fail(UGEMM_ERR_TUNE_FILE, "cannot write %s", path);
    return UGEMM_OK;
}

int ugemm_hip_tune_load(const char *path) {
    if (!path) return fail(UGEMM_ERR_ARG, "tune_load: NULL path");
    int rc = ensure_device();
    if (rc) return rc;
    return ugemm::tune_load_file(path);
}

}  // extern "C"

namespace ugemm {

int tune_load_file(const char *path) {
    FILE *f = fopen(path, "r");
    if (!f) return fail(UGEMM_ERR_TUNE_FILE, "cannot read %s", path);
    char arch[64], farch[64];
    int cus, fcus, ver;
    arch_info(arch, sizeof arch, &cus);
    if (fscanf(f, "ugemm-hip-tune %d %63s %d", &ver, farch, &fcus) != 3 || ver < 1 || ver > 3) {
        fclose(f);
        return fail(UGEMM_ERR_TUNE_FILE, "%s: not a version-1, -2 or -3 tuning cache", path);
    }
    auto split_form = [](int mode) {
        return mode == UGEMM_MODE_BF16X1 || mode == UGEMM_MODE_BF16X3 || mode == UGEMM_MODE_BF16X6 ||
               mode == UGEMM_MODE_FP16X3;
    };
    int loaded = 0;
    if (!strcmp(arch, farch) && cus == fcus) {
        int m, a, b, lm, ln, lk, v, sp = 0;
        float ms;
        while (fscanf(f, "%d %d %d %d %d %d %d %f", &m, &a, &b, &lm, &ln, &lk, &v, &ms) == 8) {
            if (ver >= 2 && fscanf(f, "%d", &sp) != 1) break;
            if (ver < 3 && split_form(m == TUNE_KEY_GROUPED ? b : m)) continue;   // older variant numbering
            const int nv = (m == TUNE_KEY_GROUPED) ? grouped_variant_count(b)
                                                   : ((m < -1 || m >= UGEMM_MODE_COUNT) ? 0 : ugemm_hip_variant_count(m));
            if (v < 0 || v >= nv || sp < 0 || sp > 64) continue;
            std::lock_guard<std::mutex> lk2(g_mu);
            g_table[Key(m, a, b, lm, ln, lk)] = Entry{v, ms, sp};
            ++loaded;
        }
    }
    fclose(f);
    return loaded;
}

}  // namespace ugemm
