// Runtime of libugemm_hip.so: status reporting, device set-up, argument
// normalisation, host-pointer staging, and the exported entry points.
//
// Boundary replaced (reference -> here):
//   sgemm_cpu / sgemm_c / sgemm_avx uut signature (ugemm.h:287-301,
//   check_sgemm.c:96-103)                  -> sgemm_hip, ugemm_hip_sgemm
//   dgemm_cpu / dgemm_avx (ugemm.h:162-176) -> dgemm_hip
//   sgemm_ocl_init / sgemm_ocl / sgemm_ocl_finish (sgemm_ocl2.h:158-224)
//                                          -> ugemm_hip_init / sgemm_hip_packed / ugemm_hip_finish
#include "internal.h"

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <mutex>
#include <utility>
#include <vector>

namespace ugemm {

// ---------------------------------------------------------------------------
// Status: per-thread code + message.
static thread_local int t_status = UGEMM_OK;
static thread_local char t_msg[256] = "";

int fail(int status, const char *fmt, ...) {
    t_status = status;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(t_msg, sizeof t_msg, fmt, ap);
    va_end(ap);
    return status;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(e == hipErrorOutOfMemory ? UGEMM_ERR_NOMEM : UGEMM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

static int ok() {
    t_status = UGEMM_OK;
    t_msg[0] = 0;
    return UGEMM_OK;
}

// ---------------------------------------------------------------------------
// Static device scratch of ugemm_hip_sgemm_noalloc_dev. Each device holds its
// own copy from the load of this code object until the process ends. In host
// code `scratch` names a host-side stand-in, never the device memory: host
// code uses the device address looked up once per device at initialisation
// (g_scratch_addr), and kernels receive that address.
static __device__ __align__(256) unsigned char scratch[UGEMM_HIP_NOALLOC_SCRATCH_BYTES];
constexpr int kMaxDevices = 64;
static void *g_scratch_addr[kMaxDevices];        // device address of `scratch` per device
static hipError_t g_scratch_err[kMaxDevices];    // result of the lookup

// Called with `device` current, under the device mutex.
static void lookup_scratch(int device) {
    if (device < 0 || device >= kMaxDevices || g_scratch_addr[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_err[device] = e;
    if (e == hipSuccess) g_scratch_addr[device] = p;
    else (void)hipGetLastError();   // reported by the entry point that needs it
}

// ---------------------------------------------------------------------------
// Device state. One device per process image of the library; the host-pointer
// entry points run on it, the device-pointer ones on the caller's stream.
struct DeviceState {
    std::mutex mu;              // guards the fields below and the staging buffer
    bool ready = false;
    int device = 0;
    int cus = 256;
    hipStream_t stream = nullptr;   // internal stream for host-pointer calls (copies)
    hipStream_t kstream = nullptr;  // second internal stream: compute of pipelined host-pointer calls
    void *staging = nullptr;        // device staging for host-pointer calls
    size_t staging_bytes = 0;
};
static DeviceState g_dev;

static int init_locked(int device, size_t max_bytes) {
    if (g_dev.ready) {
        if (device >= 0 && device != g_dev.device)
            return fail(UGEMM_ERR_ARG, "ugemm_hip_init: already initialised on device %d", g_dev.device);
    } else {
        if (device < 0) {
            const char *e = getenv("UGEMM_HIP_DEVICE");
            device = e ? atoi(e) : 0;
        }
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(UGEMM_ERR_NODEV, "no HIP device");
        if (device >= n) return fail(UGEMM_ERR_NODEV, "device %d of %d", device, n);
        UGEMM_HIP_TRY(hipSetDevice(device));
        hipDeviceProp_t prop;
        UGEMM_HIP_TRY(hipGetDeviceProperties(&prop, device));
        if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            return fail(UGEMM_ERR_NODEV, "device %d is %s; this library is built for gfx950 only", device,
                        prop.gcnArchName);
        g_dev.device = device;
        g_dev.cus = prop.multiProcessorCount;
        lookup_scratch(device);
        UGEMM_HIP_TRY(hipStreamCreateWithFlags(&g_dev.stream, hipStreamNonBlocking));
        UGEMM_HIP_TRY(hipStreamCreateWithFlags(&g_dev.kstream, hipStreamNonBlocking));
        // Keep freed workspace blocks in the default pool (no release at
        // synchronisation points), so that repeated calls do not go back to
        // the driver for memory.
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
            uint64_t keep = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
        g_dev.ready = true;
        // A saved tuning table for uut callers, who cannot call
        // ugemm_hip_tune_load themselves; a missing file is not an error.
        const char *tc = getenv("UGEMM_HIP_TUNE_CACHE");
        if (tc && *tc) (void)tune_load_file(tc);
    }
    if (max_bytes > g_dev.staging_bytes) {
        if (g_dev.staging) {
            UGEMM_HIP_TRY(hipFree(g_dev.staging));
            g_dev.staging = nullptr;
            g_dev.staging_bytes = 0;
        }
        UGEMM_HIP_TRY(hipMalloc(&g_dev.staging, max_bytes));
        g_dev.staging_bytes = max_bytes;
    }
    return UGEMM_OK;
}

int ensure_device() {
    if (g_dev.ready) return UGEMM_OK;
    std::lock_guard<std::mutex> lk(g_dev.mu);
    return init_locked(-1, 0);
}

int device_cu_count() { return g_dev.ready ? g_dev.cus : 256; }

// While a no-allocation call runs on this thread, workspace requests are
// carved from the scratch in order (one request per call today) and never
// reach the allocator; releasing is a no-op.
struct ScratchArena {
    char *base;
    size_t cap, used;
};
static thread_local ScratchArena *t_arena = nullptr;

int stream_alloc(void **ptr, size_t bytes, hipStream_t s) {
    *ptr = nullptr;
    if (bytes == 0) return UGEMM_OK;
    if (ScratchArena *a = t_arena) {
        const size_t off = (a->used + 255) / 256 * 256;
        if (off > a->cap || bytes > a->cap - off)
            return fail(UGEMM_ERR_SCRATCH,
                        "no-allocation call: workspace of %zu bytes exceeds the %zu-byte static scratch", off + bytes,
                        a->cap);
        *ptr = a->base + off;
        a->used = off + bytes;
        return UGEMM_OK;
    }
    hipError_t e = hipMallocAsync(ptr, bytes, s);
    if (e != hipSuccess) return hip_fail(e, "hipMallocAsync(workspace)");
    return UGEMM_OK;
}

int stream_free(void *ptr, hipStream_t s) {
    if (!ptr || t_arena) return UGEMM_OK;
    hipError_t e = hipFreeAsync(ptr, s);
    if (e != hipSuccess) return hip_fail(e, "hipFreeAsync(workspace)");
    return UGEMM_OK;
}

// ---------------------------------------------------------------------------
// Launch timing: a per-thread list of event pairs (events are reused).
struct TimingState {
    bool on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pairs;
    size_t used = 0;
};
static thread_local TimingState t_timing;

void timing_begin(hipStream_t s) {
    TimingState &T = t_timing;
    if (!T.on) return;
    if (T.used == T.pairs.size()) {
        hipEvent_t a = nullptr, b = nullptr;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
            if (a) (void)hipEventDestroy(a);
            T.on = false;
            return;
        }
        T.pairs.emplace_back(a, b);
    }
    (void)hipEventRecord(T.pairs[T.used].first, s);
}

void timing_end(hipStream_t s) {
    TimingState &T = t_timing;
    if (!T.on || T.used >= T.pairs.size()) return;
    (void)hipEventRecord(T.pairs[T.used].second, s);
    ++T.used;
}

// ---------------------------------------------------------------------------
// C <- beta * C kernels (alpha == 0 or K == 0). beta == 0 writes zeros; the
// FP32 version then applies the optional epilogue.
template <typename T>
__global__ void scale_kernel(int M, int N, T beta, T *__restrict__ C, long ldc, Epilogue epi) {
    const long total = (long)M * N;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const long m = i / N, n = i - m * N;
        T *p = C + m * ldc + n;
        T v = (beta == T(0)) ? T(0) : beta * *p;
        if constexpr (sizeof(T) == 4) v = apply_epilogue(epi, v, (int)m, (int)n);
        *p = v;
    }
}

template <typename T>
static int launch_scale(int M, int N, T beta, T *C, long ldc, hipStream_t s, const Epilogue &epi) {
    const long total = (long)M * N;
    const int threads = 256;
    long blocks = (total + threads - 1) / threads;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(scale_kernel<T>, dim3((unsigned)blocks), dim3(threads), 0, s, M, N, beta, C, ldc, epi);
    UGEMM_HIP_TRY(hipGetLastError());
    return UGEMM_OK;
}
int launch_scale_f32(int M, int N, float beta, float *C, long ldc, hipStream_t s, const Epilogue &epi) {
    return launch_scale(M, N, beta, C, ldc, s, epi);
}
int launch_scale_f64(int M, int N, double beta, double *C, long ldc, hipStream_t s) {
    return launch_scale(M, N, beta, C, ldc, s, Epilogue());
}

// ---------------------------------------------------------------------------
// Argument normalisation: validate, then express every call as a row-major
// GEMM. Column-major C (M x N) is row-major C^T (N x M) = op(B)^T op(A)^T, so
// a column-major call maps to row-major (transb, transa, N, M, K, B, A).
static int parse_trans(char t, bool *out) {
    switch (t) {
        case 'N': case 'n': *out = false; return 1;
        case 'T': case 't': case 'C': case 'c': *out = true; return 1;
        default: return 0;
    }
}

template <typename T>
static int normalize(char major, char transa, char transb, int M, int N, int K, T alpha, const T *A, int lda,
                     const T *B, int ldb, T beta, T *C, int ldc, Gemm<T> &g) {
    bool ta, tb, row;
    if (major == 'R' || major == 'r') row = true;
    else if (major == 'C' || major == 'c') row = false;
    else return fail(UGEMM_ERR_ARG, "major must be 'R' or 'C' (got 0x%02x)", (unsigned char)major);
    if (!parse_trans(transa, &ta)) return fail(UGEMM_ERR_ARG, "bad transa 0x%02x", (unsigned char)transa);
    if (!parse_trans(transb, &tb)) return fail(UGEMM_ERR_ARG, "bad transb 0x%02x", (unsigned char)transb);
    if (M < 0 || N < 0 || K < 0) return fail(UGEMM_ERR_ARG, "negative size M=%d N=%d K=%d", M, N, K);
    // Leading dimensions as in BLAS: at least the length of a stored row
    // (row-major) or column (column-major), and at least 1.
    auto need = [](int x) { return x > 1 ? x : 1; };
    const int la = row ? (ta ? M : K) : (ta ? K : M);
    const int lb = row ? (tb ? K : N) : (tb ? N : K);
    const int lc = row ? N : M;
    if (lda < need(la)) return fail(UGEMM_ERR_ARG, "lda=%d < %d", lda, need(la));
    if (ldb < need(lb)) return fail(UGEMM_ERR_ARG, "ldb=%d < %d", ldb, need(lb));
    if (ldc < need(lc)) return fail(UGEMM_ERR_ARG, "ldc=%d < %d", ldc, need(lc));
    if (M > 0 && N > 0) {
        if (!C) return fail(UGEMM_ERR_ARG, "C is NULL");
        if (K > 0 && alpha != T(0) && (!A || !B)) return fail(UGEMM_ERR_ARG, "A or B is NULL");
    }
    if (row) {
        g = Gemm<T>{M, N, K, alpha, beta, A, lda, B, ldb, C, ldc, ta, tb};
    } else {
        g = Gemm<T>{N, M, K, alpha, beta, B, ldb, A, lda, C, ldc, tb, ta};
    }
    return UGEMM_OK;
}

// ---------------------------------------------------------------------------
// Device-pointer execution of a normalised GEMM.
// Variant and split-K count: forced (ugemm_hip_force_variant/_splits), else
// the tuning table's entry for the shape bucket, else the built-in choice.
static int choose_variant(const Gemm<float> &g, int mode, int *splits) {
    int v = forced_variant(mode), tsplits = 0;
    const int fs = forced_splits(mode);
    if (v < 0) v = tuned_variant(mode, g.ta, g.tb, g.M, g.N, g.K, &tsplits);
    *splits = fs > 0 ? fs : tsplits;
    if (v >= 0) return v;
    switch (mode) {
        case UGEMM_MODE_FP32: return fp32_pick_variant(g);
        case UGEMM_MODE_FP32_MFMA: return m32_pick_variant(g);
        case UGEMM_MODE_FP8_BLOCK: return -1;   // launch_fp8 picks by grid size
        default: return mfma_pick_variant(g, mode);
    }
}

int run_sgemm(const Gemm<float> &g, int mode, int variant, hipStream_t s) {
    if (g.M == 0 || g.N == 0) return UGEMM_OK;
    if (g.alpha == 0.0f || g.K == 0) {
        if (g.beta == 1.0f && !g.epi.active()) return UGEMM_OK;
        return launch_scale_f32(g.M, g.N, g.beta, g.C, g.ldc, s, g.epi);
    }
    if (g.epi.active() && (mode == UGEMM_MODE_FP8_BLOCK))
        return fail(UGEMM_ERR_UNSUPPORTED, "bias/activation epilogue: not provided for the FP8 form");
    Gemm<float> gs = g;
    if (variant < 0) {
        variant = choose_variant(g, mode, &gs.splits);
    } else if (gs.splits == 0 && forced_splits(mode) > 0) {
        gs.splits = forced_splits(mode);
    }
    switch (mode) {
        case UGEMM_MODE_FP32: return launch_fp32(gs, variant, s);
        case UGEMM_MODE_FP32_MFMA: return launch_m32(gs, variant, s);
        case UGEMM_MODE_BF16X1:
        case UGEMM_MODE_BF16X3:
        case UGEMM_MODE_BF16X6:
        case UGEMM_MODE_FP16X3: return launch_mfma(gs, mode, variant, s);
        case UGEMM_MODE_FP8_BLOCK: return launch_fp8(gs, variant, s);
        default: return fail(UGEMM_ERR_ARG, "unknown mode %d", mode);
    }
}

int run_dgemm(const Gemm<double> &g, int variant, hipStream_t s) {
    if (g.M == 0 || g.N == 0) return UGEMM_OK;
    if (g.alpha == 0.0 || g.K == 0) {
        if (g.beta == 1.0) return UGEMM_OK;
        return launch_scale_f64(g.M, g.N, g.beta, g.C, g.ldc, s);
    }
    if (variant < 0) {
        variant = forced_variant(-1);
        if (variant < 0) variant = dgemm_pick_variant(g);
    }
    Gemm<double> gs = g;
    if (gs.splits == 0) gs.splits = forced_splits(-1);
    return launch_dgemm(gs, variant, s);
}

// ---------------------------------------------------------------------------
// Host-pointer execution: stage A, B (and C when it is read) into the device
// staging buffer with 2-D copies that keep the caller's leading dimensions,
// run on the internal stream, copy C back row by row (the gaps between rows
// of a padded C are never written), and wait.
static size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

template <typename T>
static int run_host(const Gemm<T> &h, int mode) {
    int st = ensure_device();
    if (st) return st;
    if (h.M == 0 || h.N == 0) return UGEMM_OK;
    if (h.beta == T(1) && (h.alpha == T(0) || h.K == 0)) return UGEMM_OK;
    std::lock_guard<std::mutex> lk(g_dev.mu);
    UGEMM_HIP_TRY(hipSetDevice(g_dev.device));
    const bool need_ab = h.alpha != T(0) && h.K > 0;
    const bool need_c = h.beta != T(0);
    // Stored extents (rows x contiguous length) of the three operands.
    const long a_rows = h.ta ? h.K : h.M, a_len = h.ta ? h.M : h.K;
    const long b_rows = h.tb ? h.N : h.K, b_len = h.tb ? h.K : h.N;
    const size_t a_bytes = need_ab ? round_up((size_t)a_rows * h.lda * sizeof(T), 256) : 0;
    const size_t b_bytes = need_ab ? round_up((size_t)b_rows * h.ldb * sizeof(T), 256) : 0;
    const size_t c_bytes = round_up((size_t)h.M * h.ldc * sizeof(T), 256);
    const size_t total = a_bytes + b_bytes + c_bytes;
    if (total > g_dev.staging_bytes) {
        if (g_dev.staging) {
            UGEMM_HIP_TRY(hipFree(g_dev.staging));
            g_dev.staging = nullptr;
            g_dev.staging_bytes = 0;
        }
        UGEMM_HIP_TRY(hipMalloc(&g_dev.staging, total));
        g_dev.staging_bytes = total;
    }
    char *base = (char *)g_dev.staging;
    T *dA = (T *)base, *dB = (T *)(base + a_bytes), *dC = (T *)(base + a_bytes + b_bytes);
    hipStream_t s = g_dev.stream;
    const size_t es = sizeof(T);
    // Inputs: one contiguous copy of the span an operand occupies (its last
    // row ends at (rows - 1) * ld + len, inside the caller's buffer); the
    // gaps between rows travel too, which costs less than a strided copy.
    auto span = [](long rows, long ld, long len) { return (size_t)((rows - 1) * ld + len); };
    int rc = UGEMM_OK;
    // Large calls whose A rows and C rows are contiguous blocks: split along
    // M so that block b's GEMM (second stream) runs while block b+1's rows of
    // A come in and block b-1's rows of C go out (copy stream).
    if (!h.ta && h.ldc == h.N && need_ab && h.M >= 2048 && total >= (64u << 20)) {
        hipStream_t ks = g_dev.kstream;
        const int nb = (int)std::min<long>(8, std::max<long>(2, h.M / 2048));
        const long Mb = (long)round_up((size_t)((h.M + nb - 1) / nb), 256);
        std::vector<hipEvent_t> evA, evC;
        auto cleanup = [&]() {
            (void)hipStreamSynchronize(ks);
            (void)hipStreamSynchronize(s);
            for (hipEvent_t e : evA) (void)hipEventDestroy(e);
            for (hipEvent_t e : evC) (void)hipEventDestroy(e);
        };
        rc = staged_h2d(dB, h.B, span(b_rows, h.ldb, b_len) * es, s);
        long prev0 = -1, prev1 = -1;
        for (long m0 = 0; !rc && m0 < h.M; m0 += Mb) {
            const long m1 = std::min<long>(h.M, m0 + Mb);
            hipEvent_t ea = nullptr, ec = nullptr;
            if (hipEventCreateWithFlags(&ea, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&ec, hipEventDisableTiming) != hipSuccess) {
                if (ea) (void)hipEventDestroy(ea);
                rc = fail(UGEMM_ERR_HIP, "pipelined copy: events");
                break;
            }
            evA.push_back(ea);
            evC.push_back(ec);
            rc = staged_h2d(dA + m0 * h.lda, h.A + m0 * h.lda, span(m1 - m0, h.lda, a_len) * es, s);
            if (!rc && need_c) rc = staged_h2d(dC + m0 * h.N, h.C + m0 * h.N, (size_t)(m1 - m0) * h.N * es, s);
            if (rc) break;
            if (hipEventRecord(ea, s) != hipSuccess || hipStreamWaitEvent(ks, ea, 0) != hipSuccess) {
                rc = fail(UGEMM_ERR_HIP, "pipelined copy: ordering");
                break;
            }
            Gemm<T> d = h;
            d.M = (int)(m1 - m0);
            d.A = dA + m0 * h.lda;
            d.B = dB;
            d.C = dC + m0 * h.N;
            if constexpr (sizeof(T) == 4) rc = run_sgemm(d, mode, -1, ks);
            else rc = run_dgemm(d, -1, ks);
            if (rc) break;
            if (hipEventRecord(ec, ks) != hipSuccess) {
                rc = fail(UGEMM_ERR_HIP, "pipelined copy: ordering");
                break;
            }
            if (prev0 >= 0) {
                if (hipStreamWaitEvent(s, evC[evC.size() - 2], 0) != hipSuccess) {
                    rc = fail(UGEMM_ERR_HIP, "pipelined copy: ordering");
                    break;
                }
                rc = staged_d2h(h.C + prev0 * h.N, dC + prev0 * h.N, (size_t)(prev1 - prev0) * h.N * es, s);
            }
            prev0 = m0;
            prev1 = m1;
        }
        if (!rc && prev0 >= 0) {
            if (hipStreamWaitEvent(s, evC.back(), 0) != hipSuccess) rc = fail(UGEMM_ERR_HIP, "pipelined copy: ordering");
            else rc = staged_d2h(h.C + prev0 * h.N, dC + prev0 * h.N, (size_t)(prev1 - prev0) * h.N * es, s);
        }
        cleanup();
        return rc;
    }
    if (need_ab) {
        rc = staged_h2d(dA, h.A, span(a_rows, h.lda, a_len) * es, s);
        if (!rc) rc = staged_h2d(dB, h.B, span(b_rows, h.ldb, b_len) * es, s);
    }
    if (!rc && need_c) rc = staged_h2d(dC, h.C, span(h.M, h.ldc, h.N) * es, s);
    if (rc) {
        (void)hipStreamSynchronize(s);
        return rc;
    }
    Gemm<T> d = h;
    d.A = dA;
    d.B = dB;
    d.C = dC;
    if constexpr (sizeof(T) == 4) rc = run_sgemm(d, mode, -1, s);
    else rc = run_dgemm(d, -1, s);
    if (rc) {
        (void)hipStreamSynchronize(s);
        return rc;
    }
    // Output: contiguous when C has no gaps; otherwise row by row, so that
    // the gaps between C's rows keep the caller's values.
    if (h.ldc == h.N) {
        rc = staged_d2h(h.C, dC, (size_t)h.M * h.N * es, s);
    } else {
        hipError_t e = hipMemcpy2DAsync(h.C, h.ldc * es, dC, h.ldc * es, h.N * es, h.M, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) rc = hip_fail(e, "copy out");
    }
    hipError_t e = hipStreamSynchronize(s);
    if (!rc && e != hipSuccess) rc = hip_fail(e, "synchronize");
    return rc;
}

static int env_mode() {
    // Read once (thread-safe static initialisation).
    static const int mode = [] {
        const char *e = getenv("UGEMM_HIP_MODE");
        if (e) {
            if (!strcmp(e, "bf16x3")) return (int)UGEMM_MODE_BF16X3;
            if (!strcmp(e, "bf16x1")) return (int)UGEMM_MODE_BF16X1;
            if (!strcmp(e, "bf16x6")) return (int)UGEMM_MODE_BF16X6;
            if (!strcmp(e, "fp8")) return (int)UGEMM_MODE_FP8_BLOCK;
            if (!strcmp(e, "fp16x3")) return (int)UGEMM_MODE_FP16X3;
            if (!strcmp(e, "fp32_mfma")) return (int)UGEMM_MODE_FP32_MFMA;
        }
        return (int)UGEMM_MODE_FP32;
    }();
    return mode;
}

// Implicit-GEMM convolution: M = ch, N = hcol * wcol, K = ich * k * k.
int conv_shape(int ich, int w, int h, int k, int pad, int stride, int ch, int *hcol, int *wcol) {
    if (ich < 1 || w < 1 || h < 1 || k < 1 || ch < 1 || stride < 1 || pad < 0 || k > 32767)
        return fail(UGEMM_ERR_ARG, "conv: bad shape ich=%d w=%d h=%d k=%d pad=%d stride=%d ch=%d", ich, w, h, k, pad,
                    stride, ch);
    if (h + 2 * pad < k || w + 2 * pad < k) return fail(UGEMM_ERR_ARG, "conv: kernel larger than the padded image");
    *hcol = (h + 2 * pad - k) / stride + 1;
    *wcol = (w + 2 * pad - k) / stride + 1;
    if ((long)ich * k * k > 0x7fffffffL || (long)*hcol * *wcol > 0x7fffffffL || (long)ich * h * w > 0x7fffffffL)
        return fail(UGEMM_ERR_ARG, "conv: sizes exceed 32-bit indexing");
    return UGEMM_OK;
}

__global__ void conv_ktab_kernel(int2 *__restrict__ ktab, int K, int k, int H, int W) {
    const int kk = blockIdx.x * blockDim.x + threadIdx.x;
    if (kk >= K) return;
    const int c = kk / (k * k), r = kk - c * k * k, i = r / k, j = r - i * k;
    ktab[kk] = make_int2(c * H * W + i * W + j, (i << 16) | j);
}

static int run_conv(const float *d_in, int ich, int w, int h, const float *d_weights, int k, int pad, int stride,
                    float *d_out, int ch, const Epilogue &epi, int mode, hipStream_t s) {
    int hcol, wcol;
    int rc = conv_shape(ich, w, h, k, pad, stride, ch, &hcol, &wcol);
    if (rc) return rc;
    if (mode < 0 || mode >= UGEMM_MODE_COUNT) return fail(UGEMM_ERR_ARG, "conv: unknown mode %d", mode);
    if (mode == UGEMM_MODE_FP8_BLOCK) return fail(UGEMM_ERR_UNSUPPORTED, "conv: the FP8 form is not provided");
    if (!d_in || !d_weights || !d_out) return fail(UGEMM_ERR_ARG, "conv: NULL buffer");
    const int K = ich * k * k;
    void *ws = nullptr;
    rc = stream_alloc(&ws, (size_t)K * sizeof(int2), s);
    if (rc) return rc;
    int2 *ktab = (int2 *)ws;
    hipLaunchKernelGGL(conv_ktab_kernel, dim3((unsigned)((K + 255) / 256)), dim3(256), 0, s, ktab, K, k, h, w);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = hip_fail(e, "conv: offset table");
    if (!rc) {
        Gemm<float> g{ch, hcol * wcol, K, 1.0f, 0.0f, d_weights, K, nullptr, 0, d_out, (long)hcol * wcol, false, false};
        g.epi = epi;
        g.splits = forced_splits(mode);
        const ConvDesc cd{d_in, ktab, h, w, wcol, stride, pad};
        const int v = forced_variant(mode);
        rc = mode == UGEMM_MODE_FP32        ? launch_conv_fp32(g, cd, v, s)
             : mode == UGEMM_MODE_FP32_MFMA ? launch_conv_m32(g, cd, v, s)
                                            : launch_mfma_conv(g, cd, mode, v, s);
    }
    const int rc2 = stream_free(ws, s);
    return rc ? rc : rc2;
}

static int axpy_dispatch(int N, float alpha, const float *x, int incx, float *y, int incy, hipStream_t s) {
    return launch_saxpy(N, alpha, x, incx, y, incy, s);
}
static int axpy_dispatch(int N, double alpha, const double *x, int incx, double *y, int incy, hipStream_t s) {
    return launch_daxpy(N, alpha, x, incx, y, incy, s);
}
static int gemv_dispatch(char trans, int M, int N, float alpha, const float *A, int lda, const float *x, int incx,
                         float beta, float *y, int incy, hipStream_t s) {
    return launch_sgemv(trans, M, N, alpha, A, lda, x, incx, beta, y, incy, s);
}
static int gemv_dispatch(char trans, int M, int N, double alpha, const double *A, int lda, const double *x, int incx,
                         double beta, double *y, int incy, hipStream_t s) {
    return launch_dgemv(trans, M, N, alpha, A, lda, x, incx, beta, y, incy, s);
}

// Host-pointer versions (the argument lists of saxpy_cpu, daxpy_cpu,
// sgemv_cpu, dgemv_cpu; ugemm.h:76-160): stage the spans the call touches,
// run on the internal stream, copy y back, wait.
template <typename T>
static void host_axpy(int N, T alpha, const T *x, int incx, T *y, int incy) {
    int rc = ensure_device();
    if (rc) return;
    if (N < 0 || incx < 1 || incy < 1) {
        fail(UGEMM_ERR_ARG, "axpy: N=%d incx=%d incy=%d", N, incx, incy);
        return;
    }
    if (N == 0) {
        ok();
        return;
    }
    if (!x || !y) {
        fail(UGEMM_ERR_ARG, "axpy: NULL vector");
        return;
    }
    std::lock_guard<std::mutex> lk(g_dev.mu);
    hipStream_t s = g_dev.stream;
    const size_t nx = (size_t)(N - 1) * incx + 1, ny = (size_t)(N - 1) * incy + 1;
    void *ws = nullptr;
    if (hipSetDevice(g_dev.device) != hipSuccess || stream_alloc(&ws, (nx + ny) * sizeof(T) + 16, s)) return;
    T *dx = (T *)ws, *dy = dx + nx;
    hipError_t e = hipMemcpyAsync(dx, x, nx * sizeof(T), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dy, y, ny * sizeof(T), hipMemcpyHostToDevice, s);
    rc = e == hipSuccess ? axpy_dispatch(N, alpha, (const T *)dx, incx, dy, incy, s) : hip_fail(e, "axpy: copy in");
    if (!rc && (e = hipMemcpyAsync(y, dy, ny * sizeof(T), hipMemcpyDeviceToHost, s)) != hipSuccess)
        rc = hip_fail(e, "axpy: copy out");
    const int rc2 = stream_free(ws, s);
    if ((e = hipStreamSynchronize(s)) != hipSuccess && !rc) rc = hip_fail(e, "axpy: synchronize");
    if (!rc && !rc2) ok();
}

// Host-pointer three-operand axpy (the argument list of saxpy_avx /
// daxpy_avx, ugemm.h:41-74): stage x and y, run, copy z back, wait.
static int axpy3_dispatch(int N, float a, const float *x, const float *y, float *z, hipStream_t s) {
    return launch_saxpy3(N, a, x, y, z, s);
}
static int axpy3_dispatch(int N, double a, const double *x, const double *y, double *z, hipStream_t s) {
    return launch_daxpy3(N, a, x, y, z, s);
}
template <typename T>
static void host_axpy3(int N, T a, const T *x, const T *y, T *z) {
    int rc = ensure_device();
    if (rc) return;
    if (N < 0) {
        fail(UGEMM_ERR_ARG, "axpy3: N=%d", N);
        return;
    }
    if (N == 0) {
        ok();
        return;
    }
    if (!x || !y || !z) {
        fail(UGEMM_ERR_ARG, "axpy3: NULL vector");
        return;
    }
    std::lock_guard<std::mutex> lk(g_dev.mu);
    hipStream_t s = g_dev.stream;
    const size_t n = (size_t)N, span = (n * sizeof(T) + 255) / 256 * 256;
    void *ws = nullptr;
    if (hipSetDevice(g_dev.device) != hipSuccess || stream_alloc(&ws, 2 * span, s)) return;
    T *dx = (T *)ws, *dy = (T *)((char *)ws + span);
    hipError_t e = hipMemcpyAsync(dx, x, n * sizeof(T), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dy, y, n * sizeof(T), hipMemcpyHostToDevice, s);
    // z overwrites y's device copy (the kernel reads y[i] before writing z[i]).
    rc = e == hipSuccess ? axpy3_dispatch(N, a, dx, dy, dy, s) : hip_fail(e, "axpy3: copy in");
    if (!rc && (e = hipMemcpyAsync(z, dy, n * sizeof(T), hipMemcpyDeviceToHost, s)) != hipSuccess)
        rc = hip_fail(e, "axpy3: copy out");
    const int rc2 = stream_free(ws, s);
    if ((e = hipStreamSynchronize(s)) != hipSuccess && !rc) rc = hip_fail(e, "axpy3: synchronize");
    if (!rc && !rc2) ok();
}

template <typename T>
static void host_gemv(char trans, int M, int N, T alpha, const T *A, int lda, const T *x, int incx, T beta, T *y,
                      int incy) {
    int rc = ensure_device();
    if (rc) return;
    const bool t = !(trans == 'N' || trans == 'n');
    if (M < 0 || N < 0 || incx < 1 || incy < 1 || lda < std::max(1, t ? N : M)) {
        fail(UGEMM_ERR_ARG, "gemv: M=%d N=%d lda=%d incx=%d incy=%d", M, N, lda, incx, incy);
        return;
    }
    if (M == 0) {
        ok();
        return;
    }
    if (!y || (N > 0 && (!A || !x))) {
        fail(UGEMM_ERR_ARG, "gemv: NULL operand");
        return;
    }
    std::lock_guard<std::mutex> lk(g_dev.mu);
    hipStream_t s = g_dev.stream;
    // A spans (cols - 1) * lda + rows elements: rows x cols column-major ('N': M x N, 'T': N x M).
    const size_t na = N > 0 ? (size_t)((t ? M : N) - 1) * lda + (t ? N : M) : 0;
    const size_t nx = N > 0 ? (size_t)(N - 1) * incx + 1 : 0, ny = (size_t)(M - 1) * incy + 1;
    void *ws = nullptr;
    if (hipSetDevice(g_dev.device) != hipSuccess || stream_alloc(&ws, (na + nx + ny) * sizeof(T) + 16, s)) return;
    T *dA = (T *)ws, *dx = dA + na, *dy = dx + nx;
    hipError_t e = hipSuccess;
    if (na) e = hipMemcpyAsync(dA, A, na * sizeof(T), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && nx) e = hipMemcpyAsync(dx, x, nx * sizeof(T), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && beta != T(0)) e = hipMemcpyAsync(dy, y, ny * sizeof(T), hipMemcpyHostToDevice, s);
    rc = e == hipSuccess ? gemv_dispatch(trans, M, N, alpha, (const T *)dA, lda, (const T *)dx, incx, beta, dy, incy, s)
                         : hip_fail(e, "gemv: copy in");
    if (!rc) {
        // Only y's elements are written back (the gaps of a strided y keep the caller's values).
        e = incy == 1 ? hipMemcpyAsync(y, dy, (size_t)M * sizeof(T), hipMemcpyDeviceToHost, s)
                      : hipMemcpy2DAsync(y, incy * sizeof(T), dy, incy * sizeof(T), sizeof(T), M,
                                         hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) rc = hip_fail(e, "gemv: copy out");
    }
    const int rc2 = stream_free(ws, s);
    if ((e = hipStreamSynchronize(s)) != hipSuccess && !rc) rc = hip_fail(e, "gemv: synchronize");
    if (!rc && !rc2) ok();
}

}  // namespace ugemm

using namespace ugemm;

// ---------------------------------------------------------------------------
// Exported C ABI.
extern "C" {

const char *ugemm_hip_version(void) { return "ugemm-hip 0.1 (gfx950)"; }

int ugemm_hip_last_error(void) { return t_status; }
const char *ugemm_hip_last_error_message(void) { return t_msg; }

const char *ugemm_hip_status_string(int status) {
    switch (status) {
        case UGEMM_OK: return "ok";
        case UGEMM_ERR_ARG: return "invalid argument";
        case UGEMM_ERR_HIP: return "HIP runtime error";
        case UGEMM_ERR_NOMEM: return "out of memory";
        case UGEMM_ERR_NODEV: return "no usable gfx950 device";
        case UGEMM_ERR_UNSUPPORTED: return "unsupported mode or layout";
        case UGEMM_ERR_TUNE_FILE: return "tuning cache file error";
        case UGEMM_ERR_SCRATCH: return "workspace exceeds the static scratch of the no-allocation entry point";
        default: return "unknown status";
    }
}

int ugemm_hip_init(int device, size_t max_bytes) {
    std::lock_guard<std::mutex> lk(g_dev.mu);
    int rc = init_locked(device, max_bytes);
    return rc ? rc : ok();
}

void ugemm_hip_finish(void) {
    std::lock_guard<std::mutex> lk(g_dev.mu);
    if (!g_dev.ready) return;
    (void)hipSetDevice(g_dev.device);
    if (g_dev.stream) (void)hipStreamSynchronize(g_dev.stream);
    release_staging();
    if (g_dev.staging) (void)hipFree(g_dev.staging);
    if (g_dev.stream) (void)hipStreamDestroy(g_dev.stream);
    if (g_dev.kstream) {
        (void)hipStreamSynchronize(g_dev.kstream);
        (void)hipStreamDestroy(g_dev.kstream);
    }
    g_dev.kstream = nullptr;
    g_dev.staging = nullptr;
    g_dev.staging_bytes = 0;
    g_dev.stream = nullptr;
    g_dev.ready = false;
}

int ugemm_hip_sgemm(char major, char transa, char transb, int M, int N, int K, float alpha, const float *A, int lda,
                    const float *B, int ldb, float beta, float *C, int ldc, int mode) {
    if (mode < 0 || mode >= UGEMM_MODE_COUNT) return fail(UGEMM_ERR_ARG, "unknown mode %d", mode);
    Gemm<float> g;
    int rc = normalize(major, transa, transb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, g);
    if (rc) return rc;
    rc = run_host(g, mode);
    return rc ? rc : ok();
}

void sgemm_hip(char major, char transa, char transb, int M, int N, int K, float alpha, const float *A, int lda,
               const float *B, int ldb, float beta, float *C, int ldc) {
    ugemm_hip_sgemm(major, transa, transb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, env_mode());
}

void dgemm_hip(char major, char transa, char transb, int M, int N, int K, double alpha, const double *A, int lda,
               const double *B, int ldb, double beta, double *C, int ldc) {
    Gemm<double> g;
    int rc = normalize(major, transa, transb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, g);
    if (rc) return;
    rc = run_host(g, 0);
    if (!rc) ok();
}

void sgemm_hip_packed(char ta, char tb, int m, int n, int k, float alpha, float *a, float *b, float beta, float *c) {
    // Packed row-major operands, as sgemm_ocl (sgemm_ocl2.h:166-219): no
    // leading dimensions, a 'T' operand is stored transposed and packed.
    bool tA = !(ta == 'N' || ta == 'n'), tB = !(tb == 'N' || tb == 'n');
    ugemm_hip_sgemm('R', ta, tb, m, n, k, alpha, a, tA ? m : k, b, tB ? k : n, beta, c, n, env_mode());
}

int ugemm_hip_sgemm_dev(char major, char transa, char transb, int M, int N, int K, float alpha, const float *dA,
                        int lda, const float *dB, int ldb, float beta, float *dC, int ldc, int mode, void *stream) {
    if (mode < 0 || mode >= UGEMM_MODE_COUNT) return fail(UGEMM_ERR_ARG, "unknown mode %d", mode);
    int rc = ensure_device();
    if (rc) return rc;
    Gemm<float> g;
    rc = normalize(major, transa, transb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, g);
    if (rc) return rc;
    rc = run_sgemm(g, mode, -1, (hipStream_t)stream);
    return rc ? rc : ok();
}

int ugemm_hip_sgemm_noalloc_dev(char major, char transa, char transb, int M, int N, int K, float alpha,
                                const float *dA, int lda, const float *dB, int ldb, float beta, float *dC, int ldc,
                                int mode, void *stream) {
    if (mode < 0 || mode >= UGEMM_MODE_COUNT) return fail(UGEMM_ERR_ARG, "unknown mode %d", mode);
    int rc = ensure_device();
    if (rc) return rc;
    const int dev = g_dev.device;
    if (dev >= kMaxDevices) return fail(UGEMM_ERR_NODEV, "no-allocation call: device %d has no scratch slot", dev);
    if (!g_scratch_addr[dev])
        return hip_fail(g_scratch_err[dev] != hipSuccess ? g_scratch_err[dev] : hipErrorInvalidSymbol,
                        "no-allocation call: hipGetSymbolAddress(scratch)");
    Gemm<float> g;
    rc = normalize(major, transa, transb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, g);
    if (rc) return rc;
    ScratchArena arena{(char *)g_scratch_addr[dev], (size_t)UGEMM_HIP_NOALLOC_SCRATCH_BYTES, 0};
    t_arena = &arena;
    rc = run_sgemm(g, mode, -1, (hipStream_t)stream);
    t_arena = nullptr;
    return rc ? rc : ok();
}

int ugemm_hip_sgemm_ex_dev(char major, char transa, char transb, int M, int N, int K, float alpha,
                           const float *dA, int lda, const float *dB, int ldb, float beta, float *dC, int ldc,
                           int mode, const float *d_bias, int bias_dim, int act, float slope, void *stream) {
    if (mode < 0 || mode >= UGEMM_MODE_COUNT) return fail(UGEMM_ERR_ARG, "unknown mode %d", mode);
    if (bias_dim != UGEMM_BIAS_ROWS && bias_dim != UGEMM_BIAS_COLS) return fail(UGEMM_ERR_ARG, "bad bias_dim %d", bias_dim);
    if (act != UGEMM_ACT_NONE && act != UGEMM_ACT_LEAKY_RELU) return fail(UGEMM_ERR_ARG, "bad activation %d", act);
    int rc = ensure_device();
    if (rc) return rc;
    Gemm<float> g;
    rc = normalize(major, transa, transb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, g);
    if (rc) return rc;
    const bool row = major == 'R' || major == 'r';
    // Column-major calls run transposed (normalize): a per-row bias of the
    // caller's C is then a per-column bias of the kernel's C.
    g.epi.bias = d_bias;
    g.epi.bias_rows = (bias_dim == UGEMM_BIAS_ROWS) == row ? 1 : 0;
    g.epi.act = act;
    g.epi.slope = slope;
    rc = run_sgemm(g, mode, -1, (hipStream_t)stream);
    return rc ? rc : ok();
}

int ugemm_hip_conv2d_dev(const float *d_in, int ich, int w, int h, const float *d_weights, int k, int pad, int stride,
                         float *d_out, int ch, const float *d_bias, int act, float slope, int mode, void *stream) {
    if (act != UGEMM_ACT_NONE && act != UGEMM_ACT_LEAKY_RELU) return fail(UGEMM_ERR_ARG, "bad activation %d", act);
    int rc = ensure_device();
    if (rc) return rc;
    Epilogue epi;
    epi.bias = d_bias;
    epi.bias_rows = 1;
    epi.act = act;
    epi.slope = slope;
    rc = run_conv(d_in, ich, w, h, d_weights, k, pad, stride, d_out, ch, epi, mode, (hipStream_t)stream);
    return rc ? rc : ok();
}

void convolution_LReLU_hip(float *inputs, int ich, int w, int h, float *weights, int k, int pad, int stride,
                           float *outputs, int ch, float *bias) {
    // Host pointers, synchronous: the argument list of gl_convolution_LReLU
    // (sgemm_gl1.h:191) and ocl_convolution_LReLU (sgemm_ocl1.h:302).
    int rc = ensure_device();
    int hcol = 0, wcol = 0;
    if (!rc) rc = conv_shape(ich, w, h, k, pad, stride, ch, &hcol, &wcol);
    if (!rc && (!inputs || !weights || !outputs || !bias)) rc = fail(UGEMM_ERR_ARG, "conv: NULL buffer");
    if (rc) return;
    std::lock_guard<std::mutex> lk(g_dev.mu);
    if (hipSetDevice(g_dev.device) != hipSuccess) {
        hip_fail(hipGetLastError(), "hipSetDevice");
        return;
    }
    hipStream_t s = g_dev.stream;
    const size_t n_in = (size_t)ich * h * w, n_w = (size_t)ch * ich * k * k, n_out = (size_t)ch * hcol * wcol;
    void *ws = nullptr;
    rc = stream_alloc(&ws, (n_in + n_w + n_out + (size_t)ch) * sizeof(float), s);
    if (rc) return;
    float *d_in = (float *)ws, *d_w = d_in + n_in, *d_out = d_w + n_w, *d_b = d_out + n_out;
    hipError_t e = hipMemcpyAsync(d_in, inputs, n_in * sizeof(float), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_w, weights, n_w * sizeof(float), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_b, bias, (size_t)ch * sizeof(float), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) rc = hip_fail(e, "conv: copy in");
    if (!rc) {
        Epilogue epi;
        epi.bias = d_b;
        epi.act = UGEMM_ACT_LEAKY_RELU;
        epi.slope = 0.1f;
        rc = run_conv(d_in, ich, w, h, d_w, k, pad, stride, d_out, ch, epi, UGEMM_MODE_FP32, s);
    }
    if (!rc && (e = hipMemcpyAsync(outputs, d_out, n_out * sizeof(float), hipMemcpyDeviceToHost, s)) != hipSuccess)
        rc = hip_fail(e, "conv: copy out");
    const int rc2 = stream_free(ws, s);
    if ((e = hipStreamSynchronize(s)) != hipSuccess && !rc) rc = hip_fail(e, "conv: synchronize");
    if (!rc && !rc2) ok();
}

// ---- BLAS-1/2 (blas12.hip) ----
int ugemm_hip_saxpy_dev(int N, float alpha, const float *dx, int incx, float *dy, int incy, void *stream) {
    int rc = ensure_device();
    if (!rc) rc = launch_saxpy(N, alpha, dx, incx, dy, incy, (hipStream_t)stream);
    return rc ? rc : ok();
}
int ugemm_hip_daxpy_dev(int N, double alpha, const double *dx, int incx, double *dy, int incy, void *stream) {
    int rc = ensure_device();
    if (!rc) rc = launch_daxpy(N, alpha, dx, incx, dy, incy, (hipStream_t)stream);
    return rc ? rc : ok();
}
int ugemm_hip_sgemv_dev(char trans, int M, int N, float alpha, const float *dA, int lda, const float *dx, int incx,
                        float beta, float *dy, int incy, void *stream) {
    int rc = ensure_device();
    if (!rc) rc = launch_sgemv(trans, M, N, alpha, dA, lda, dx, incx, beta, dy, incy, (hipStream_t)stream);
    return rc ? rc : ok();
}
int ugemm_hip_dgemv_dev(char trans, int M, int N, double alpha, const double *dA, int lda, const double *dx,
                        int incx, double beta, double *dy, int incy, void *stream) {
    int rc = ensure_device();
    if (!rc) rc = launch_dgemv(trans, M, N, alpha, dA, lda, dx, incx, beta, dy, incy, (hipStream_t)stream);
    return rc ? rc : ok();
}

void saxpy_hip(int N, float alpha, const float *x, int incx, float *y, int incy) {
    host_axpy<float>(N, alpha, x, incx, y, incy);
}
void daxpy_hip(int N, double alpha, const double *x, int incx, double *y, int incy) {
    host_axpy<double>(N, alpha, x, incx, y, incy);
}
void saxpy_avx_hip(int N, float a, float *x, float *y, float *z) { host_axpy3<float>(N, a, x, y, z); }
void daxpy_avx_hip(int N, double a, double *x, double *y, double *z) { host_axpy3<double>(N, a, x, y, z); }
int ugemm_hip_saxpy3_dev(int N, float a, const float *dx, const float *dy, float *dz, void *stream) {
    int rc = ensure_device();
    if (!rc) rc = launch_saxpy3(N, a, dx, dy, dz, (hipStream_t)stream);
    return rc ? rc : ok();
}
int ugemm_hip_daxpy3_dev(int N, double a, const double *dx, const double *dy, double *dz, void *stream) {
    int rc = ensure_device();
    if (!rc) rc = launch_daxpy3(N, a, dx, dy, dz, (hipStream_t)stream);
    return rc ? rc : ok();
}
void sgemv_hip(char trans, int M, int N, float alpha, const float *A, int lda, const float *x, int incx, float beta,
               float *y, int incy) {
    host_gemv<float>(trans, M, N, alpha, A, lda, x, incx, beta, y, incy);
}
void dgemv_hip(char trans, int M, int N, double alpha, const double *A, int lda, const double *x, int incx,
               double beta, double *y, int incy) {
    host_gemv<double>(trans, M, N, alpha, A, lda, x, incx, beta, y, incy);
}

int ugemm_hip_dgemm_dev(char major, char transa, char transb, int M, int N, int K, double alpha, const double *dA,
                        int lda, const double *dB, int ldb, double beta, double *dC, int ldc, void *stream) {
    int rc = ensure_device();
    if (rc) return rc;
    Gemm<double> g;
    rc = normalize(major, transa, transb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, g);
    if (rc) return rc;
    rc = run_dgemm(g, -1, (hipStream_t)stream);
    return rc ? rc : ok();
}

int ugemm_hip_mfma_ceiling(const void *d_operands, int fp16, int rounds, double *tflops) {
    int rc = ensure_device();
    if (!rc) rc = mfma_ceiling(d_operands, fp16, rounds, tflops);
    return rc ? rc : ok();
}

int ugemm_hip_timing_enable(int on) {
    TimingState &T = t_timing;
    T.on = on != 0;
    T.used = 0;
    return UGEMM_OK;
}

int ugemm_hip_timing_collect(double *total_ms) {
    TimingState &T = t_timing;
    double sum = 0;
    int n = 0;
    for (size_t i = 0; i < T.used; ++i) {
        if (hipEventSynchronize(T.pairs[i].second) != hipSuccess) return fail(UGEMM_ERR_HIP, "timing: sync");
        float ms = 0;
        if (hipEventElapsedTime(&ms, T.pairs[i].first, T.pairs[i].second) != hipSuccess)
            return fail(UGEMM_ERR_HIP, "timing: elapsed");
        sum += ms;
        ++n;
    }
    T.used = 0;
    if (total_ms) *total_ms = sum;
    return n;
}

int ugemm_hip_variant_count(int mode) {
    switch (mode) {
        case UGEMM_MODE_FP32: return fp32_variant_count();
        case UGEMM_MODE_FP32_MFMA: return m32_variant_count();
        case UGEMM_MODE_FP8_BLOCK: return fp8_variant_count();
        case -1: return dgemm_variant_count();
        case UGEMM_MODE_BF16X1: case UGEMM_MODE_BF16X3: case UGEMM_MODE_BF16X6: case UGEMM_MODE_FP16X3:
            return mfma_variant_count(mode);
        default: return 0;
    }
}

const char *ugemm_hip_variant_name(int mode, int v) {
    switch (mode) {
        case UGEMM_MODE_FP32: return fp32_variant_name(v);
        case UGEMM_MODE_FP32_MFMA: return m32_variant_name(v);
        case UGEMM_MODE_FP8_BLOCK: return fp8_variant_name(v);
        case -1: return dgemm_variant_name(v);
        case UGEMM_MODE_BF16X1: case UGEMM_MODE_BF16X3: case UGEMM_MODE_BF16X6: case UGEMM_MODE_FP16X3:
            return mfma_variant_name(mode, v);
        default: return nullptr;
    }
}

}  // extern "C"
