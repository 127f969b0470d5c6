/*
 * check_sgemm_hip.c -- the uut drop-in, exercised from C the way the
 * reference's harness drives its CPU kernels (check_sgemm.c:87-143,
 * check_dgemm.c:86-142): a function pointer with the sgemm_cpu argument list
 * (ugemm.h:287-301) is called on nIter distinct inputs, each call timed, and
 * every result is compared with a reference product.
 *
 * Command line: key=value arguments as check_sgemm.c:145-237 takes them
 *   M= N= K= alpha= beta= lda= ldb= ldc=      (the reference's keys)
 * plus this harness's own
 *   major=R|C transa=N|T transb=N|T           layout (default R, N, N), or
 *   layout=RNN|RNT|RTN|...                    the same in one key
 *   mode=fp32|bf16x3|bf16x1|fp8|bf16x6|fp16x3|fp32_mfma   (default: sgemm_hip
 *                                             itself, whose form is fp32 unless
 *                                             UGEMM_HIP_MODE says otherwise)
 *   seed=<integer>                            SplitMix64 seed of the inputs
 *   iter=<integer>                            distinct inputs (default 11)
 * Defaults: M=128 N=361 K=1152 alpha=1 beta=0, ld = the stored row length.
 * A bad key value or an ld below the stored row length prints "Bad
 * parameter ..." and exits 1, as the reference does.
 *
 * Per iteration it prints the reference's cmp_results line (stdErr/stdRef,
 * the largest error and where, and its FAIL/Sucks verdict; check_sgemm.c:
 * 56-85) against the long-double product rounded to float, and decides
 * pass/fail with the elementwise bound
 *   |C - exact| <= (2(K+1) eps + extra(mode)) (|alpha||A||B| + |beta||C0|)
 * (extra: the split form's per-product representation error).  Then one
 * dgemm_hip call, one packed sgemm_ocl-style call (sgemm_hip_packed) and one
 * invalid call are checked.  Exit status 0 when every check passes.
 *
 * Built with gcc against include/ugemm_hip.h and libugemm_hip.so; nothing here
 * comes from the reference.
 */
#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <time.h>

#include "ugemm_hip.h"

typedef void (*sgemm_uut)(char, char, char, int, int, int, float, const float *, int, const float *, int,
                          float, float *, int);
typedef void (*dgemm_uut)(char, char, char, int, int, int, double, const double *, int, const double *, int,
                          double, double *, int);

/* SplitMix64: a named, seeded generator (the reference uses unseeded rand()). */
static uint64_t sm_state = 0x55474D4D00ULL;
static uint64_t splitmix64(void) {
    uint64_t z = (sm_state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double uniform_pm1(void) { return ((double)(splitmix64() >> 40) / 16777216.0) * 2.0 - 1.0; }

/* Element accessors of op(A), op(B) and C for both majors (ugemm.h:310-349). */
static long a_at(char major, char ta, int m, int k, int lda) {
    int t = !(ta == 'N' || ta == 'n');
    if (major == 'R') return t ? (long)k * lda + m : (long)m * lda + k;
    return t ? (long)m * lda + k : (long)k * lda + m;
}
static long b_at(char major, char tb, int k, int n, int ldb) {
    int t = !(tb == 'N' || tb == 'n');
    if (major == 'R') return t ? (long)n * ldb + k : (long)k * ldb + n;
    return t ? (long)k * ldb + n : (long)n * ldb + k;
}
static long c_at(char major, int m, int n, int ldc) { return major == 'R' ? (long)m * ldc + n : (long)n * ldc + m; }

/* Stored extents: elements of the array that holds an operand. */
static long stored_elems(char major, char trans, int rows, int cols, int ld) {
    int t = !(trans == 'N' || trans == 'n');
    int r = t ? cols : rows, c = t ? rows : cols;  /* op(X) is rows x cols; stored X is r x c */
    return major == 'R' ? (long)r * ld : (long)c * ld;
}
/* Length of a stored row (row-major) or column (column-major): the least ld. */
static int stored_len(char major, char trans, int rows, int cols) {
    int t = !(trans == 'N' || trans == 'n');
    return major == 'R' ? (t ? rows : cols) : (t ? cols : rows);
}

static double now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* The statistic of the reference's cmp_results (check_sgemm.c:56-85) over the
 * logical M x N result: RMS error, the spread of the reference values, the
 * largest error and where, and the verdict strings it prints. */
static void cmp_results_line(char major, int M, int N, const float *ref, const float *res, int ldc) {
    double max_err = 0, s2_err = 0, s1_ref = 0, s2_ref = 0;
    int mi = 0, ni = 0;
    for (int m = 0; m < M; ++m)
        for (int n = 0; n < N; ++n) {
            long c = c_at(major, m, n, ldc);
            double r = ref[c], x = res[c], e = x - r;
            if (max_err < fabs(e)) {
                max_err = fabs(e);
                mi = m;
                ni = n;
            }
            s2_err += e * e;
            s1_ref += r;
            s2_ref += r * r;
        }
    double mn = (double)M * N;
    double std_err = sqrt(s2_err / mn);
    double v = s2_ref * mn - s1_ref * s1_ref;
    double std_ref = sqrt(v > 0 ? v : 0) / mn;
    long c = c_at(major, mi, ni, ldc);
    printf("%.3e/%.3e=%.3e. %.3e at [%3d,%3d] %18.10e vs %18.10e %s\n", std_err, std_ref,
           std_ref > 0 ? std_err / std_ref : 0.0, max_err, mi, ni, (double)ref[c], (double)res[c],
           max_err > std_ref * 1e-5 ? "FAIL !!!" : (max_err > std_ref * 3e-5 || std_err > std_ref * 1e-6 ? "Sucks !" : ""));
}

/* Long-double reference product and error scale of one result; returns the
 * worst |res - exact| / (tol * scale) over C (> 1 is a violation) and writes
 * the reference rounded to T into ref (C's layout). */
#define CHECK_BODY(T, EPS)                                                                              \
    double worst = 0.0;                                                                                 \
    for (int m = 0; m < M; ++m)                                                                         \
        for (int n = 0; n < N; ++n) {                                                                   \
            long double s = 0.0L, sc = 0.0L;                                                            \
            for (int k = 0; k < K; ++k) {                                                               \
                long double p = (long double)A[a_at(major, ta, m, k, lda)] * B[b_at(major, tb, k, n, ldb)]; \
                s += p;                                                                                 \
                sc += fabsl(p);                                                                         \
            }                                                                                           \
            long c = c_at(major, m, n, ldc);                                                            \
            long double c0 = beta == 0 ? 0.0L : (long double)C0[c];                                     \
            long double exact = (long double)alpha * s + (long double)beta * c0;                        \
            long double scale = fabsl((long double)alpha) * sc + fabsl((long double)beta * c0);         \
            if (ref) ref[c] = (T)exact;                                                                 \
            double tol = (2.0 * (K + 1) * (EPS) + extra) * (double)scale;                               \
            double err = fabs((double)((long double)res[c] - exact));                                   \
            double r = tol > 0 ? err / tol : (err > 0 ? INFINITY : 0.0);                                \
            if (!(r <= worst)) worst = r;                                                               \
        }                                                                                               \
    return worst;

static double check_s(char major, char ta, char tb, int M, int N, int K, float alpha, const float *A, int lda,
                      const float *B, int ldb, float beta, const float *C0, const float *res, int ldc, double extra,
                      float *ref) {
    CHECK_BODY(float, ldexp(1.0, -24))
}
static double check_d(char major, char ta, char tb, int M, int N, int K, double alpha, const double *A, int lda,
                      const double *B, int ldb, double beta, const double *C0, const double *res, int ldc,
                      double extra, double *ref) {
    CHECK_BODY(double, ldexp(1.0, -53))
}

/* mode=: the form, and its extra per-product error term (relative to |a||b|). */
static const struct {
    const char *name;
    int mode;
    double extra;
} kModes[] = {
    {"fp32", UGEMM_MODE_FP32, 0.0},           {"fp32_mfma", UGEMM_MODE_FP32_MFMA, 0.0},
    {"bf16x3", UGEMM_MODE_BF16X3, 0x1p-14},   {"bf16x6", UGEMM_MODE_BF16X6, 0x1p-21},
    {"fp16x3", UGEMM_MODE_FP16X3, 0x1p-19},   {"bf16x1", UGEMM_MODE_BF16X1, 0x1p-6},
    {"fp8", UGEMM_MODE_FP8_BLOCK, 0x1p-3},    /* two e4m3 roundings of 2^-4 each */
};
static int g_mode = -1;
/* The uut for mode=: ugemm_hip_sgemm with the selected form, the uut argument list. */
static void sgemm_hip_mode(char major, char ta, char tb, int M, int N, int K, float alpha, const float *A, int lda,
                           const float *B, int ldb, float beta, float *C, int ldc) {
    ugemm_hip_sgemm(major, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, g_mode);
}

static int bad(const char *arg, const char *why) {
    fprintf(stderr, "Bad parameter '%s'. %s\n", arg, why);
    return 1;
}

int main(int argc, char **argv) {
    int M = 128, N = 361, K = 1152, lda = 0, ldb = 0, ldc = 0, nIter = 11;
    float alpha = 1.0f, beta = 0.0f;
    char major = 'R', ta = 'N', tb = 'N';
    double extra = 0.0;
    const char *mode_name = "sgemm_hip";

    for (int i = 1; i < argc; ++i) {
        const char *arg = argv[i], *eq = strchr(arg, '=');
        if (!eq) return bad(arg, "Expected key=value.");
        size_t klen = (size_t)(eq - arg);
        const char *val = eq + 1;
        char *endp;
#define KEY(s) (klen == strlen(s) && strncasecmp(arg, s, klen) == 0)
        if (KEY("alpha") || KEY("beta")) {
            double v = strtod(val, &endp);
            if (endp == val || *endp) return bad(arg, "Not a number.");
            if (KEY("alpha")) alpha = (float)v;
            else beta = (float)v;
        } else if (KEY("M") || KEY("N") || KEY("K") || KEY("lda") || KEY("ldb") || KEY("ldc") || KEY("iter")) {
            long v = strtol(val, &endp, 0);
            if (endp == val || *endp || v <= 0 || v > 0x7fffffffL) return bad(arg, "Not a positive number.");
            if (KEY("M")) M = (int)v;
            else if (KEY("N")) N = (int)v;
            else if (KEY("K")) K = (int)v;
            else if (KEY("lda")) lda = (int)v;
            else if (KEY("ldb")) ldb = (int)v;
            else if (KEY("ldc")) ldc = (int)v;
            else nIter = (int)v;
        } else if (KEY("seed")) {
            unsigned long long v = strtoull(val, &endp, 0);
            if (endp == val || *endp) return bad(arg, "Not an integer.");
            sm_state = (uint64_t)v;
        } else if (KEY("layout")) {
            /* RNN, RNT, RTN (sgemm_test.c's names), any of [RC][NT][NT], or [NT][NT] (row-major) */
            size_t n = strlen(val);
            const char *t = n == 3 ? val + 1 : val;
            if (!((n == 3 && (val[0] == 'R' || val[0] == 'C')) || n == 2) || (t[0] != 'N' && t[0] != 'T') ||
                (t[1] != 'N' && t[1] != 'T'))
                return bad(arg, "layout is [R|C][N|T][N|T].");
            major = n == 3 ? val[0] : 'R';
            ta = t[0];
            tb = t[1];
        } else if (KEY("major") || KEY("transa") || KEY("transb")) {
            char c = val[0];
            if (val[1] != 0) return bad(arg, "One letter expected.");
            if (KEY("major")) {
                if (c != 'R' && c != 'C') return bad(arg, "major is R or C.");
                major = c;
            } else {
                if (c != 'N' && c != 'T') return bad(arg, "trans is N or T.");
                if (KEY("transa")) ta = c;
                else tb = c;
            }
        } else if (KEY("mode")) {
            size_t j = 0;
            for (; j < sizeof kModes / sizeof kModes[0]; ++j)
                if (strcmp(val, kModes[j].name) == 0) break;
            if (j == sizeof kModes / sizeof kModes[0]) return bad(arg, "Unknown mode.");
            g_mode = kModes[j].mode;
            extra = kModes[j].extra;
            mode_name = kModes[j].name;
        } else {
            return bad(arg, "Unknown key.");
        }
#undef KEY
    }
    const int la = stored_len(major, ta, M, K), lb = stored_len(major, tb, K, N), lc = major == 'R' ? N : M;
    if (lda == 0) lda = la;
    if (ldb == 0) ldb = lb;
    if (ldc == 0) ldc = lc;
    if (lda < la) {
        fprintf(stderr, "Bad parameter lda=%d. Should be greater or equal to %d\n", lda, la);
        return 1;
    }
    if (ldb < lb) {
        fprintf(stderr, "Bad parameter ldb=%d. Should be greater or equal to %d\n", ldb, lb);
        return 1;
    }
    if (ldc < lc) {
        fprintf(stderr, "Bad parameter ldc=%d. Should be greater or equal to %d\n", ldc, lc);
        return 1;
    }
    if (g_mode < 0) {
        /* sgemm_hip's own form: UGEMM_HIP_MODE (fp32 when unset). */
        const char *e = getenv("UGEMM_HIP_MODE");
        for (size_t j = 0; e && j < sizeof kModes / sizeof kModes[0]; ++j)
            if (strcmp(e, kModes[j].name) == 0) extra = kModes[j].extra;
    }
    const int tA = !(ta == 'N' || ta == 'n'), tB = !(tb == 'N' || tb == 'n');
    const long na = stored_elems(major, ta, M, K, lda), nb = stored_elems(major, tb, K, N, ldb);
    const long nc = major == 'R' ? (long)M * ldc : (long)N * ldc;

    printf("%s\nRunning SGEMM with M=%d, N=%d, K=%d, alpha=%f, lda=%d, ldb=%d, beta=%f, ldc=%d, major=%c, "
           "transa=%c, transb=%c, uut=%s, seed=0x%llx, iter=%d\n",
           ugemm_hip_version(), M, N, K, alpha, lda, ldb, beta, ldc, major, ta, tb, mode_name,
           (unsigned long long)sm_state, nIter);
    fflush(stdout);
    if (ugemm_hip_init(-1, 0) != UGEMM_OK) {
        printf("ugemm_hip_init: %s\n", ugemm_hip_last_error_message());
        return 1;
    }
    int fail = 0;

    /* --- SGEMM through the uut function pointer, nIter distinct inputs --------- */
    {
        float *A = malloc(sizeof(float) * na * nIter), *B = malloc(sizeof(float) * nb * nIter);
        float *C = malloc(sizeof(float) * nc * nIter), *C0 = malloc(sizeof(float) * nc * nIter);
        float *ref = malloc(sizeof(float) * nc);
        double *ms = malloc(sizeof(double) * nIter);
        if (!A || !B || !C || !C0 || !ref || !ms) {
            printf("out of host memory\n");
            return 1;
        }
        for (long i = 0; i < na * nIter; ++i) A[i] = (float)uniform_pm1();
        for (long i = 0; i < nb * nIter; ++i) B[i] = (float)uniform_pm1();
        for (long i = 0; i < nc * nIter; ++i) C[i] = C0[i] = (float)uniform_pm1();
        sgemm_uut uut = g_mode < 0 ? sgemm_hip : sgemm_hip_mode;
        int st = UGEMM_OK;
        for (int it = 0; it < nIter; ++it) {
            double t0 = now();
            uut(major, ta, tb, M, N, K, alpha, A + it * na, lda, B + it * nb, ldb, beta, C + it * nc, ldc);
            ms[it] = 1e3 * (now() - t0);
            if (st == UGEMM_OK) st = ugemm_hip_last_error();
        }
        double mean = 0;
        for (int it = 0; it < nIter; ++it) {
            printf(" %.3f", ms[it]);
            mean += ms[it];
        }
        mean /= nIter;
        printf(": mean %.3f ms (host pointers, copies included). %.2f GFLOPS\n", mean, 2.0 * M * N * K / mean / 1e6);
        if (st != UGEMM_OK) {
            printf("status %d: %s\n", st, ugemm_hip_last_error_message());
            fail = 1;
        }
        for (int it = 0; it < nIter; ++it) {
            memcpy(ref, C0 + it * nc, sizeof(float) * nc);
            double w = check_s(major, ta, tb, M, N, K, alpha, A + it * na, lda, B + it * nb, ldb, beta, C0 + it * nc,
                               C + it * nc, ldc, extra, ref);
            cmp_results_line(major, M, N, ref, C + it * nc, ldc);
            printf("iter %d: worst error / bound = %.3g\n", it, w);
            if (!(w <= 1.0)) fail = 1;
        }
        /* Packed sgemm_ocl-style call: row-major, ld = row length. */
        if (major == 'R') {
            int pa = tA ? M : K, pb = tB ? K : N;
            float *a = malloc(sizeof(float) * (long)M * K), *b = malloc(sizeof(float) * (long)K * N);
            float *c = malloc(sizeof(float) * (long)M * N), *c0 = malloc(sizeof(float) * (long)M * N);
            for (long i = 0; i < (long)M * K; ++i) a[i] = (float)uniform_pm1();
            for (long i = 0; i < (long)K * N; ++i) b[i] = (float)uniform_pm1();
            for (long i = 0; i < (long)M * N; ++i) c[i] = c0[i] = (float)uniform_pm1();
            sgemm_hip_packed(ta, tb, M, N, K, alpha, a, b, beta, c);
            int pst = ugemm_hip_last_error();
            double w = check_s('R', ta, tb, M, N, K, alpha, a, pa, b, pb, beta, c0, c, N, extra, NULL);
            printf("sgemm_hip_packed: status %d, worst error / bound = %.3g\n", pst, w);
            if (pst != UGEMM_OK || !(w <= 1.0)) fail = 1;
            free(a); free(b); free(c); free(c0);
        }
        free(A); free(B); free(C); free(C0); free(ref); free(ms);
    }

    /* --- DGEMM through the uut function pointer --------------------------------- */
    {
        double *A = malloc(sizeof(double) * na), *B = malloc(sizeof(double) * nb);
        double *C = malloc(sizeof(double) * nc), *C0 = malloc(sizeof(double) * nc);
        for (long i = 0; i < na; ++i) A[i] = uniform_pm1();
        for (long i = 0; i < nb; ++i) B[i] = uniform_pm1();
        for (long i = 0; i < nc; ++i) C[i] = C0[i] = uniform_pm1();
        dgemm_uut uut = dgemm_hip;
        uut(major, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
        int st = ugemm_hip_last_error();
        double w = check_d(major, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C0, C, ldc, 0.0, NULL);
        printf("dgemm_hip: status %d, worst error / bound = %.3g\n", st, w);
        if (st != UGEMM_OK || !(w <= 1.0)) fail = 1;
        free(A); free(B); free(C); free(C0);
    }

    /* --- An invalid call reports through the status, and nothing is written ---- */
    {
        float x = 7.0f;
        sgemm_hip('R', 'N', 'N', 1, 1, 1, 1.0f, &x, 0, &x, 1, 0.0f, &x, 1); /* lda < K */
        int st = ugemm_hip_last_error();
        printf("invalid lda: status %d (%s)\n", st, ugemm_hip_status_string(st));
        if (st != UGEMM_ERR_ARG || x != 7.0f) fail = 1;
    }

    ugemm_hip_finish();
    printf(fail ? "FAIL\n" : "OK\n");
    return fail;
}
