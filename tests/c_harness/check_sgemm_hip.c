/*
 * check_sgemm_hip.c -- the uut drop-in, exercised from C exactly as the
 * reference's harness drives its CPU kernels (check_sgemm.c:87-143,
 * check_dgemm.c:86-142): a function pointer with the sgemm_cpu argument list
 * (ugemm.h:287-301) is called on nIter distinct inputs, then every result is
 * compared with a naive loop.
 *
 * Built with gcc against include/ugemm_hip.h and libugemm_hip.so; nothing here
 * comes from the reference.  Usage:
 *   check_sgemm_hip [M N K [alpha beta [major ta tb]]]
 * Exit status 0 when every iteration of sgemm_hip and dgemm_hip is within
 * 2*(K+1)*eps*(|alpha||A||B| + |beta||C0|) elementwise of a long-double loop
 * and the packed sgemm_ocl-style call agrees with it too; 1 otherwise.
 */
#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
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

/* Stored extents: rows x ld of the array that holds an operand. */
static long stored_elems(char major, char trans, int rows, int cols, int ld) {
    int t = !(trans == 'N' || trans == 'n');
    int r = t ? cols : rows, c = t ? rows : cols;  /* op(X) is rows x cols; stored X is r x c */
    return major == 'R' ? (long)r * ld : (long)c * ld;
}

static double now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* Worst |res - exact| / (tol * scale) over C; > 1 means a violation. */
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
            double tol = 2.0 * (K + 1) * (EPS) * (double)scale;                                         \
            double err = fabs((double)((long double)res[c] - exact));                                   \
            double r = tol > 0 ? err / tol : (err > 0 ? INFINITY : 0.0);                                \
            if (!(r <= worst)) worst = r;                                                               \
        }                                                                                               \
    return worst;

static double check_s(char major, char ta, char tb, int M, int N, int K, float alpha, const float *A, int lda,
                      const float *B, int ldb, float beta, const float *C0, const float *res, int ldc) {
    CHECK_BODY(float, ldexp(1.0, -24))
}
static double check_d(char major, char ta, char tb, int M, int N, int K, double alpha, const double *A, int lda,
                      const double *B, int ldb, double beta, const double *C0, const double *res, int ldc) {
    CHECK_BODY(double, ldexp(1.0, -53))
}

int main(int argc, char **argv) {
    int M = argc > 3 ? atoi(argv[1]) : 128, N = argc > 3 ? atoi(argv[2]) : 361, K = argc > 3 ? atoi(argv[3]) : 1152;
    float alpha = argc > 5 ? (float)atof(argv[4]) : 1.5f, beta = argc > 5 ? (float)atof(argv[5]) : 0.5f;
    char major = argc > 8 ? argv[6][0] : 'R', ta = argc > 8 ? argv[7][0] : 'N', tb = argc > 8 ? argv[8][0] : 'N';
    const int nIter = 3;
    int tA = !(ta == 'N' || ta == 'n'), tB = !(tb == 'N' || tb == 'n');
    /* Leading dimensions with a little padding, as a caller's sub-matrix would have. */
    int lda = (major == 'R' ? (tA ? M : K) : (tA ? K : M)) + 3;
    int ldb = (major == 'R' ? (tB ? K : N) : (tB ? N : K)) + 5;
    int ldc = (major == 'R' ? N : M) + 2;
    long na = stored_elems(major, ta, M, K, lda), nb = stored_elems(major, tb, K, N, ldb);
    long nc = major == 'R' ? (long)M * ldc : (long)N * ldc;

    printf("%s  M=%d N=%d K=%d alpha=%g beta=%g major=%c %c%c\n", ugemm_hip_version(), M, N, K, alpha, beta,
           major, ta, tb);
    if (ugemm_hip_init(-1, 0) != UGEMM_OK) {
        printf("ugemm_hip_init: %s\n", ugemm_hip_last_error_message());
        return 1;
    }
    int fail = 0;

    /* --- SGEMM through the uut function pointer --------------------------------- */
    {
        float *A = malloc(sizeof(float) * na * nIter), *B = malloc(sizeof(float) * nb * nIter);
        float *C = malloc(sizeof(float) * nc * nIter), *C0 = malloc(sizeof(float) * nc * nIter);
        for (long i = 0; i < na * nIter; ++i) A[i] = (float)uniform_pm1();
        for (long i = 0; i < nb * nIter; ++i) B[i] = (float)uniform_pm1();
        for (long i = 0; i < nc * nIter; ++i) C[i] = C0[i] = (float)uniform_pm1();
        sgemm_uut uut = sgemm_hip;
        for (int it = 0; it < nIter; ++it) {
            double t0 = now();
            uut(major, ta, tb, M, N, K, alpha, A + it * na, lda, B + it * nb, ldb, beta, C + it * nc, ldc);
            double t1 = now();
            int st = ugemm_hip_last_error();
            double w = check_s(major, ta, tb, M, N, K, alpha, A + it * na, lda, B + it * nb, ldb, beta, C0 + it * nc,
                               C + it * nc, ldc);
            printf("sgemm_hip  iter %d: status %d, %.3f ms (host pointers), worst error / bound = %.3g\n", it, st,
                   1e3 * (t1 - t0), w);
            if (st != UGEMM_OK || !(w <= 1.0)) fail = 1;
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
            int st = ugemm_hip_last_error();
            double w = check_s('R', ta, tb, M, N, K, alpha, a, pa, b, pb, beta, c0, c, N);
            printf("sgemm_hip_packed: status %d, worst error / bound = %.3g\n", st, w);
            if (st != UGEMM_OK || !(w <= 1.0)) fail = 1;
            free(a); free(b); free(c); free(c0);
        }
        free(A); free(B); free(C); free(C0);
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
        double w = check_d(major, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C0, C, ldc);
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
