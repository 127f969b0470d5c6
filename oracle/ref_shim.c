/* Builds the reference's header-only GEMM library (ugemm.h and the CPU
 * backends it includes: gemm_cpu.h, sgemm_avx256.h, sgemm_sse.h, dgemm_avx.h)
 * into a shared object, unchanged, so that the tests and the CPU-baseline leg
 * of bench.py can call sgemm_cpu, sgemm_c, sgemm_avx, sgemm_sse, dgemm_cpu,
 * _dgemm_c and dgemm_avx.  Test infrastructure only. */
#include <stdlib.h>
#include "ugemm.h"

/* sgemm_sse is file-static in sgemm_sse.h; give it an exported name. */
void ref_sgemm_sse(char major, char transa, char transb, int M, int N, int K, float alpha,
                   const float *A, int lda, const float *B, int ldb, float beta, float *C, int ldc)
{
	sgemm_sse(major, transa, transb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}
