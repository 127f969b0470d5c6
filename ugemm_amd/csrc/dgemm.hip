// DGEMM on FP64 MFMA (placeholder until the kernels land).
#include "internal.h"
namespace ugemm {
int dgemm_variant_count() { return 0; }
const char *dgemm_variant_name(int) { return nullptr; }
int dgemm_pick_variant(const Gemm<double> &) { return 0; }
int launch_dgemm(const Gemm<double> &, int, hipStream_t) { return fail(UGEMM_ERR_UNSUPPORTED, "dgemm not built yet"); }
}  // namespace ugemm
