// FP8 block-scaled SGEMM (placeholder until the kernels land).
#include "internal.h"
namespace ugemm {
int fp8_variant_count() { return 0; }
const char *fp8_variant_name(int) { return nullptr; }
int launch_fp8(const Gemm<float> &, int, hipStream_t) { return fail(UGEMM_ERR_UNSUPPORTED, "fp8 not built yet"); }
}  // namespace ugemm
