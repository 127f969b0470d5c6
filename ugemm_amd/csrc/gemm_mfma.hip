// Split-BF16 SGEMM on MFMA (placeholder until the kernels land).
#include "internal.h"
namespace ugemm {
int mfma_variant_count(int) { return 0; }
const char *mfma_variant_name(int, int) { return nullptr; }
int mfma_pick_variant(const Gemm<float> &, int) { return 0; }
int launch_mfma(const Gemm<float> &, int mode, int, hipStream_t) {
    return fail(UGEMM_ERR_UNSUPPORTED, "mode %d not built yet", mode);
}
}  // namespace ugemm
