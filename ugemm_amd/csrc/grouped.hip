// Grouped persistent SGEMM (placeholder until the kernels land).
#include "internal.h"
using namespace ugemm;
extern "C" {
int ugemm_hip_grouped_plan_create(ugemm_grouped_plan_t **plan, int, const ugemm_problem_t *, int, int, void *) {
    if (plan) *plan = nullptr;
    return fail(UGEMM_ERR_UNSUPPORTED, "grouped not built yet");
}
int ugemm_hip_grouped_run(ugemm_grouped_plan_t *, void *) { return fail(UGEMM_ERR_UNSUPPORTED, "grouped not built yet"); }
void ugemm_hip_grouped_plan_destroy(ugemm_grouped_plan_t *) {}
int ugemm_hip_sgemm_grouped(int, const ugemm_problem_t *, int, void *) { return fail(UGEMM_ERR_UNSUPPORTED, "grouped not built yet"); }
int ugemm_hip_tune_grouped(ugemm_grouped_plan_t *) { return fail(UGEMM_ERR_UNSUPPORTED, "grouped not built yet"); }
}
