// Sustained matrix-core rate of the device: a measurement the roofline of
// the split forms is read against.  The published dense peak (2.5 PFLOP/s
// for bf16 and fp16) assumes the top clock; with every CU issuing MFMAs the
// part runs into its power limit and the clock falls, so the rate that any
// GEMM core can reach is lower and depends on the operands' bit patterns
// (the multipliers' switching).  This kernel issues nothing but independent
// v_mfma_f32_16x16x32 instructions on operands taken from a caller buffer
// (real planes, so that the switching matches), on every SIMD, and reports
// the rate.  No counterpart in the reference.
#include <algorithm>

#include "internal.h"
#include "kernel_common.h"

namespace ugemm {
namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float float4v __attribute__((ext_vector_type(4)));

constexpr int kFrags = 1024;   // 16-byte operand fragments in the caller's buffer
constexpr int kAcc = 16;       // independent accumulators per wave

// Each wave holds 8 A and 8 B fragments (lane-dependent picks from ops) and
// runs kAcc MFMAs per round, accumulator j taking A fragment j % 8 and B
// fragment (j / 2) % 8, so consecutive MFMAs never depend on each other.  The
// accumulators are summed into out (one float per thread) so that nothing is
// dead code.
template <bool F16>
__global__ __launch_bounds__(256) void mfma_ceiling_kernel(const uint4 *__restrict__ ops, float *__restrict__ out,
                                                           int rounds) {
    const int lane = threadIdx.x & 63;
    uint4 a[8], b[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        a[i] = ops[(lane + 64 * i) & (kFrags - 1)];
        b[i] = ops[(lane * 3 + 64 * i + kFrags / 2) & (kFrags - 1)];
    }
    float4v acc[kAcc];
#pragma unroll
    for (int j = 0; j < kAcc; ++j) acc[j] = float4v{0.0f, 0.0f, 0.0f, 0.0f};
    for (int it = 0; it < rounds; ++it) {
#pragma unroll
        for (int j = 0; j < kAcc; ++j) {
            if constexpr (F16)
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(*reinterpret_cast<const half8 *>(&a[j & 7]),
                                                               *reinterpret_cast<const half8 *>(&b[(j >> 1) & 7]),
                                                               acc[j], 0, 0, 0);
            else
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<const bf8 *>(&a[j & 7]),
                                                                *reinterpret_cast<const bf8 *>(&b[(j >> 1) & 7]),
                                                                acc[j], 0, 0, 0);
        }
    }
    float sum = 0.0f;
#pragma unroll
    for (int j = 0; j < kAcc; ++j) sum += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
    out[(long)blockIdx.x * blockDim.x + threadIdx.x] = sum;
}

}  // namespace

// Eight blocks of four waves per CU (eight waves per SIMD), one untimed
// launch of rounds / 4, then one timed launch; *tflops = flop / time.
int mfma_ceiling(const void *ops, int fp16, int rounds, double *tflops) {
    if (!ops || !tflops || rounds <= 0) return fail(UGEMM_ERR_ARG, "mfma ceiling: bad argument");
    const unsigned blocks = (unsigned)device_cu_count() * 8;
    float *out = nullptr;
    UGEMM_HIP_TRY(hipMalloc((void **)&out, (size_t)blocks * 256 * sizeof(float)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    auto kern = fp16 ? mfma_ceiling_kernel<true> : mfma_ceiling_kernel<false>;
    float ms = 0.0f;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, 0, (const uint4 *)ops, out, std::max(1, rounds / 4));
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e0, 0);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, 0, (const uint4 *)ops, out, rounds);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, 0);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (e != hipSuccess) return hip_fail(e, "mfma ceiling");
    const double flop = 2.0 * 16 * 16 * 32 * kAcc * (double)rounds * blocks * 4;
    *tflops = ms > 0.0f ? flop / (ms * 1e-3) / 1e12 : 0.0;
    return UGEMM_OK;
}

}  // namespace ugemm
