// FP32 VALU issue-rate calibration (zdt_calib_valu, used by
// bench.py --valu-calibration): the rate the compute roofline of the bench
// line quotes, and how it depends on the number of independent dependency
// chains per lane.  Each thread runs NCH independent chains of one
// instruction form over 16384 blocks of 256 threads; the rate is lane
// operations per second (an FMA is one lane operation, two flops).  Forms:
//   0 vop3_sgpr  a = a*b + c with b, c uniform (v_fma_f32 with SGPR operands)
//   1 vop3_vgpr  a = a*b + c with b, c per lane (v_fma_f32, VGPR operands)
//   2 vop2_fmac  a = b*c + a with b, c per lane (v_fmac_f32_e32)
//   3 vop2_add   a = a + b (v_add_f32_e32)
// The library's chirp-Z recurrences are VOP2 chains, two or four per lane.
#include <hip/hip_runtime.h>

#include "zdt_internal.h"

namespace zdt {
namespace {

template <int K, int NCH>
__global__ __launch_bounds__(256) void valu_chains(float* out, float b0, float c0, int iters) {
  const float b = (K == 0) ? b0 : b0 + threadIdx.x * 1e-9f;
  const float c = (K == 0) ? c0 : c0 - threadIdx.x * 1e-9f;
  float a[NCH];
#pragma unroll
  for (int q = 0; q < NCH; ++q) a[q] = threadIdx.x + q;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < 128 / NCH; ++u) {
#pragma unroll
      for (int q = 0; q < NCH; ++q) {
        if constexpr (K <= 1) {
          a[q] = fmaf(a[q], b, c);
        } else if constexpr (K == 2) {
          a[q] = fmaf(b, c, a[q]);
        } else {
          a[q] += (q & 1) ? c : b;
        }
      }
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int q = 0; q < NCH; ++q) s += a[q];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

template <int K>
void launch_form(int nch, float* out, int blocks, int iters, hipStream_t st) {
  switch (nch) {
    case 1: valu_chains<K, 1><<<blocks, 256, 0, st>>>(out, 0.999f, 0.001f, iters); break;
    case 2: valu_chains<K, 2><<<blocks, 256, 0, st>>>(out, 0.999f, 0.001f, iters); break;
    case 4: valu_chains<K, 4><<<blocks, 256, 0, st>>>(out, 0.999f, 0.001f, iters); break;
    default: valu_chains<K, 8><<<blocks, 256, 0, st>>>(out, 0.999f, 0.001f, iters); break;
  }
}

}  // namespace
}  // namespace zdt

extern "C" int zdt_calib_valu(int form, int chains, double* lane_ops_per_s) {
  using namespace zdt;
  if (form < 0 || form > 3) return fail(ZDT_EINVAL, "calib_valu: form %d not in 0..3", form);
  if (chains != 1 && chains != 2 && chains != 4 && chains != 8)
    return fail(ZDT_EINVAL, "calib_valu: chains must be 1, 2, 4 or 8");
  if (lane_ops_per_s == nullptr) return fail(ZDT_EINVAL, "calib_valu: NULL result");
  constexpr int blocks = 16384, iters = 250;  // 128 operations per iteration and lane
  float* out = nullptr;
  ZDT_TRY_HIP(hipMalloc(&out, sizeof(float) * blocks * 256));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t err = hipEventCreate(&e0);
  if (err == hipSuccess) err = hipEventCreate(&e1);
  float best = 1e30f;
  for (int rep = 0; rep < 3 && err == hipSuccess; ++rep) {
    err = hipEventRecord(e0, nullptr);
    if (err != hipSuccess) break;
    if (form == 0) launch_form<0>(chains, out, blocks, iters, nullptr);
    if (form == 1) launch_form<1>(chains, out, blocks, iters, nullptr);
    if (form == 2) launch_form<2>(chains, out, blocks, iters, nullptr);
    if (form == 3) launch_form<3>(chains, out, blocks, iters, nullptr);
    err = hipGetLastError();
    if (err == hipSuccess) err = hipEventRecord(e1, nullptr);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    if (err == hipSuccess && ms < best) best = ms;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(out);
  if (err != hipSuccess) return hip_fail(err, "calib_valu");
  *lane_ops_per_s = static_cast<double>(blocks) * 256 * iters * 128 / (best * 1e-3);
  return ZDT_OK;
}
