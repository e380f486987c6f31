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
//
// LDS bank calibration (zdt_calib_lds, bench.py --lds-calibration): the cost
// of 8-byte LDS accesses in the row layouts and strides of the fused
// kernels' 512-point FFT passes (see the patterns below), timed over a
// full-device launch; rocprofv3's SQ_LDS_BANK_CONFLICT on the same command
// gives the conflict cycles per pattern.  The k2 row layout (q64 rows, the
// first pass's output at lsw) was chosen from these.
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

// LDS patterns: entry index of lane l (0..63) for access r (0..7) of a
// 512-entry row stored in a 576-entry slot; reads for patterns < 6, writes
// from 6 on.
constexpr int LDS_PATTERNS = 13;
__device__ __forceinline__ int cal_lsw(int i) { return i + (i >> 3); }
__device__ __forceinline__ int cal_q64(int i) { return i + ((i >> 6) << 3); }
__device__ __forceinline__ int cal_p16(int i) { return i + (i >> 4); }
__device__ __forceinline__ int cal_p32(int i) { return i + (i >> 5); }
template <int PAT>
__device__ __forceinline__ int lds_index(int l, int r) {
  const int c = l + 64 * r;                               // contiguous run of lanes
  const int s8 = 8 * l + r;                               // first FFT pass writes
  const int p2 = ((l >> 3) << 6) + (l & 7) + 8 * r;       // second FFT pass writes
  switch (PAT) {
    case 0: return c;            // read, unpadded
    case 1: return cal_lsw(c);   // read, one pad in 8
    case 2: return cal_p16(c);   // read, one pad in 16
    case 3: return cal_p32(c);   // read, one pad in 32
    case 4: return cal_q64(c);   // read, eight pads in 64
    case 5: return cal_lsw(s8);  // read, stride 8, one pad in 8
    case 6: return c;            // write, unpadded
    case 7: return cal_lsw(c);   // write, one pad in 8
    case 8: return cal_q64(c);   // write, eight pads in 64
    case 9: return cal_lsw(s8);  // write, first-pass stride 8, one pad in 8
    case 10: return cal_q64(s8); // write, first-pass stride 8, eight pads in 64
    case 11: return cal_q64(p2); // write, second pass, eight pads in 64
    default: return cal_lsw(p2); // write, second pass, one pad in 8
  }
}

template <int PAT>
__global__ __launch_bounds__(256) void lds_pattern(float* out, int iters) {
  constexpr bool WRITE = PAT >= 6;
  __shared__ c32 s[8 * 576];
  for (int i = threadIdx.x; i < 8 * 576; i += 256) s[i] = {static_cast<float>(i), 1.0f};
  __syncthreads();
  float acc = 0.0f;
  for (int it = 0; it < iters; ++it) {
    int l = threadIdx.x & 63;
    asm volatile("" : "+v"(l));
    const int row = ((it + (threadIdx.x >> 6)) & 7) * 576;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int i = row + lds_index<PAT>(l, r);
      if constexpr (WRITE) {
        s[i] = {acc, static_cast<float>(r)};
      } else {
        const c32 v = s[i];
        acc += v.x * v.y;
      }
    }
  }
  out[blockIdx.x * 256 + threadIdx.x] = acc + s[threadIdx.x].x;
}

template <int PAT>
void launch_lds(int pattern, float* out, int blocks, int iters, hipStream_t st) {
  if constexpr (PAT < LDS_PATTERNS) {
    if (pattern == PAT) {
      lds_pattern<PAT><<<blocks, 256, 0, st>>>(out, iters);
    } else {
      launch_lds<PAT + 1>(pattern, out, blocks, iters, st);
    }
  }
}

}  // namespace
}  // namespace zdt

extern "C" int zdt_calib_lds(int pattern, double* ns_per_access) {
  using namespace zdt;
  if (pattern < 0 || pattern >= LDS_PATTERNS)
    return fail(ZDT_EINVAL, "calib_lds: pattern %d not in 0..%d", pattern, LDS_PATTERNS - 1);
  if (ns_per_access == nullptr) return fail(ZDT_EINVAL, "calib_lds: NULL result");
  int dev = 0, ncu = 256;
  ZDT_TRY_HIP(hipGetDevice(&dev));
  ZDT_TRY_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int blocks = 4 * ncu, iters = 2048;  // 8 accesses per iteration and wavefront
  float* out = nullptr;
  ZDT_TRY_HIP(hipMalloc(&out, sizeof(float) * blocks * 256));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t err = hipEventCreate(&e0);
  if (err == hipSuccess) err = hipEventCreate(&e1);
  float best = 1e30f;
  for (int rep = 0; rep < 3 && err == hipSuccess; ++rep) {
    err = hipEventRecord(e0, nullptr);
    if (err != hipSuccess) break;
    launch_lds<0>(pattern, out, blocks, iters, nullptr);
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
  if (err != hipSuccess) return hip_fail(err, "calib_lds");
  // wavefront accesses per CU: blocks/ncu blocks of 4 wavefronts, iters x 8 each
  const double per_cu = static_cast<double>(blocks) / ncu * 4.0 * iters * 8.0;
  *ns_per_access = best * 1e6 / per_cu;
  return ZDT_OK;
}

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
