// FP32 VALU issue-rate calibration for gfx950 (a standalone program, not part
// of libzdt_hip.so).  Each thread runs eight independent chains of one
// instruction form over 16384 blocks of 256 threads; the rate is lane
// operations per second (an FMA is one lane operation, two flops).  The forms:
//   vop3_sgpr  a = a*b + c with b, c uniform (v_fma_f32 with SGPR operands)
//   vop3_vgpr  a = a*b + c with b, c per lane (v_fma_f32, VGPR operands)
//   vop2_fmac  a = b*c + a with b, c per lane (v_fmac_f32_e32)
//   vop2_add   a = a + b (v_add_f32_e32)
// Built with the same flags as the library (packed FP32 disabled):
//   hipcc --offload-arch=gfx950 -O3 -Xclang -target-feature -Xclang -packed-fp32-ops valu_rate.hip -o valu_rate
#include <hip/hip_runtime.h>

#include <cstdio>

template <int K>
__global__ __launch_bounds__(256) void chains(float* out, float b0, float c0, int iters) {
  const float b = (K == 0) ? b0 : b0 + threadIdx.x * 1e-9f;
  const float c = (K == 0) ? c0 : c0 - threadIdx.x * 1e-9f;
  float a0 = threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, a4 = a0 + 4, a5 = a0 + 5, a6 = a0 + 6, a7 = a0 + 7;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      if constexpr (K <= 1) {
        a0 = fmaf(a0, b, c); a1 = fmaf(a1, b, c); a2 = fmaf(a2, b, c); a3 = fmaf(a3, b, c);
        a4 = fmaf(a4, b, c); a5 = fmaf(a5, b, c); a6 = fmaf(a6, b, c); a7 = fmaf(a7, b, c);
      } else if constexpr (K == 2) {
        a0 = fmaf(b, c, a0); a1 = fmaf(b, c, a1); a2 = fmaf(b, c, a2); a3 = fmaf(b, c, a3);
        a4 = fmaf(b, c, a4); a5 = fmaf(b, c, a5); a6 = fmaf(b, c, a6); a7 = fmaf(b, c, a7);
      } else {
        a0 += b; a1 += c; a2 += b; a3 += c; a4 += b; a5 += c; a6 += b; a7 += c;
      }
    }
  }
  out[blockIdx.x * 256 + threadIdx.x] = a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7;
}

int main() {
  const int blocks = 16384, iters = 2000;
  float* out = nullptr;
  if (hipMalloc(&out, sizeof(float) * blocks * 256) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
  const double ops = double(blocks) * 256 * iters * 16 * 8;
  const char* names[4] = {"vop3_sgpr", "vop3_vgpr", "vop2_fmac", "vop2_add"};
  for (int k = 0; k < 4; ++k) {
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      (void)hipEventRecord(e0);
      if (k == 0) chains<0><<<blocks, 256>>>(out, 0.999f, 0.001f, iters);
      if (k == 1) chains<1><<<blocks, 256>>>(out, 0.999f, 0.001f, iters);
      if (k == 2) chains<2><<<blocks, 256>>>(out, 0.999f, 0.001f, iters);
      if (k == 3) chains<3><<<blocks, 256>>>(out, 0.999f, 0.001f, iters);
      (void)hipEventRecord(e1);
      if (hipEventSynchronize(e1) != hipSuccess) return 1;
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) best = ms;
    }
    printf("%-10s %8.3f ms  %6.1f T lane-ops/s  (%.1f TFLOP/s if FMA)\n", names[k], best, ops / best / 1e9,
           2 * ops / best / 1e9);
  }
  (void)hipFree(out);
  return 0;
}
