// Generic batched mixed-radix FFT engine (see fft_generic.hip).
#pragma once

#include <vector>

#include "zdt_internal.h"

namespace zdt {

struct FftPlan {
  int64_t N = 0;
  bool dbl = false;
  std::vector<int> radices;  // passes, in order
  void* tw = nullptr;        // device table W_N^m, m < N (c32 or c64)
  // four-step split N = N1 * N2 of the FP32 contiguous transforms (N1 = 0: none):
  // N1-point column DFTs and N2-point row DFTs, each in LDS, TC columns / TR rows
  // per block; twiddles W_N^m = thi[m >> 10] * tlo[m & 1023]
  int N1 = 0, N2 = 0, TC = 0, TR = 0;
  std::vector<int> rad1, rad2;
  c32* tlo = nullptr;
  c32* thi = nullptr;
};

// Factor N into the supported radices, largest power-of-two radix first and
// none above maxr (16, 8 or 4); false if a prime factor is unsupported.
bool fft_factorize(int64_t N, std::vector<int>* radices, int maxr = 16);

// Cached plan (per device, length, precision and radix cap).  Plans live
// until exit.
int fft_plan_get(int64_t N, bool dbl, const FftPlan** out, int maxr = 16);

// Batched complex transform.  Element n of batch b of `in` is at
// in[b*ibs + n*ies] (complex elements); same for out.  s0/s1 are scratch
// buffers of N*batch complex elements each (of the plan's precision), distinct
// from in/out.  in == out is allowed.  The result is multiplied by `scale`.
int fft_c2c(const FftPlan* p, const void* in, int64_t ies, int64_t ibs, void* out, int64_t oes,
            int64_t obs, int64_t batch, bool inverse, double scale, void* s0, void* s1,
            hipStream_t st);

// Complex elements needed in each of the three scratch buffers of fft_r2c/fft_c2r
// (N/2*batch for even N, N*batch for odd N).
size_t fft_real_scratch_elems(int64_t N, int64_t batch);

// Real-to-complex along the contiguous axis: in is N x batch (column stride
// ldi floats), out is (N/2+1) x batch (column stride ldo complex).
int fft_r2c(int64_t N, const float* in, int64_t ldi, c32* out, int64_t ldo, int64_t batch, void* s0,
            void* s1, void* s2, hipStream_t st, int maxr = 16);

// Complex-to-real inverse: the normalized inverse (1/N) times `scale`; the imaginary
// parts of the DC and (even N) Nyquist bins are ignored.
int fft_c2r(int64_t N, const c32* in, int64_t ldi, float* out, int64_t ldo, int64_t batch, double scale,
            void* s0, void* s1, void* s2, hipStream_t st, int maxr = 16);

}  // namespace zdt
