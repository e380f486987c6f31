// The fused paths' plan (device buffers and tables) and the host helpers the
// fused path (zdt_fast.hip) and the fused output (zdt_fout.hip) share.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "zdt_fast.h"
#include "zdt_internal.h"

namespace zdt {

// strided FFT length N2 of k1/k3: 1024 (32 x 32) when that leaves N1 = M/N2 >= 64,
// else 512 (32 x 16)
constexpr int fast_n2(int64_t Nf) { return Nf >= (int64_t(1) << 17) ? 1024 : 512; }

struct FastPlan {
  int64_t Nf = 0, Nt = 0, Nr = 0, M = 0;
  int N1 = 0, logN1 = 0, N2 = 0, logN2 = 0, NT = 0;
  float dr = 0;
  c32* G = nullptr;     // [Nt][N2][N1]
  c32* B = nullptr;     // [Nr][N1/16][N2][16]
  c32* tw1 = nullptr;   // W_N1^m
  c32* tw2 = nullptr;   // W_N2^m
  c32* tabq = nullptr;  // e^{i pi dr j}
  c64* tab64 = nullptr;  // e^{i pi dr j} in FP64
  c32* yhalf = nullptr;  // [Nr] bin M/2 term of the c2r (k2 -> k3)
  c32* Yc = nullptr;     // fused output only: [Nr][N2][N1] spectra
  c32* yM = nullptr;     // fused output only: [Nr] Nyquist bin
  double* part = nullptr;  // k3's per-wavefront FDR statistics (fast_stats_doubles)
  bool paired = false;     // the last k2 wrote B in mirrored rate pairs
  bool have_B = false;
  int64_t resident_blocks = 512;  // persistent k1/k3 grid: 2 blocks per CU
  int precision = 0;              // ZDT_PRECISION_* of the DFT stages
  // set by a caller around one run: k3 then also writes the FDR's run
  // maxima there (Nr * Nf/FDR_RUN floats, see K3Args::smax)
  float* smax = nullptr;
};

int ilog2i(int64_t n);
// device copy of a host table
int upload_table(c32** dst, const std::vector<c32>& h);
// W_N^m, m < N
std::vector<c32> twiddles(int N);
// allocates p->part (fast_stats_doubles) unless rc already holds an error
void alloc_part(FastPlan* p, int& rc);
// k1_input over Nt time steps of spec into G
int launch_k1(const FastPlan* p, const float* spec, c32* G, int64_t Nt, hipStream_t st);

// split-precision (matrix-core) k1/k3 for p->precision != ZDT_PRECISION_FP32
// (zdt_mfma.hip); they need N2 = 1024
struct K1Args;
struct K3Args;
bool mfma_supported(const FastPlan* p);
int launch_k1_mfma(const FastPlan* p, const K1Args& a, hipStream_t st);
int launch_k3_mfma(const FastPlan* p, const K3Args& a, hipStream_t st);

}  // namespace zdt
