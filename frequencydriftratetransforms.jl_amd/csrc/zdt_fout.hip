// Fused output for power-of-two Nf in [2^16, 2^22] and Nt <= FOUT_MAX_NT,
// where the chirp-Z coefficients of a bin do not fit k2's registers (Nt > 16)
// or the rows do not fit its LDS (Nf > 2^20):
//   k1_input + k2_input   input!: the four-step r2c into Ft[t][k2][k1]
//   czt_reinsch_col       the chirp-Z per bin -> Yc[j][k2][k1]
//   k2_c2r / k2_c2r_wave  c2r pre-pass and inverse DFT_N1 -> B (as k2)
//   k3_output             as in the fused path (fast_output)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "zdt_fast_dev.h"
#include "zdt_fast_plan.h"

namespace zdt {

// ---------------------------------------------------------------- fused output for Nt > 16
//
// For power-of-two Nf with Nt up to FOUT_MAX_NT (more coefficients than k2
// keeps in registers; above 64 the chirp-Z runs in 64-step chunks) the
// chirp-Z runs per bin (czt_reinsch_col) and the output irfft reuses the
// second half of k2 and k3:
//   czt_reinsch_col  Ys[k, j] -> Yc[j][k2][k1] (k = k2 + N2 k1 < M), Ys[M, j] -> yM[j]
//   k2_c2r           per Hermitian column pair: c2r pre-pass, inverse DFT_N1
//                    over k1 (one wavefront per row) -> B, bin M/2 -> yhalf
//   k3_output        as in the fused path
// which moves the spectra through HBM twice instead of once per FFT pass.

struct CztColArgs {
  const c32* F;  // Nk x Nt, natural layout (column stride ldF), or with ftrans
                 // [Nt][N2][N1] (bin k2 + N2 k1 at [t][k2][k1]) then [Nt] bin M
  int64_t ldF;
  int ftrans;
  c32* Yc;       // [Nr][N2][N1]
  c32* yM;       // [Nr]
  int64_t M, Nf;
  int logN1, logN2;
  int Nt, Nr;
  double hd2;    // dr/Nf
  int wkind;
  const float* wcustom;
  float scale;
  double hc;     // czt_reinsch_col: rc/Nf, the coefficients' pre-rotation (centre rate rc)
  double hz;     // czt_reinsch_col: (r0 - rc)/Nf
};

// The bin of a thread: one thread per bin; a block covers 16 columns k2 x 16
// rows k1 (so the Nr stores of a wavefront are four 128-byte rows of Yc); the
// last block's thread 0 takes the Nyquist bin M.  Returns false for the idle
// threads of the last block.
struct ColBin {
  int64_t k;
  c32* dst;
  int64_t stride;  // between rates in dst
  int64_t fo, fs;  // F offset of the bin and its step in t
};

__device__ __forceinline__ bool col_bin(const CztColArgs& a, ColBin& b) {
  const int ntk1 = (1 << a.logN1) >> 4;
  const int tile = blockIdx.x;
  if (tile * 256 >= a.M) {
    if (threadIdx.x != 0) return false;
    b.k = a.M;
    b.dst = a.yM;
    b.stride = 1;
    b.fo = a.ftrans ? a.M * a.Nt : a.M;
    b.fs = a.ftrans ? 1 : a.ldF;
  } else {
    const int k2 = ((tile / ntk1) << 4) + (threadIdx.x >> 4);
    const int k1 = ((tile % ntk1) << 4) + (threadIdx.x & 15);
    b.k = k2 + (static_cast<int64_t>(k1) << a.logN2);
    b.dst = a.Yc + (static_cast<int64_t>(k2) << a.logN1) + k1;
    b.stride = a.M;
    b.fo = a.ftrans ? (static_cast<int64_t>(k2) << a.logN1) + k1 : b.k;
    b.fs = a.ftrans ? a.M : a.ldF;
  }
  return true;
}

// The chirp-Z of a bin for the rates in mirrored pairs: with the centre rate
// rc = r0 + (Nr-1) dr/2 folded into the coefficients (c_t e^{2 pi i k t rc/Nf}),
// rates j and Nr-1-j are the sums at z = e^{i theta} and conj(z),
// theta = 2 pi k (r_j - rc)/Nf, and come from one Reinsch-modified Goertzel
// recurrence (C. Reinsch, 1967; J. Oliver, "An error analysis of the modified
// Clenshaw method", 1977) over the real and the imaginary parts of the
// coefficients.  With s = sign(cos theta) and lambda = 2 (cos theta - s),
//   D_t = a_t + lambda U_{t+1} + s D_{t+1},  U_t = D_t + s U_{t+1}  (t = Nt-1 .. 1),
//   sum a_t cos(t theta) = a_0 + (lambda/2) U_1 + s D_1,  sum a_t sin(t theta) = U_1 sin theta:
// three FMAs per term and part, six per term for both rates against eight
// for two complex Horner evaluations.  lambda comes from the FP64 cos theta,
// so no accuracy is lost near theta = 0 or pi (where Goertzel's rounded
// 2 cos theta would lose it).  The middle rate of an odd Nr is its own
// partner.
template <int NT>
__global__ __launch_bounds__(256) void czt_reinsch_col(CztColArgs a) {
  ColBin bin;
  if (!col_bin(a, bin)) return;
  const double kd = static_cast<double>(bin.k);
  const float w = fast_window(a.wkind, a.wcustom, bin.k, a.Nf) * a.scale;
  float cr[NT], ci[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    if (t < a.Nt) {
      c32 f = a.F[bin.fo + static_cast<int64_t>(t) * bin.fs];
      if (a.hc != 0.0) f = cmul(f, cis_nh(kd * static_cast<double>(t), a.hc));  // uniform
      cr[t] = f.x * w;
      ci[t] = f.y * w;
    } else {
      cr[t] = 0.0f;
      ci[t] = 0.0f;
    }
  }
  const c64 g = cis64f(kd * a.hd2);
  c64 z = cis64f(kd * a.hz);  // e^{2 pi i k (r_0 - rc)/Nf}
  const int npair = (a.Nr + 1) / 2;  // j < npair; partner Nr-1-j
  // two rate pairs (j, j+1) per iteration: four independent recurrences
  int j = 0;
  for (; j + 1 < npair; j += 2) {
    const c64 z1 = cmul(z, g);
    c32 y0, y0c, y1, y1c;
    reinsch_quad<NT>([&](int t) { return cr[t]; }, [&](int t) { return ci[t]; }, z, [&](int t) { return cr[t]; },
                     [&](int t) { return ci[t]; }, z1, y0, y0c, y1, y1c);
    const int jp = a.Nr - 1 - j;  // > j + 1: both partners are distinct
    bin.dst[static_cast<int64_t>(j) * bin.stride] = y0;
    bin.dst[static_cast<int64_t>(jp) * bin.stride] = y0c;
    bin.dst[static_cast<int64_t>(j + 1) * bin.stride] = y1;
    if (jp - 1 != j + 1) bin.dst[static_cast<int64_t>(jp - 1) * bin.stride] = y1c;
    z = cmul(z1, g);
  }
  if (j < npair) {
    const int jp = a.Nr - 1 - j;
    c32 yz, yc;
    reinsch_split<NT>(cr, ci, z, yz, yc);
    bin.dst[static_cast<int64_t>(j) * bin.stride] = yz;
    if (jp != j) bin.dst[static_cast<int64_t>(jp) * bin.stride] = yc;
  }
}

// Nt > 64: the sum in chunks of 64 time steps, sum_t c_t z^t =
// sum_q z^{64q} P_q(z), with NC (a power of two >= ceil(Nt/64)) adjacent
// lanes per bin: lane q evaluates P_q for every rate pair (the recurrence of
// czt_reinsch_col, two pairs at a time), scales it by z^{64q} (conj for the
// partner rate; FP64 recurrence over the rates) and the NC partial sums are
// added across the lanes (butterfly shuffles), so each Yc entry is written
// once.  Bins are numbered g = 0 .. M (g = M: the Nyquist bin) in col_bin's
// tile order.
template <int NC>
__global__ __launch_bounds__(256) void czt_reinsch_col_split(CztColArgs a) {
  constexpr int NT = 64;
  constexpr int BPB = 256 / NC;  // bins per block
  const int q = threadIdx.x % NC;
  const int64_t g = static_cast<int64_t>(blockIdx.x) * BPB + threadIdx.x / NC;
  if (g > a.M) return;  // whole groups of NC lanes leave together
  ColBin bin;
  if (g == a.M) {
    bin.k = a.M;
    bin.dst = a.yM;
    bin.stride = 1;
    bin.fo = a.ftrans ? a.M * a.Nt : a.M;
    bin.fs = a.ftrans ? 1 : a.ldF;
  } else {
    const int ntk1 = (1 << a.logN1) >> 4;
    const int64_t tile = g >> 8, w = g & 255;
    const int k2 = static_cast<int>(((tile / ntk1) << 4) + (w >> 4));
    const int k1 = static_cast<int>(((tile % ntk1) << 4) + (w & 15));
    bin.k = k2 + (static_cast<int64_t>(k1) << a.logN2);
    bin.dst = a.Yc + (static_cast<int64_t>(k2) << a.logN1) + k1;
    bin.stride = a.M;
    bin.fo = a.ftrans ? (static_cast<int64_t>(k2) << a.logN1) + k1 : bin.k;
    bin.fs = a.ftrans ? a.M : a.ldF;
  }
  const double kd = static_cast<double>(bin.k);
  const float w = fast_window(a.wkind, a.wcustom, bin.k, a.Nf) * a.scale;
  const int t0 = NT * q;
  float cr[NT], ci[NT];
#pragma unroll
  for (int u = 0; u < NT; ++u) {
    const int t = t0 + u;
    if (t < a.Nt) {
      c32 f = a.F[bin.fo + static_cast<int64_t>(t) * bin.fs];
      if (a.hc != 0.0) f = cmul(f, cis_nh(kd * static_cast<double>(t), a.hc));  // uniform
      cr[u] = f.x * w;
      ci[u] = f.y * w;
    } else {
      cr[u] = 0.0f;
      ci[u] = 0.0f;
    }
  }
  const c64 g1 = cis64f(kd * a.hd2);                              // z step between rates
  const c64 gq = cis64f(kd * a.hd2 * static_cast<double>(t0));   // gamma^{t0}
  c64 z = cis64f(kd * a.hz);                                      // e^{2 pi i k (r_0 - rc)/Nf}
  c64 zq = cis64f(kd * a.hz * static_cast<double>(t0));           // z^{t0} of rate 0
  auto reduce = [](c32 v) {
#pragma unroll
    for (int m = 1; m < NC; m <<= 1) {
      v.x += __shfl_xor(v.x, m, 64);
      v.y += __shfl_xor(v.y, m, 64);
    }
    return v;
  };
  auto store = [&](int j, c32 y) {
    if (q == 0) bin.dst[static_cast<int64_t>(j) * bin.stride] = y;
  };
  const int npair = (a.Nr + 1) / 2;  // j < npair; partner Nr-1-j
  int j = 0;
  for (; j + 1 < npair; j += 2) {
    const c64 z1 = cmul(z, g1);
    const c64 zq1 = cmul(zq, gq);
    c32 y0, y0c, y1, y1c;
    reinsch_quad<NT>([&](int t) { return cr[t]; }, [&](int t) { return ci[t]; }, z, [&](int t) { return cr[t]; },
                     [&](int t) { return ci[t]; }, z1, y0, y0c, y1, y1c);
    const c32 s0 = {static_cast<float>(zq.x), static_cast<float>(zq.y)};
    const c32 s1 = {static_cast<float>(zq1.x), static_cast<float>(zq1.y)};
    const int jp = a.Nr - 1 - j;  // > j + 1
    store(j, reduce(cmul(y0, s0)));
    store(jp, reduce(cmulc(y0c, s0)));
    store(j + 1, reduce(cmul(y1, s1)));
    const c32 r1c = reduce(cmulc(y1c, s1));
    if (jp - 1 != j + 1) store(jp - 1, r1c);
    z = cmul(z1, g1);
    zq = cmul(zq1, gq);
  }
  if (j < npair) {
    const int jp = a.Nr - 1 - j;
    c32 yz, yc;
    reinsch_split<NT>(cr, ci, z, yz, yc);
    const c32 s0 = {static_cast<float>(zq.x), static_cast<float>(zq.y)};
    store(j, reduce(cmul(yz, s0)));
    const c32 rc = reduce(cmulc(yc, s0));
    if (jp != j) store(jp, rc);
  }
}

struct K2cArgs {
  const c32* Yc;  // [Nr][N2][N1]
  const c32* yM;  // [Nr]
  c32* B;         // [Nr][N1/16][N2][16]
  c32* yhalf;     // [Nr]
  const c32* tw1;
  int Nr;
  int64_t M, Nf;
  float oscale;   // output scale / M
};

// the rate loop of k2_czt with the spectra read from Yc instead of evaluated
template <int N1, int N2, int ROWS>
__global__ __launch_bounds__(N1 < 512 ? N1 : 512, k2_waves_per_eu(N1)) void k2_c2r(K2cArgs a) {
  constexpr int T = N1 < 512 ? N1 : 512;
  static_assert(T == N1, "one Hermitian pair per thread");
  constexpr int P = N1 + N1 / 8;
  constexpr int HALF = N1 / 2;
  __shared__ c32 buf[ROWS * P];
  __shared__ c32 buf2[ROWS * P];
  __shared__ c32 twl[N1 + N1 / 8];  // W_N1^m at lsw(m)
  for (int i = threadIdx.x; i < N1; i += T) twl[lsw(i)] = a.tw1[i];
  __syncthreads();
  const int tid = threadIdx.x;
  const int p = blockIdx.x;
  const bool special = (p == 0);
  const int colA = special ? 0 : p;
  const int colB = special ? N2 / 2 : N2 - p;
  const HPair hp = hpair<N1, N2>(p, tid);
  const int rowA = hp.rowA, rowB = hp.rowB, k1A = hp.k1A, k1B = hp.k1B, cA = hp.cA, cB = hp.cB;
  const bool nyq = hp.nyq;
  const int64_t kA = cA + static_cast<int64_t>(N2) * k1A;
  const c32 WA = cis_rev(-static_cast<float>(kA) / static_cast<float>(a.Nf));
  const float h = 0.5f * a.oscale;
  if (special) {
    for (int j = tid; j < a.Nr; j += T) {
      const c32 y = a.Yc[static_cast<int64_t>(j) * a.M + HALF];  // bin M/2: k2 = 0, k1 = N1/2
      a.yhalf[j] = {y.x * a.oscale, -y.y * a.oscale};
    }
  }
  const c32* srcA = a.Yc + static_cast<int64_t>(cA) * N1 + k1A;
  const c32* srcB = a.Yc + static_cast<int64_t>(cB) * N1 + (nyq ? 0 : k1B);
  constexpr int JG = ROWS / 2;
  constexpr int NW = T / 64;
  constexpr int RPW = ROWS / NW;
  static_assert(RPW * NW == ROWS && RPW * fft_max_nr<N1>() <= 64, "one butterfly per lane and pass");
  const int wv = tid >> 6, lane = tid & 63;
  c32* cur = buf;
  c32* nxt = buf2;
  // the next group's spectra are loaded right after this group's barrier, so
  // their latency overlaps the inverse DFT
  c32 ya[JG], yb[JG];
  auto load = [&](int j0n) {
    const int njn = min(JG, a.Nr - j0n);
#pragma unroll
    for (int jj = 0; jj < JG; ++jj) {
      if (jj < njn) {
        const int64_t off = static_cast<int64_t>(j0n + jj) * a.M;
        ya[jj] = srcA[off];
        yb[jj] = nyq ? a.yM[j0n + jj] : srcB[off];
      }
    }
  };
  load(0);
  for (int j0 = 0; j0 < a.Nr; j0 += JG) {
    const int nj = min(JG, a.Nr - j0);
#pragma unroll
    for (int jj = 0; jj < JG; ++jj) {
      if (jj < nj) {
        const c32 YA = {ya[jj].x * h, ya[jj].y * h}, YB = {yb[jj].x * h, yb[jj].y * h};
        if (nyq) {
          cur[(2 * jj + rowA) * P] = {YA.x + YB.x, YA.x - YB.x};
          cur[(2 * jj + rowA) * P + lsw(HALF)] = {0.0f, 0.0f};
        } else {
          const c32 E = {YA.x + YB.x, YA.y - YB.y};
          const c32 D = {YA.x - YB.x, YA.y + YB.y};
          const c32 O = cmulc(D, WA);
          cur[(2 * jj + rowA) * P + lsw(k1A)] = {E.x - O.y, E.y + O.x};
          cur[(2 * jj + rowB) * P + lsw(k1B)] = {E.x + O.y, -E.y + O.x};
        }
      }
    }
    __syncthreads();
    if (j0 + JG < a.Nr) load(j0 + JG);
    const int nrw = min(RPW, max(0, 2 * nj - wv * RPW));
    wave_fft<N1, true, 1, true>(cur + wv * RPW * P, P, nrw, twl, lane, [&](int f, int jx, auto& v) {
      constexpr int RL = sizeof(v) / sizeof(c32);
      const int row = wv * RPW + f;
      c32* dst = a.B + static_cast<int64_t>(j0 + (row >> 1)) * (N1 * N2) + ((row & 1) ? colB : colA) * 16;
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const int n1 = jx + r * (N1 / RL);
        dst[(n1 >> 4) * (N2 * 16) + (n1 & 15)] = v[r];
      }
    });
    c32* t = cur;
    cur = nxt;
    nxt = t;
  }
}


// In-place inverse/forward FFT of one row of N entries owned by one
// wavefront, N/(64 R) radix-R butterflies per lane and pass (rows of 2^10 and
// 2^11 entries, too long for wave_fft's one butterfly per lane).  A pass
// issues all of its reads before any of its writes, as in wave_pass; the last
// pass calls emit(j, v) for each butterfly j of the lane (elements
// j + r N/R) instead of writing back.  Twiddles: the table tw of W_N^m.
template <int N, bool INV, int NS = 1, typename Emit>
__device__ __forceinline__ void wave_fft_row(c32* x, const c32* __restrict__ tw, int lane, Emit&& emit) {
  constexpr int R = fft_radix<N, NS>();
  constexpr int NR = N / R;
  constexpr int BPL = NR / 64;
  constexpr int LOGN = ilog2c(N), LOGR = ilog2c(R), LOGNS = ilog2c(NS);
  static_assert(BPL * 64 == NR && NR % 8 == 0, "whole butterflies per lane");
  c32 v[BPL][R];
#pragma unroll
  for (int b = 0; b < BPL; ++b) {
    const c32* xr = x + lsw(lane + 64 * b);
#pragma unroll
    for (int r = 0; r < R; ++r) v[b][r] = xr[r * (NR + NR / 8)];
  }
#pragma unroll
  for (int b = 0; b < BPL; ++b) {
    if constexpr (NS > 1) {
      const int k = (lane + 64 * b) & (NS - 1);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const c32 w = tw[(r * k) << (LOGN - LOGNS - LOGR)];
        v[b][r] = INV ? cmulc(v[b][r], w) : cmul(v[b][r], w);
      }
    }
    dftd<INV, float, R>(v[b]);
  }
  if constexpr (NS * R < N) {
#pragma unroll
    for (int b = 0; b < BPL; ++b) {
      const int j = lane + 64 * b;
      const int k = j & (NS - 1);
      c32* yy = x + lsw(((j >> LOGNS) << (LOGNS + LOGR)) + k);
      if constexpr (NS == 1) {
#pragma unroll
        for (int r = 0; r < R; ++r) yy[r + (r >> 3)] = v[b][r];
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) yy[r * (NS + (NS >> 3))] = v[b][r];
      }
    }
    wave_sync();
    wave_fft_row<N, INV, NS * R>(x, tw, lane, emit);
  } else {
#pragma unroll
    for (int b = 0; b < BPL; ++b) emit(lane + 64 * b, v[b]);
  }
}

// Inverse DFT of one 2048-entry row owned by one wavefront, in place and
// with 32 row entries per lane but at most 16 live at a time (a Stockham row
// FFT keeps all 32 live, which does not fit beside k2_c2r's state).  Stage 1
// is a radix-16 Stockham pass from natural order (butterfly j reads entries
// j + 128 r) whose outputs go to the digit-reversed positions of an in-place
// decimation-in-time FFT (group g = j/16 + 8 (j mod 16), entries 16 g + r');
// stages 2 (radix 8, span 16) and 3 (radix 16, span 128) are in-place DIT
// butterflies on entries g L + k + r L', one butterfly at a time, and stage 3
// calls emit(k, v) with v[r] = output k + 128 r.  Rows are stored with the
// swizzle s2048 (bits 7-10 folded into bits 0-3), which keeps every stage's
// accesses within two-way bank conflicts; tw is the LDS table of W_2048^m.
__device__ __forceinline__ int s2048(int i) { return i ^ ((i >> 7) & 15); }

template <typename Emit>
__device__ __forceinline__ void wave_ifft_2048(c32* x, const c32* tw, int lane, Emit&& emit) {
  // opaque to the optimizer: keeps the ~100 lane-dependent LDS addresses from
  // being hoisted out of the caller's rate loop (where they would spill)
  asm volatile("" : "+v"(lane));
  {
    c32 v[2][16];
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) v[b][r] = x[s2048(lane + 64 * b + 128 * r)];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      dftd<true, float, 16>(v[b]);
      const int j = lane + 64 * b;
      const int g = (j >> 4) + 8 * (j & 15);
#pragma unroll
      for (int r = 0; r < 16; ++r) x[s2048(16 * g + r)] = v[b][r];
    }
  }
  wave_sync();
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int bb = lane + 64 * b;
    const int base = (bb >> 4) * 128 + (bb & 15);
    const int k = bb & 15;
    c32 v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = x[s2048(base + 16 * r)];
#pragma unroll
    for (int r = 1; r < 8; ++r) v[r] = cmulc(v[r], tw[16 * r * k]);  // W_128^{-r k}
    dftd<true, float, 8>(v);
#pragma unroll
    for (int r = 0; r < 8; ++r) x[s2048(base + 16 * r)] = v[r];
  }
  wave_sync();
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int k = lane + 64 * b;
    c32 v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = x[s2048(k + 128 * r)];
#pragma unroll
    for (int r = 1; r < 16; ++r) v[r] = cmulc(v[r], tw[r * k]);  // W_2048^{-r k}
    dftd<true, float, 16>(v);
    emit(k, v);
  }
}

// k2_c2r for N1 = 1024, 2048 (Nf = 2^21, 2^22): 512 threads, PP = N1/512
// pairs each, groups of 4 rates (8 rows, one per wavefront) in one LDS
// buffer; each wavefront runs the inverse DFT_N1 of its row in place
// (wave_fft_row) and stores it, so a group costs two barriers.  The next
// group's spectra are loaded into registers before the DFTs, so their
// latency overlaps them (their addresses are formed inside the rate loop:
// hoisted out of it, they spilled).
template <int N1, int N2>
__global__ __launch_bounds__(512) void k2_c2r_wave(K2cArgs a) {
  constexpr int T = 512;
  constexpr int PP = N1 / T;
  constexpr bool BIG = N1 == 2048;  // wave_ifft_2048 layout (no row padding)
  constexpr int P = BIG ? N1 : N1 + N1 / 8;
  constexpr int HALF = N1 / 2;
  constexpr int ROWS = 8;
  constexpr int JG = ROWS / 2;
  __shared__ c32 buf[ROWS * P];
  __shared__ c32 twl[BIG ? N1 : 1];
  if constexpr (BIG) {
    for (int i = threadIdx.x; i < N1; i += T) twl[i] = a.tw1[i];
  }
  auto pos = [](int i) { return BIG ? s2048(i) : lsw(i); };
  const int tid = threadIdx.x;
  const int p = blockIdx.x;
  const bool special = (p == 0);
  const int colA = special ? 0 : p;
  const int colB = special ? N2 / 2 : N2 - p;
  int rowA[PP], rowB[PP], k1A[PP], k1B[PP];
  bool nyq[PP];
  c32 WA[PP];
  int offA[PP], offB[PP];  // within a rate's M spectra
#pragma unroll
  for (int i = 0; i < PP; ++i) {
    const HPair h = hpair<N1, N2>(p, tid + i * T);
    rowA[i] = h.rowA;
    rowB[i] = h.rowB;
    k1A[i] = h.k1A;
    k1B[i] = h.k1B;
    nyq[i] = h.nyq;
    const int cA = h.cA, cB = h.cB;
    const int64_t kA = cA + static_cast<int64_t>(N2) * k1A[i];
    WA[i] = cis_rev(-static_cast<float>(kA) / static_cast<float>(a.Nf));
    offA[i] = cA * N1 + k1A[i];
    offB[i] = cB * N1 + (nyq[i] ? 0 : k1B[i]);
  }
  const float h = 0.5f * a.oscale;
  if (special) {
    for (int j = tid; j < a.Nr; j += T) {
      const c32 y = a.Yc[static_cast<int64_t>(j) * a.M + HALF];
      a.yhalf[j] = {y.x * a.oscale, -y.y * a.oscale};
    }
  }
  c32 ya[JG][PP], yb[JG][PP];
  auto load = [&](int j0n) {
    const int njn = min(JG, a.Nr - j0n);
#pragma unroll
    for (int jj = 0; jj < JG; ++jj) {
      if (jj < njn) {
        const c32* y = a.Yc + static_cast<int64_t>(j0n + jj) * a.M;  // uniform
#pragma unroll
        for (int i = 0; i < PP; ++i) {
          // opaque offsets: the 64-bit addresses are formed here, not hoisted
          // out of the rate loop (where they would be spilled)
          int oa = offA[i], ob = offB[i];
          asm volatile("" : "+v"(oa), "+v"(ob));
          ya[jj][i] = y[oa];
          yb[jj][i] = nyq[i] ? a.yM[j0n + jj] : y[ob];
        }
      }
    }
  };
  load(0);
  const int wv = tid >> 6, lane = tid & 63;
  for (int j0 = 0; j0 < a.Nr; j0 += JG) {
    const int nj = min(JG, a.Nr - j0);
#pragma unroll
    for (int jj = 0; jj < JG; ++jj) {
      if (jj < nj) {
#pragma unroll
        for (int i = 0; i < PP; ++i) {
          const c32 YA = {ya[jj][i].x * h, ya[jj][i].y * h}, YB = {yb[jj][i].x * h, yb[jj][i].y * h};
          if (nyq[i]) {
            buf[(2 * jj + rowA[i]) * P] = {YA.x + YB.x, YA.x - YB.x};
            buf[(2 * jj + rowA[i]) * P + pos(HALF)] = {0.0f, 0.0f};
          } else {
            const c32 E = {YA.x + YB.x, YA.y - YB.y};
            const c32 D = {YA.x - YB.x, YA.y + YB.y};
            const c32 O = cmulc(D, WA[i]);
            buf[(2 * jj + rowA[i]) * P + pos(k1A[i])] = {E.x - O.y, E.y + O.x};
            buf[(2 * jj + rowB[i]) * P + pos(k1B[i])] = {E.x + O.y, -E.y + O.x};
          }
        }
      }
    }
    __syncthreads();
    if (j0 + JG < a.Nr) load(j0 + JG);
    if (wv < 2 * nj) {  // uniform per wavefront: wavefront wv owns row wv
      c32* dst = a.B + static_cast<int64_t>(j0 + (wv >> 1)) * (N1 * N2) + ((wv & 1) ? colB : colA) * 16;
      auto emit = [&](int jx, auto& v) {
        constexpr int RL = sizeof(v) / sizeof(c32);
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          const int n1 = jx + r * (N1 / RL);
          dst[(n1 >> 4) * (N2 * 16) + (n1 & 15)] = v[r];
        }
      };
      if constexpr (BIG) {
        wave_ifft_2048(buf + wv * P, twl, lane, emit);
      } else {
        int ln = lane;
        asm volatile("" : "+v"(ln));  // keeps the row FFT's LDS addresses inside the rate loop
        wave_fft_row<N1, true>(buf + wv * P, a.tw1, ln, emit);
      }
    }
    __syncthreads();  // the rows are rewritten by the next group
  }
}

// ---------------------------------------------------------------- fused-output input

struct K2iArgs {
  c32* G;          // in: [Nt][N2][N1] column FFTs of k1_input; out: X_t(k2 + N2 k1) at [t][k2][k1]
  c32* XM;         // [Nt] X_t(M)
  const c32* tw1;  // W_N1^m
  int Nt;
  int64_t M, Nf;
};

// input! for the fused output: the second half of the four-step r2c, in
// place over k1_input's G.  Block per Hermitian column pair as in k2_czt
// (regular block p: columns p and N2-p; block 0: columns 0 and N2/2); for each
// group of TG = ROWS/2 time steps the 2 TG rows get the twiddle W_M^{n1 k2}
// and a forward DFT_N1, then each thread untangles its pairs (k, M-k) into
// the r2c spectrum and writes both bins back over the rows it read.
template <int N1, int N2, int ROWS>
__global__ __launch_bounds__(N1 < 1024 ? N1 : 1024) void k2_input(K2iArgs a) {
  constexpr int T = N1 < 1024 ? N1 : 1024;
  constexpr int PP = N1 / T;
  constexpr int P = N1 + N1 / 8;
  constexpr int HALF = N1 / 2;
  constexpr int TG = ROWS / 2;
  __shared__ c32 buf[ROWS * P];
  __shared__ c32 buf2[ROWS * P];
  const int tid = threadIdx.x;
  const int p = blockIdx.x;
  const bool special = (p == 0);
  const int colA = special ? 0 : p;
  const int colB = special ? N2 / 2 : N2 - p;

  int rowA[PP], rowB[PP], k1A[PP], k1B[PP];
  bool nyq[PP];
  c32 WA[PP], twS0[PP], twS1[PP];
  const float invM = 1.0f / static_cast<float>(a.M);
#pragma unroll
  for (int i = 0; i < PP; ++i) {
    const int q = tid + i * T;
    const HPair h = hpair<N1, N2>(p, q);
    rowA[i] = h.rowA;
    rowB[i] = h.rowB;
    k1A[i] = h.k1A;
    k1B[i] = h.k1B;
    nyq[i] = h.nyq;
    const int cA = h.cA;
    const int64_t kA = cA + static_cast<int64_t>(N2) * k1A[i];
    WA[i] = cis_rev(-static_cast<float>(kA) / static_cast<float>(a.Nf));  // W_Nf^{kA}
    const int64_t n1 = q;
    twS0[i] = cis_rev(static_cast<float>((n1 * colA) & (a.M - 1)) * invM);
    twS1[i] = cis_rev(static_cast<float>((n1 * colB) & (a.M - 1)) * invM);
  }

  for (int t0 = 0; t0 < a.Nt; t0 += TG) {
    const int nt = min(TG, a.Nt - t0);
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      if (r < 2 * nt) {
        const c32* src = a.G + (static_cast<int64_t>(t0 + (r >> 1)) * N2 + ((r & 1) ? colB : colA)) * N1;
#pragma unroll
        for (int i = 0; i < PP; ++i)
          buf[r * P + lsw(tid + i * T)] = cmulc(src[tid + i * T], (r & 1) ? twS1[i] : twS0[i]);
      }
    }
    __syncthreads();
    const c32* fz = lds_fft_pp<N1, T, false>(buf, buf2, P, 2 * nt, a.tw1);
#pragma unroll
    for (int uu = 0; uu < TG; ++uu) {
      if (uu < nt) {
        const int64_t t = t0 + uu;
        c32* rA = a.G + (t * N2 + colA) * N1;
        c32* rB = a.G + (t * N2 + colB) * N1;
#pragma unroll
        for (int i = 0; i < PP; ++i) {
          const c32 Za = fz[(2 * uu + rowA[i]) * P + lsw(k1A[i])];
          c32* dA = rowA[i] ? rB : rA;
          c32* dB = rowB[i] ? rB : rA;
          if (nyq[i]) {
            dA[0] = {Za.x + Za.y, 0.0f};
            a.XM[t] = {Za.x - Za.y, 0.0f};
          } else {
            const c32 Zb = fz[(2 * uu + rowB[i]) * P + lsw(k1B[i])];
            const c32 E = {0.5f * (Za.x + Zb.x), 0.5f * (Za.y - Zb.y)};
            const c32 O = {0.5f * (Za.y + Zb.y), -0.5f * (Za.x - Zb.x)};
            const c32 WO = cmul(WA[i], O);
            dA[k1A[i]] = cadd(E, WO);
            dB[k1B[i]] = {E.x - WO.x, -(E.y - WO.y)};
          }
        }
        if (special && tid == 0) {
          // bin M/2 (column 0, k1 = N1/2) is its own partner: X = conj(Z)
          const c32 Zh = fz[(2 * uu) * P + lsw(HALF)];
          rA[HALF] = cconj(Zh);
        }
      }
    }
    __syncthreads();
  }
}

// F (natural Nk x Nt) from the fused-output input layout
__global__ __launch_bounds__(256) void ft_to_natural(const c32* __restrict__ Ft, c32* __restrict__ F, int64_t Nk,
                                                     int Nt, int logN1, int logN2) {
  const int64_t M = Nk - 1;
  const int64_t total = Nk * Nt;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t t = g / Nk, k = g - t * Nk;
    const int64_t k2 = k & ((int64_t(1) << logN2) - 1), k1 = k >> logN2;
    F[g] = (k == M) ? Ft[M * Nt + t] : Ft[t * M + (k2 << logN1) + k1];
  }
}

}  // namespace zdt

namespace zdt {

// ---------------------------------------------------------------- fused output (host)

bool fast_out_supported(int64_t Nf, int64_t Nt, int64_t Nr) {
  if (Nf < (1 << 16) || Nf > (1 << 22) || (Nf & (Nf - 1)) != 0) return false;
  return Nt >= 1 && Nt <= FOUT_MAX_NT && Nr >= 1 && Nr <= 65535;
}

int64_t fast_out_bytes(int64_t Nf, int64_t Nr) {
  const int64_t M = Nf / 2;
  const int64_t N2 = fast_n2(Nf);
  return 16 * M * Nr + 8 * (M / N2 + N2 + 2 * Nr) + 8 * fast_stats_doubles(Nf, Nr);
}

int fast_out_create(int64_t Nf, int64_t Nr, int precision, FastPlan** out) {
  *out = nullptr;
  if (!fast_out_supported(Nf, 1, Nr)) return fail(ZDT_ENOTSUP, "fused output: unsupported shape");
  if (!fast_precision_ok(Nf, 0, precision))
    return fail(ZDT_ENOTSUP, "fused output: precision %d not available for Nf=%lld", precision, (long long)Nf);
  FastPlan* p = new FastPlan();
  p->precision = precision;
  p->Nf = Nf;
  p->Nt = 0;
  p->Nr = Nr;
  p->M = Nf / 2;
  p->N2 = fast_n2(Nf);
  p->logN2 = ilog2i(p->N2);
  p->N1 = static_cast<int>(p->M / p->N2);
  p->logN1 = ilog2i(p->N1);
  {
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    p->resident_blocks = 2 * static_cast<int64_t>(ncu);
  }
  int rc = ZDT_OK;
  auto alloc = [&](c32** q, size_t n) {
    if (rc != ZDT_OK) return;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(q), n * sizeof(c32));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      rc = fail(ZDT_ENOMEM, "hipMalloc(%zu bytes) failed: %s", n * sizeof(c32), hipGetErrorString(e));
    }
  };
  alloc(&p->Yc, static_cast<size_t>(Nr * p->M));
  alloc(&p->B, static_cast<size_t>(Nr * p->M));
  alloc_part(p, rc);
  alloc(&p->yM, static_cast<size_t>(Nr));
  alloc(&p->yhalf, static_cast<size_t>(Nr));
  if (rc == ZDT_OK) rc = upload_table(&p->tw1, twiddles(p->N1));
  if (rc == ZDT_OK) rc = upload_table(&p->tw2, twiddles(p->N2));
  if (rc != ZDT_OK) {
    fast_destroy(p);
    return rc;
  }
  *out = p;
  return ZDT_OK;
}

template <int N1, int N2>
static int launch_k2c(const K2cArgs& a, hipStream_t st) {
  if constexpr (N1 <= 512) {
    k2_c2r<N1, N2, k2_rows(N1)><<<N2 / 2, N1, 0, st>>>(a);
  } else {
    k2_c2r_wave<N1, N2><<<N2 / 2, 512, 0, st>>>(a);
  }
  ZDT_CHECK_LAUNCH("k2_c2r");
  return ZDT_OK;
}

template <int N1, int N2>
static void launch_k2i(const K2iArgs& a, hipStream_t st) {
  // two ping-pong buffers of ~36 KB each: 8 rows up to N1 = 512, then 4, 2
  constexpr int ROWS = N1 <= 512 ? 8 : (N1 == 1024 ? 4 : 2);
  k2_input<N1, N2, ROWS><<<N2 / 2, (N1 < 1024 ? N1 : 1024), 0, st>>>(a);
}

int fast_out_input(FastPlan* p, const float* spec, int64_t Nt, c32* Ft, hipStream_t st) {
  if (Nt < 1 || Nt > FOUT_MAX_NT) return fail(ZDT_ENOTSUP, "fused output: Nt=%lld", (long long)Nt);
  ZDT_TRY(launch_k1(p, spec, Ft, Nt, st));
  K2iArgs a;
  a.G = Ft;
  a.XM = Ft + p->M * Nt;
  a.tw1 = p->tw1;
  a.Nt = static_cast<int>(Nt);
  a.M = p->M;
  a.Nf = p->Nf;
  ProfScope ps(st, "k2_input", 16.0 * p->M * Nt);
  switch (p->N1 * 2048 + p->N2) {
    case 64 * 2048 + 512:
      launch_k2i<64, 512>(a, st);
      break;
    case 64 * 2048 + 1024:
      launch_k2i<64, 1024>(a, st);
      break;
    case 128 * 2048 + 1024:
      launch_k2i<128, 1024>(a, st);
      break;
    case 256 * 2048 + 1024:
      launch_k2i<256, 1024>(a, st);
      break;
    case 512 * 2048 + 1024:
      launch_k2i<512, 1024>(a, st);
      break;
    case 1024 * 2048 + 1024:
      launch_k2i<1024, 1024>(a, st);
      break;
    case 2048 * 2048 + 1024:
      launch_k2i<2048, 1024>(a, st);
      break;
    default:
      return fail(ZDT_ENOTSUP, "fused output input: N1=%d N2=%d", p->N1, p->N2);
  }
  ZDT_CHECK_LAUNCH("k2_input");
  return ZDT_OK;
}

int fast_out_unpack_F(FastPlan* p, const c32* Ft, int64_t Nt, c32* F, hipStream_t st) {
  const int64_t n = (p->M + 1) * Nt;
  int64_t g = (n + 255) / 256;
  if (g > (1 << 20)) g = 1 << 20;
  ft_to_natural<<<static_cast<unsigned>(g), 256, 0, st>>>(Ft, F, p->M + 1, static_cast<int>(Nt), p->logN1, p->logN2);
  ZDT_CHECK_LAUNCH("ft_to_natural");
  return ZDT_OK;
}

int fast_out_run(FastPlan* p, const c32* F, int64_t ldF, bool ftrans, int Nt, float r0, float dr, int wkind,
                 const float* window, float* dest, double scale, hipStream_t st, double* colmean, double* colvar) {
  if (Nt < 1 || Nt > FOUT_MAX_NT) return fail(ZDT_ENOTSUP, "fused output: Nt=%d", Nt);
  CztColArgs c;
  c.F = F;
  c.ldF = ldF;
  c.ftrans = ftrans ? 1 : 0;
  c.Yc = p->Yc;
  c.yM = p->yM;
  c.M = p->M;
  c.Nf = p->Nf;
  c.logN1 = p->logN1;
  c.logN2 = p->logN2;
  c.Nt = Nt;
  c.Nr = static_cast<int>(p->Nr);
  c.hd2 = static_cast<double>(dr) / static_cast<double>(p->Nf);
  c.wkind = wkind;
  c.wcustom = window;
  c.scale = 1.0f;
  // rates in mirrored pairs around the centre rate (a single rate is its own partner)
  const double rc = static_cast<double>(r0) + 0.5 * static_cast<double>(p->Nr - 1) * static_cast<double>(dr);
  c.hc = rc / static_cast<double>(p->Nf);
  c.hz = (static_cast<double>(r0) - rc) / static_cast<double>(p->Nf);
  if (Nt <= 64) {
    const unsigned grid = static_cast<unsigned>(p->M / 256 + 1);
    ProfScope ps(st, "czt_reinsch_col", 8.0 * (p->M + 1) * (Nt + p->Nr));  // F read, Yc written, once each
    if (Nt <= 16) {
      czt_reinsch_col<16><<<grid, 256, 0, st>>>(c);
    } else if (Nt <= 32) {
      czt_reinsch_col<32><<<grid, 256, 0, st>>>(c);
    } else {
      czt_reinsch_col<64><<<grid, 256, 0, st>>>(c);
    }
    ZDT_CHECK_LAUNCH("czt_reinsch_col");
  } else {
    // NC lanes per bin: the 64-step chunks, a power of two (FOUT_MAX_NT = 1024: at most 16)
    int nc = 2;
    while (nc * 64 < Nt) nc *= 2;
    const int64_t bins = p->M + 1;
    const unsigned grid = static_cast<unsigned>((bins + 256 / nc - 1) / (256 / nc));
    ProfScope ps(st, "czt_reinsch_col_split", 8.0 * (p->M + 1) * (Nt + p->Nr));
    switch (nc) {
      case 2:
        czt_reinsch_col_split<2><<<grid, 256, 0, st>>>(c);
        break;
      case 4:
        czt_reinsch_col_split<4><<<grid, 256, 0, st>>>(c);
        break;
      case 8:
        czt_reinsch_col_split<8><<<grid, 256, 0, st>>>(c);
        break;
      case 16:
        czt_reinsch_col_split<16><<<grid, 256, 0, st>>>(c);
        break;
      default:
        return fail(ZDT_ENOTSUP, "fused output: Nt=%d above %lld", Nt, (long long)FOUT_MAX_NT);
    }
    ZDT_CHECK_LAUNCH("czt_reinsch_col_split");
  }
  K2cArgs a;
  a.Yc = p->Yc;
  a.yM = p->yM;
  a.B = p->B;
  a.yhalf = p->yhalf;
  a.tw1 = p->tw1;
  a.Nr = static_cast<int>(p->Nr);
  a.M = p->M;
  a.Nf = p->Nf;
  a.oscale = static_cast<float>(scale / static_cast<double>(p->M));
  {
    ProfScope ps(st, "k2_c2r", 16.0 * p->M * p->Nr);
    switch (p->N1 * 2048 + p->N2) {
      case 64 * 2048 + 512:
        ZDT_TRY((launch_k2c<64, 512>(a, st)));
        break;
      case 64 * 2048 + 1024:
        ZDT_TRY((launch_k2c<64, 1024>(a, st)));
        break;
      case 128 * 2048 + 1024:
        ZDT_TRY((launch_k2c<128, 1024>(a, st)));
        break;
      case 256 * 2048 + 1024:
        ZDT_TRY((launch_k2c<256, 1024>(a, st)));
        break;
      case 512 * 2048 + 1024:
        ZDT_TRY((launch_k2c<512, 1024>(a, st)));
        break;
      case 1024 * 2048 + 1024:
        ZDT_TRY((launch_k2c<1024, 1024>(a, st)));
        break;
      case 2048 * 2048 + 1024:
        ZDT_TRY((launch_k2c<2048, 1024>(a, st)));
        break;
      default:
        return fail(ZDT_ENOTSUP, "fused output: N1=%d N2=%d", p->N1, p->N2);
    }
  }
  p->have_B = true;
  if (dest != nullptr) ZDT_TRY(fast_output(p, dest, scale, st, colmean, colvar));
  return ZDT_OK;
}

}  // namespace zdt
