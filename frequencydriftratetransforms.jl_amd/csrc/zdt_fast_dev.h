// Device building blocks of the fused kernels (zdt_fast.hip, zdt_fout.hip):
// LDS and wavefront FFT passes, the register-resident column FFT of k1/k3,
// window weights, FP64 phases, the Hermitian column pairs of k2 and the
// chirp-Z evaluations.
#pragma once

#include <hip/hip_runtime.h>

#include "fft_butterflies.h"
#include "zdt_internal.h"

namespace zdt {

// ---------------------------------------------------------------- LDS FFT

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }

// radix of the pass at stride NS in the LDS FFTs: 1024 = 16*8*8,
// 512 = 8*8*8, 256 = 16*16, 128 = 16*8, 64 = 8*8
template <int N, int NS>
constexpr int fft_radix() {
  return (N / NS == 16 || (NS == 1 && ilog2c(N) % 3 != 0)) ? 16 : (N / NS >= 8 ? 8 : N / NS);
}

// LDS row swizzle: one pad entry after every 8, so the strided Stockham
// writes of the early passes spread over the banks.  Rows of length N occupy
// N + N/8 entries.
__device__ __forceinline__ int lsw(int i) { return i + (i >> 3); }

// Second row layout (LAY = 1), for 512-point rows of the radix 8*8*8
// passes: rows hold entry i at q64(i) = i + 8 (i/64) (pitch 576), except the
// output of the first pass, which is written (and read by the second pass)
// at lsw(i).  Measured on gfx950 (SQ_LDS_BANK_CONFLICT), 8-byte LDS reads are
// served 32 lanes per cycle over 64 banks and 8-byte writes 16 lanes per
// cycle over 32 banks.  Under that model, q64 keeps the contiguous reads and
// writes (j + 64 r, runs of lanes) and the second-pass writes (64 (j/8) + j%8
// + 8 r) conflict-free; no padding of this kind does so for the first-pass
// writes (8 j + r), which lsw does, at the cost of two-way conflicts on the
// second pass's reads -- one conflicted access pattern per 512-point FFT
// where lsw throughout has four.  Offsets stay immediate: q64(j + 64 r) =
// q64(j) + 72 r = lsw-alike, q64(d0 + 8 r) = q64(d0) + 8 r for d0 % 64 < 8,
// lsw(8 j + r) = lsw(8 j) + r.  Twiddles come from per-pass tables
// (lay1_load_twiddles), read at contiguous or broadcast addresses.
__device__ __forceinline__ int q64(int i) { return i + ((i >> 6) << 3); }

template <int LAY>
__device__ __forceinline__ int row_index(int i) {
  return LAY == 1 ? q64(i) : lsw(i);
}

// Per-pass twiddle table of the LAY = 1 passes (N = 512): the pass at stride
// Ns (8 or 64) reads W^{(r k) * 512 / (8 Ns)} at (Ns - 8) + (r - 1) Ns + k,
// k < Ns, r = 1..7 (504 entries), W^m = tw[m].
constexpr int LAY1_TW = 504;

__device__ __forceinline__ void lay1_load_twiddles(c32* twp, const c32* __restrict__ tw, int nthreads) {
  for (int e = threadIdx.x; e < LAY1_TW; e += nthreads) {
    const int Ns = e < 56 ? 8 : 64;
    const int q = e - (Ns - 8);
    const int r = q / Ns + 1, k = q % Ns;
    twp[e] = tw[(r * k) * (64 / Ns)];
  }
}

// Twiddle tables in LDS are read at (r k) << sh, a stride of 2^sh entries
// across the lanes of a pass; with SW the passes read a table stored
// swizzled like the rows (entry m at lsw(m), N + N/8 entries), which spreads
// those strided reads over the banks.

// Ping-pong variant: radix-R butterflies stream from buffer `in` to buffer
// `out` (one butterfly live per thread), one barrier per pass.  Rows are
// swizzled (lsw); since NR and Ns are multiples of 8 or the write base is,
// every butterfly needs the swizzle only once for its reads and once for its
// writes.  Twiddles come from the LDS table tw (W_N^m, m < N); INV selects the
// inverse butterflies and conjugate twiddles.
template <int N, int R, int T, bool INV, bool SW = false, int LAY = 0>
__device__ __forceinline__ void lds_pass_pp(const c32* in, c32* out, int P, int nrows, int Ns,
                                            const c32* __restrict__ tw) {
  constexpr int NR = N / R;
  constexpr int LOGNR = ilog2c(NR);
  constexpr int LOGN = ilog2c(N);
  constexpr int LOGR = ilog2c(R);
  static_assert(NR % 8 == 0, "pass layout assumes N/R is a multiple of 8");
  static_assert(LAY == 0 || (N == 512 && R == 8), "layout 1 is the 512 = 8*8*8 schedule");
  const int logNs = 31 - __builtin_clz(Ns);
  const int total = nrows * NR;
  const int sh = LOGN - logNs - LOGR;
  for (int b = threadIdx.x; b < total; b += T) {
    const int f = b >> LOGNR;
    const int j = b & (NR - 1);
    int k = j & (Ns - 1);
    // Opaque to the optimizer: keeps the R-1 twiddle addresses from being
    // hoisted out of the caller's rate loop, where they would be spilled.
    asm volatile("" : "+v"(k));
    // layout 1: the second pass reads the first pass's lsw output
    const c32* x = in + f * P + (LAY == 1 && Ns != 8 ? q64(j) : lsw(j));
    c32 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = x[r * (NR + NR / 8)];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const c32 w = LAY == 1 ? tw[(Ns - 8) + (r - 1) * Ns + k] : tw[SW ? lsw((r * k) << sh) : (r * k) << sh];
        v[r] = INV ? cmulc(v[r], w) : cmul(v[r], w);
      }
    }
    dftd<INV, float, R>(v);
    const int d0 = ((j >> logNs) << (logNs + LOGR)) + k;
    c32* y = out + f * P;
    if (Ns == 1) {
      // d0 = j*R is a multiple of 8 (R >= 8): lsw(d0 + r) = lsw(d0) + r + r/8
      // (both layouts write the first pass at lsw)
      c32* yy = y + lsw(d0);
#pragma unroll
      for (int r = 0; r < R; ++r) yy[r + (r >> 3)] = v[r];
    } else {
      // Ns >= 8: the stride Ns keeps the swizzle offset linear (layout 1:
      // d0 % 64 < 8 for Ns = 8, d0 < 64 for Ns = 64)
      c32* yy = y + row_index<LAY>(d0);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (LAY == 1)
          yy[Ns == 8 ? 8 * r : 72 * r] = v[r];
        else
          yy[r * (Ns + (Ns >> 3))] = v[r];
      }
    }
  }
  __syncthreads();
}

// radix schedule: 1024 = 16*8*8, 512 = 8*8*8, 256 = 16*16, 128 = 16*8,
// 64 = 8*8; returns the buffer holding the result
template <int N, int T, bool INV, int NS = 1, bool SW = false, int LAY = 0>
__device__ __forceinline__ c32* lds_fft_pp(c32* a, c32* b, int P, int nrows, const c32* __restrict__ tw) {
  if constexpr (NS < N) {
    constexpr int R = fft_radix<N, NS>();
    lds_pass_pp<N, R, T, INV, SW, LAY>(a, b, P, nrows, NS, tw);
    return lds_fft_pp<N, T, INV, NS * R, SW, LAY>(b, a, P, nrows, tw);
  } else {
    return a;
  }
}

// Wavefront-level passes.  The rows [0, nrw) at `rows` (pitch P, swizzled)
// belong to one wavefront, and lane b < nrw*N/R owns butterfly b of each
// pass.  A wavefront issues all its reads of a pass before any of its
// writes and its LDS accesses complete in order, so the passes run in place
// with no workgroup barrier; wave_sync only keeps the compiler from moving
// LDS accesses across pass boundaries.  Requires nrw*N/R <= 64 for every
// radix R of the schedule.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int N, int R, bool INV, bool SW = false, int LAY = 0>
__device__ __forceinline__ void wave_pass(c32* rows, int P, int nrw, int Ns, const c32* __restrict__ tw, int lane,
                                          c32 (&v)[R], bool last) {
  constexpr int NR = N / R;
  constexpr int LOGNR = ilog2c(NR);
  constexpr int LOGN = ilog2c(N);
  constexpr int LOGR = ilog2c(R);
  static_assert(NR % 8 == 0, "pass layout assumes N/R is a multiple of 8");
  static_assert(LAY == 0 || (N == 512 && R == 8), "layout 1 is the 512 = 8*8*8 schedule");
  if (lane < nrw * NR) {
    const int logNs = 31 - __builtin_clz(Ns);
    int ln = lane;
    // the lane's row and twiddle addresses depend on nothing else: opaque,
    // they are formed here, not hoisted out of the caller's loop and spilled
    if constexpr (LAY == 1) asm volatile("" : "+v"(ln));
    const int f = ln >> LOGNR;
    const int j = ln & (NR - 1);
    int k = j & (Ns - 1);
    asm volatile("" : "+v"(k));  // keeps twiddle addresses out of the caller's loop
    c32* x = rows + f * P;
    const c32* xr = x + (LAY == 1 && Ns != 8 ? q64(j) : lsw(j));
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = xr[r * (NR + NR / 8)];
    if (Ns > 1) {
      const int sh = LOGN - logNs - LOGR;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const c32 w = LAY == 1 ? tw[(Ns - 8) + (r - 1) * Ns + k] : tw[SW ? lsw((r * k) << sh) : (r * k) << sh];
        v[r] = INV ? cmulc(v[r], w) : cmul(v[r], w);
      }
    }
    dftd<INV, float, R>(v);
    if (!last) {
      const int d0 = ((j >> logNs) << (logNs + LOGR)) + k;
      if (Ns == 1) {
        c32* yy = x + lsw(d0);
#pragma unroll
        for (int r = 0; r < R; ++r) yy[r + (r >> 3)] = v[r];
      } else {
        c32* yy = x + row_index<LAY>(d0);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if constexpr (LAY == 1)
            yy[Ns == 8 ? 8 * r : 72 * r] = v[r];
          else
            yy[r * (Ns + (Ns >> 3))] = v[r];
        }
      }
    }
  }
  wave_sync();
}

template <int N, int NS = 1>
constexpr int fft_max_nr() {
  if constexpr (NS >= N) {
    return 0;
  } else {
    constexpr int R = fft_radix<N, NS>();
    constexpr int a = N / R, b = fft_max_nr<N, NS * R>();
    return a > b ? a : b;
  }
}

// In-place inverse/forward FFT of the wavefront's rows; the last pass leaves
// its butterfly outputs in registers (lane (f, j): row f, elements
// j + r*N/R) and calls emit(f, j, v) instead of writing them back.
template <int N, bool INV, int NS = 1, bool SW = false, int LAY = 0, typename Emit>
__device__ __forceinline__ void wave_fft(c32* rows, int P, int nrw, const c32* __restrict__ tw, int lane, Emit&& emit) {
  constexpr int R = fft_radix<N, NS>();
  c32 v[R];
  if constexpr (NS * R < N) {
    wave_pass<N, R, INV, SW, LAY>(rows, P, nrw, NS, tw, lane, v, false);
    wave_fft<N, INV, NS * R, SW, LAY>(rows, P, nrw, tw, lane, emit);
  } else {
    wave_pass<N, R, INV, SW, LAY>(rows, P, nrw, NS, tw, lane, v, true);
    constexpr int NR = N / R;
    if (lane < nrw * NR) emit(lane / NR, lane % NR, v);
  }
}

// ---------------------------------------------------------------- column FFT (N2 = 32*R1)
//
// k1 and k3 transform 16 adjacent columns (n1 or k1 fixed, 128 contiguous
// bytes per row) along a strided axis of length N2 = 32*R1 (R1 = 16 or 32),
// register resident: thread (c = tid&15, r = tid>>4) holds rows r + R1*m
// (m < 32) of column c, does a DFT-32 over m, multiplies by W_N2^{r m'}, and
// after one LDS exchange the thread (c, g) does the DFT-R1 over r for
// m' = g + R1*h, giving output rows m' + 32 r'.  Inverse transforms conjugate
// on entry and exit.  Blocks are persistent over tiles.

template <int R1>
struct CF {
  static constexpr int N2 = 32 * R1;
  static constexpr int THREADS = 16 * R1;
  static constexpr int PITCH = R1 * 16 + 8;  // complex per m' slab ([r][c] + pad)
  static constexpr int LDS = 32 * PITCH;
  static constexpr int MPER = 32 / R1;       // m' values per thread in stage 2
};

template <int R1>
__device__ __forceinline__ void cf_stage1(c32 (&x)[32], int r, const c32* twl, c32* ex, int c, bool inv) {
  if (inv) {
#pragma unroll
    for (int m = 0; m < 32; ++m) x[m].y = -x[m].y;
  }
  dft32(x);
#pragma unroll
  for (int m = 1; m < 32; ++m) x[m] = cmul(x[m], twl[r * m]);  // W_N2^{r m'}
#pragma unroll
  for (int m = 0; m < 32; ++m) ex[m * CF<R1>::PITCH + r * 16 + c] = x[m];
}

// stage 2 for one m': y[r'] is output row m' + 32 r'
template <int R1>
__device__ __forceinline__ void cf_stage2(c32 (&y)[R1], const c32* ex, int c, int mp, bool inv) {
#pragma unroll
  for (int r = 0; r < R1; ++r) y[r] = ex[mp * CF<R1>::PITCH + r * 16 + c];
  dft<float, R1>(y);
  if (inv) {
#pragma unroll
    for (int r = 0; r < R1; ++r) y[r].y = -y[r].y;
  }
}

template <int N>
__device__ __forceinline__ void cf_load_twiddles(c32* twl, const c32* __restrict__ tw, int nthreads) {
  for (int i = threadIdx.x; i < N; i += nthreads) twl[i] = tw[i];
}

// ---------------------------------------------------------------- k1/k3 arguments

struct K1Args {
  const float* spec;  // Nf x Nt, column stride ldspec floats
  int64_t ldspec;
  c32* G;             // [Nt][N2][N1]
  const c32* tw2;     // W_N2^m
  int logN1;
  int64_t M;
  int ntiles;         // Nt * N1/16
};

struct K3Args {
  const c32* B;  // [Nr][N1/16][N2][16]
  const c32* yhalf;  // [Nr] bin M/2 term: B(j, k2 = 0, n1) += yhalf[j] (-1)^{n1}
  float* dest;   // Nf x Nr, column stride ld floats
  int64_t ld;
  const c32* tw2;
  int logN1;
  int Nr;
  int ntiles;    // Nr * N1/16
  double* part;  // NULL, or per-wavefront FDR statistics [Nr][N1/16][waves][3] (see k3_stats)
  int paired;    // k2 wrote the rates in mirrored pairs from the middle outwards
  // NULL, or the largest value of every run of FDR_RUN consecutive cells of
  // the FDR (Nr * Nf/FDR_RUN floats): the run maxima that let the proto-hit
  // kernels skip tiles under a threshold.  A run is the 16 complex outputs
  // n = n1 + N1 n2, n1 = 16 nb .. 16 nb + 15, of one column (cells 2n, 2n+1),
  // which 16 adjacent lanes of k3 store in one step; its maximum is
  // smax[j][nb][n2] ([Nr][N1/16][N2]).
  float* smax;
};

// Order-preserving map of a float onto an unsigned integer (0 orders below
// every float), and its inverse.
__device__ __forceinline__ unsigned fdr_ordered(float v) {
  const unsigned u = __float_as_uint(v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fdr_unordered(unsigned o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

// max(v, v of the lane that DPP control CTRL selects within the row of 16)
template <int CTRL>
__device__ __forceinline__ float dpp_max(float v) {
  return fmaxf(v, __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, false)));
}

// The maximum of v over each row of 16 lanes, in every lane of the row.
// Every lane of the wavefront must take part.  Row-local DPP moves (no LDS):
// lanes i^1, i^2 (quad permutations), then the mirrors within the half-row
// (quads 0 and 1) and the row (halves 0 and 1).
__device__ __forceinline__ float row16_max(float v) {
  v = dpp_max<0xb1>(v);   // quad_perm [1,0,3,2]
  v = dpp_max<0x4e>(v);   // quad_perm [2,3,0,1]
  v = dpp_max<0x141>(v);  // row_half_mirror
  return dpp_max<0x140>(v);  // row_mirror
}

// k3's i-th rate: the reverse of k2's write order, so the spectra written
// last are read first (Nr-1, Nr-2, ... or, for paired rates, Nr-1, 0, Nr-2, 1, ...)
__device__ __forceinline__ int k3_rate(int i, int Nr, int paired) {
  if (!paired) return Nr - 1 - i;
  return (i & 1) ? (i >> 1) : Nr - 1 - (i >> 1);
}

// Each wavefront of a k3 tile writes (S1, S2, c): the FP64 sums of (x - c) and
// (x - c)^2 over its 64 x 64 FDR values, c = lane 0's first value.  Each lane
// sums with its own first value as shift and is re-centred to c before the
// (fixed-order) wavefront reduction, so no sum cancels.
__device__ __forceinline__ void k3_stats(double* dst, double s1, double s2, double ct) {
  constexpr double n = 64.0;  // values per lane
  const double c = __shfl(ct, 0, 64);
  const double d = ct - c;
  double S1 = s1 + n * d;
  double S2 = s2 + d * (2.0 * s1 + n * d);
  for (int off = 32; off > 0; off >>= 1) {
    S1 += __shfl_xor(S1, off, 64);
    S2 += __shfl_xor(S2, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    dst[0] = S1;
    dst[1] = S2;
    dst[2] = c;
  }
}

__device__ __forceinline__ float fast_window(int wkind, const float* wcustom, int64_t k, int64_t Nf) {
  if (wkind == ZDT_WINDOW_HAMMING) {
    const float p = static_cast<float>(static_cast<double>(k) / static_cast<double>(Nf));
    return 0.53836f + 0.46164f * cis_rev(p).x;
  }
  if (wkind == ZDT_WINDOW_CUSTOM) return wcustom[k];
  return 1.0f;
}

__device__ __forceinline__ c64 cis64f(double rev) {
  const double f = rev - __builtin_rint(rev);
  double s, c;
  sincospi(2.0 * f, &s, &c);
  return {c, s};
}

// Horner: sum_u a[u] z^u
template <int NT>
__device__ __forceinline__ c32 horner(const c32* a, c32 z) {
  c32 acc = a[NT - 1];
#pragma unroll
  for (int u = NT - 2; u >= 0; --u) acc = cfma(a[u], acc, z);
  return acc;
}

// k2 occupancy target: two 512-thread blocks per CU (4 waves per SIMD) for
// N1 = 512, whose LDS (2 x 8 rows x 576 + 512 entries, 78 KB) allows it.
constexpr int k2_waves_per_eu(int n1) { return n1 >= 512 ? 4 : 2; }
// rows per group: as many as keep one butterfly per lane in the
// wavefront-level passes (8 rows for 512 = 8*8*8 over 8 wavefronts, ...)
constexpr int k2_rows(int n1) { return n1 == 256 ? 16 : 8; }

// The Hermitian pair q of column-pair block p (see k2_czt): bin
// kA = cA + N2 k1A in row rowA of the block's two rows and its partner
// M - kA = cB + N2 k1B in row rowB.  Block p > 0 holds columns p and N2-p;
// block 0 holds the self-paired columns 0 (q < N1/2, pairs k1 and N1-k1, the
// partner of q = 0 being the Nyquist bin M) and N2/2 (q >= N1/2).
struct HPair {
  int rowA, rowB, k1A, k1B, cA, cB;
  bool nyq;
};

template <int N1, int N2>
__device__ __forceinline__ HPair hpair(int p, int q) {
  HPair h;
  h.nyq = false;
  if (p != 0) {
    h.rowA = 0;
    h.rowB = 1;
    h.k1A = q;
    h.k1B = N1 - 1 - q;
    h.cA = p;
    h.cB = N2 - p;
  } else if (q < N1 / 2) {
    h.rowA = h.rowB = 0;
    h.k1A = q;
    h.k1B = N1 - q;
    h.nyq = (q == 0);
    h.cA = h.cB = 0;
  } else {
    h.rowA = h.rowB = 1;
    h.k1A = q - N1 / 2;
    h.k1B = N1 - 1 - h.k1A;
    h.cA = h.cB = N2 / 2;
  }
  return h;
}

// The Reinsch recurrence of one part (real or imaginary) of the
// coefficients, with the sign s = +-1 folded into the state:
// D~_t = s^t D_t and U~_t = s^t U_t turn
//   D_t = a_t + lambda U_{t+1} + s D_{t+1},  U_t = D_t + s U_{t+1}
// into
//   D~_t = D~_{t+1} + s^t a_t + (s lambda) U~_{t+1},  U~_t = D~_t + U~_{t+1},
// whose every update accumulates in place (v_add_f32 / v_fmac_f32, the VOP2
// forms that issue fastest); the values are those of the original recurrence
// up to the sign.  Returns D~_1 and U~_1.
template <int NT, typename A>
__device__ __forceinline__ void reinsch_part(const A& a, float s, float lamp, float& D, float& U) {
  D = ((NT - 1) & 1) ? s * a(NT - 1) : a(NT - 1);
  U = D;
#pragma unroll
  for (int t = NT - 2; t >= 1; --t) {
    if (t & 1) {
      D = fmaf(s, a(t), D);
    } else {
      D += a(t);
    }
    D = fmaf(lamp, U, D);
    U += D;
  }
}

// Sums of the coefficients (cr_t, ci_t) at z and conj(z), |z| = 1, by the
// Reinsch-modified Goertzel recurrence over the real and imaginary parts (see
// czt_reinsch_col): with s = sign(cos theta) and lambda = 2 (cos theta - s)
// from the FP64 z, sum a_t cos(t theta) = a_0 + (lambda/2) U_1 + s D_1 =
// a_0 + (s lambda/2) U~_1 + D~_1 and sum a_t sin(t theta) = U_1 sin theta =
// U~_1 s sin theta.
template <int NT, typename AR, typename AI>
__device__ __forceinline__ void reinsch_pair(const AR& ar, const AI& ai, c64 z, c32& Yz, c32& Yzc) {
  const float s = z.x >= 0.0 ? 1.0f : -1.0f;
  const float lamp = static_cast<float>(2.0 * (z.x - static_cast<double>(s))) * s;
  const float ssn = static_cast<float>(z.y) * s;
  float Dr, Ur, Di, Ui;
  reinsch_part<NT>(ar, s, lamp, Dr, Ur);
  reinsch_part<NT>(ai, s, lamp, Di, Ui);
  const float hl = 0.5f * lamp;
  const float Cr = fmaf(hl, Ur, ar(0)) + Dr;
  const float Ci = fmaf(hl, Ui, ai(0)) + Di;
  const float Sr = ssn * Ur, Si = ssn * Ui;
  Yz = {Cr - Si, Sr + Ci};
  Yzc = {Cr + Si, Ci - Sr};
}

// Two reinsch_pair evaluations, coefficients (ar, ai) at za and (br, bi) at
// zb, advanced term by term in one loop: the four recurrences are
// independent, so each term issues four independent VALU chains instead of
// two (a lone pair of dependent chains leaves the VALU waiting on results).
template <int NT, typename AR, typename AI, typename BR, typename BI>
__device__ __forceinline__ void reinsch_quad(const AR& ar, const AI& ai, c64 za, const BR& br, const BI& bi,
                                             c64 zb, c32& Ya, c32& Yac, c32& Yb, c32& Ybc) {
  const float sa = za.x >= 0.0 ? 1.0f : -1.0f;
  const float sb = zb.x >= 0.0 ? 1.0f : -1.0f;
  const float la = static_cast<float>(2.0 * (za.x - static_cast<double>(sa))) * sa;
  const float lb = static_cast<float>(2.0 * (zb.x - static_cast<double>(sb))) * sb;
  constexpr bool ODD = ((NT - 1) & 1) != 0;
  float Dar = ODD ? sa * ar(NT - 1) : ar(NT - 1);
  float Dai = ODD ? sa * ai(NT - 1) : ai(NT - 1);
  float Dbr = ODD ? sb * br(NT - 1) : br(NT - 1);
  float Dbi = ODD ? sb * bi(NT - 1) : bi(NT - 1);
  float Uar = Dar, Uai = Dai, Ubr = Dbr, Ubi = Dbi;
#pragma unroll
  for (int t = NT - 2; t >= 1; --t) {
    if (t & 1) {
      Dar = fmaf(sa, ar(t), Dar);
      Dai = fmaf(sa, ai(t), Dai);
      Dbr = fmaf(sb, br(t), Dbr);
      Dbi = fmaf(sb, bi(t), Dbi);
    } else {
      Dar += ar(t);
      Dai += ai(t);
      Dbr += br(t);
      Dbi += bi(t);
    }
    Dar = fmaf(la, Uar, Dar);
    Dai = fmaf(la, Uai, Dai);
    Dbr = fmaf(lb, Ubr, Dbr);
    Dbi = fmaf(lb, Ubi, Dbi);
    Uar += Dar;
    Uai += Dai;
    Ubr += Dbr;
    Ubi += Dbi;
  }
  {
    const float ssn = static_cast<float>(za.y) * sa, hl = 0.5f * la;
    const float Cr = fmaf(hl, Uar, ar(0)) + Dar, Ci = fmaf(hl, Uai, ai(0)) + Dai;
    const float Sr = ssn * Uar, Si = ssn * Uai;
    Ya = {Cr - Si, Sr + Ci};
    Yac = {Cr + Si, Ci - Sr};
  }
  {
    const float ssn = static_cast<float>(zb.y) * sb, hl = 0.5f * lb;
    const float Cr = fmaf(hl, Ubr, br(0)) + Dbr, Ci = fmaf(hl, Ubi, bi(0)) + Dbi;
    const float Sr = ssn * Ubr, Si = ssn * Ubi;
    Yb = {Cr - Si, Sr + Ci};
    Ybc = {Cr + Si, Ci - Sr};
  }
}

template <int NT>
__device__ __forceinline__ void reinsch2(const c32 (&c)[NT], c64 z, c32& Yz, c32& Yzc) {
  reinsch_pair<NT>([&](int t) { return c[t].x; }, [&](int t) { return c[t].y; }, z, Yz, Yzc);
}

template <int NT>
__device__ __forceinline__ void reinsch_split(const float (&cr)[NT], const float (&ci)[NT], c64 z, c32& Yz,
                                              c32& Yzc) {
  reinsch_pair<NT>([&](int t) { return cr[t]; }, [&](int t) { return ci[t]; }, z, Yz, Yzc);
}

}  // namespace zdt
