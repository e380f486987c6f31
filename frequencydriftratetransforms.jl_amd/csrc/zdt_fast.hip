// Fused ZDT path for power-of-two Nf (2^16 .. 2^20) and Nt <= 16.
//
// The frequency-axis transforms of the reference (rfft along dim 1,
// src/zdtfdr.jl:119,211, and irfft per rate, :120,330) are real transforms of
// length Nf done as complex transforms of length M = Nf/2 on the packed pairs
// z[n] = x[2n] + i x[2n+1], split four-step as M = N1*N2 (N2 = 1024, or 512 for Nf = 2^16):
//   k = k2 + N2*k1 (spectrum index),  n = n1 + N1*n2 (time/channel index).
// Three kernels, each touching HBM once per array:
//   k1_input   for each (t, n1): DFT_N2 over n2 of z[n1 + N1 n2] (strided
//              column FFT, 16 columns per block) -> G[t][k2][n1]  (input!)
//   k2_czt     one block per Hermitian column pair (k2, N2-k2): times
//              W_M^{n1 k2}, DFT_N1 over n1 -> Z[k] -> r2c untangle -> X[k,t] (= F of the reference) ->
//              chirp-Z over t for every rate (preprocess!/convolve!/
//              postprocess!) -> c2r pre-pass -> inverse DFT_N1 over k1
//              -> B[j][n1/16][k2][n1%16]
//   k3_output  for each (j, n1): times W_M^{-n1 k2}, inverse DFT_N2 over k2
//              -> z'[n1 + N1 n2] = FDR[2n, j] + i FDR[2n+1, j]  (output!)
//
// The chirp-Z transform per bin k is evaluated in its Vandermonde form
//   Ys[k,j] = w(k) sum_t a_t(k) gamma_k^{j t},  a_t = X[k,t] e^{2 pi i k t r0/Nf},
//   gamma_k = e^{2 pi i k dr/Nf}
// which equals the reference's chirp multiply / convolve / chirp multiply
// sequence (Bluestein's identity j t = (j^2 + t^2 - (j-t)^2)/2) for any Nl,
// by Horner's rule in FP32 at z = gamma_k^j, with z carried by an FP64
// recurrence.  A thread owns a bin k and its Hermitian partner M-k, for which
// gamma_{M-k}^j = e^{i pi dr j} conj(gamma_k^j), so the partner's powers cost
// one multiply by a per-rate table.  Symmetric rate sets take the paired
// evaluation of k2_czt<..., PAIRED> (reinsch2, zdt_fast_dev.h).
//
// The device building blocks live in zdt_fast_dev.h, the plan in
// zdt_fast_plan.h, and the fused output (Nt up to 64, Nf up to 2^22), which
// shares k1 and k3, in zdt_fout.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "fft_generic.h"
#include "zdt_fast_dev.h"
#include "zdt_fast_plan.h"
#include "zdt_kernels.h"

namespace zdt {

// ---------------------------------------------------------------- k1: input pass A

template <int R1>
__global__ __launch_bounds__(CF<R1>::THREADS, 512 / CF<R1>::THREADS) void k1_input(K1Args a) {
  using C = CF<R1>;
  __shared__ c32 ex[C::LDS];
  __shared__ c32 twl[C::N2];
  cf_load_twiddles<C::N2>(twl, a.tw2, C::THREADS);
  const int c = threadIdx.x & 15, r = threadIdx.x >> 4;
  const int N1 = 1 << a.logN1;
  const int nblk = N1 >> 4;
  int tile = blockIdx.x;
  if (tile >= a.ntiles) return;
  auto load = [&](int tl, c32(&x)[32]) {
    const int t = tl / nblk, nb = tl - t * nblk;
    const c32* z = reinterpret_cast<const c32*>(a.spec + static_cast<int64_t>(t) * a.ldspec);
    const unsigned o = (nb << 4) + c;
#pragma unroll
    for (int m = 0; m < 32; ++m) x[m] = z[(static_cast<unsigned>(r + R1 * m) << a.logN1) + o];
  };
  c32 x[32];
  load(tile, x);
  __syncthreads();  // twiddle table
  for (;;) {
    const int t = tile / nblk, nb = tile - t * nblk;
    cf_stage1<R1>(x, r, twl, ex, c, false);
    __syncthreads();
    // the next tile's loads fly while this one finishes
    const int next = tile + gridDim.x;
    if (next < a.ntiles) load(next, x);
    const unsigned n1 = (nb << 4) + c;
    c32* G = a.G + static_cast<int64_t>(t) * C::N2 * N1;  // uniform base, 32-bit offsets
#pragma unroll 1
    for (int h = 0; h < C::MPER; ++h) {
      const unsigned mp = r + R1 * h;
      c32 y[R1];
      cf_stage2<R1>(y, ex, c, mp, false);
#pragma unroll
      for (int rp = 0; rp < R1; ++rp) G[((mp + 32 * rp) << a.logN1) + n1] = y[rp];
    }
    __syncthreads();
    tile = next;
    if (tile >= a.ntiles) break;
  }
}

// ---------------------------------------------------------------- k3: output pass 2

// TMAX: also write the FDR's run maxima to a.smax (a separate instance, so
// that the plain one keeps its registers)
template <int R1, bool TMAX>
__global__ __launch_bounds__(CF<R1>::THREADS, 512 / CF<R1>::THREADS) void k3_output(K3Args a) {
  using C = CF<R1>;
  __shared__ c32 ex[C::LDS];
  __shared__ c32 twl[C::N2];
  cf_load_twiddles<C::N2>(twl, a.tw2, C::THREADS);
  const int c = threadIdx.x & 15, r = threadIdx.x >> 4;
  const int N1 = 1 << a.logN1;
  const int nblk = N1 >> 4;
  int tile = blockIdx.x;
  if (tile >= a.ntiles) return;
  // rates last to first: the most recently written spectra are read first
  auto load = [&](int tl, c32(&x)[32], c32& h) {
    const int jr = tl / nblk, nb = tl - jr * nblk;
    // the tile's 16 columns are one contiguous [k2][16] block of B
    const int jj = k3_rate(jr, a.Nr, a.paired);
    const c32* b = a.B + static_cast<int64_t>(jj) * C::N2 * N1 + static_cast<int64_t>(nb) * C::N2 * 16;
#pragma unroll
    for (int m = 0; m < 32; ++m) x[m] = b[static_cast<unsigned>(r + R1 * m) * 16 + c];
    h = a.yhalf[jj];
  };
  c32 x[32], h;
  load(tile, x, h);
  __syncthreads();
  for (;;) {
    const int jr = tile / nblk, nb = tile - jr * nblk;
    const int j = k3_rate(jr, a.Nr, a.paired);
    if (r == 0) x[0] = (c & 1) ? csub(x[0], h) : cadd(x[0], h);
    {
      // four-step twiddle W_M^{-n1 k2}, k2 = r + R1 m, as A[m mod 8] B[m / 8]
      // (n1 k2 < M: every phase is an exact float).  Four cis evaluations:
      // A[0], A[4], the step s = e^{2 pi i n1 R1/M} and B[1]; the rest are
      // products of at most three of them (a few ulp)
      const unsigned n1 = (nb << 4) + c;
      const float invM = 1.0f / static_cast<float>(N1 * C::N2);
      c32 A[8], Bq[4];
      A[0] = cis_rev(static_cast<float>(n1 * r) * invM);
      A[4] = cis_rev(static_cast<float>(n1 * (r + 4 * R1)) * invM);
      const c32 s1 = cis_rev(static_cast<float>(n1 * R1) * invM);
      const c32 s2 = cmul(s1, s1), s3 = cmul(s2, s1);
      A[1] = cmul(A[0], s1);
      A[2] = cmul(A[0], s2);
      A[3] = cmul(A[0], s3);
      A[5] = cmul(A[4], s1);
      A[6] = cmul(A[4], s2);
      A[7] = cmul(A[4], s3);
      Bq[0] = {1.0f, 0.0f};
      Bq[1] = cis_rev(static_cast<float>(n1 * (8 * R1)) * invM);
      Bq[2] = cmul(Bq[1], Bq[1]);
      Bq[3] = cmul(Bq[2], Bq[1]);
#pragma unroll
      for (int m = 1; m < 32; ++m) x[m] = cmul(x[m], (m < 8) ? A[m] : cmul(A[m & 7], Bq[m >> 3]));
      x[0] = cmul(x[0], A[0]);
    }
    cf_stage1<R1>(x, r, twl, ex, c, true);
    __syncthreads();
    // the next tile's loads fly while this one finishes
    const int next = tile + gridDim.x;
    if (next < a.ntiles) load(next, x, h);
    c32* out = reinterpret_cast<c32*>(a.dest + static_cast<int64_t>(j) * a.ld);
    const unsigned o = (nb << 4) + c;
    double s1 = 0.0, s2 = 0.0, ct = 0.0;
#pragma unroll
    for (int hh = 0; hh < C::MPER; ++hh) {
      const unsigned mp = r + R1 * hh;
      c32 y[R1];
      cf_stage2<R1>(y, ex, c, mp, true);
      if (a.part != nullptr) {  // uniform
        if (hh == 0) ct = static_cast<double>(y[0].x);
#pragma unroll
        for (int rp = 0; rp < R1; ++rp) {
          const double d0 = static_cast<double>(y[rp].x) - ct, d1 = static_cast<double>(y[rp].y) - ct;
          s1 += d0 + d1;
          s2 += d0 * d0 + d1 * d1;
        }
      }
      {
        // the FDR is not read again: nontemporal stores keep it from evicting
        // the next step's inputs (spectrogram, G) from the last-level cache
        typedef float f2v __attribute__((ext_vector_type(2)));
        f2v* q = reinterpret_cast<f2v*>(out + (mp << a.logN1) + o);
        const int64_t step = int64_t(32) << a.logN1;
        // run maxima (TMAX): the 16 lanes of a row store the run of outputs
        // n1 = 16 nb + c at n2 = mp + 32 rp; its maximum goes to
        // smax[j][nb][n2].  The row maximum of step rp is kept by lane
        // c = rp mod 16, which stores it after the loop, so that the block's
        // stores fill one contiguous run of N2 floats.
        float keep0 = 0.0f, keep1 = 0.0f;
#pragma unroll
        for (int rp = 0; rp < R1; ++rp) {
          const f2v v = {y[rp].x, y[rp].y};
          __builtin_nontemporal_store(v, q + rp * step);
          if constexpr (TMAX) {
            const float m = row16_max(fmaxf(y[rp].x, y[rp].y));
            if (rp < 16) keep0 = c == rp ? m : keep0;
            else keep1 = c == rp - 16 ? m : keep1;
          }
        }
        if constexpr (TMAX) {
          float* sm = a.smax + (static_cast<int64_t>(j) * (N1 >> 4) + nb) * C::N2;
          sm[mp + 32u * c] = keep0;
          if (R1 > 16) sm[mp + 32u * (c + 16)] = keep1;
        }
      }
    }
    if (a.part != nullptr) {
      constexpr int NW = C::THREADS / 64;
      k3_stats(a.part + ((static_cast<int64_t>(j) * nblk + nb) * NW + (threadIdx.x >> 6)) * 3, s1, s2, ct);
    }
    __syncthreads();
    tile = next;
    if (tile >= a.ntiles) break;
  }
}

// ---------------------------------------------------------------- k2: fused chirp-Z

struct K2Args {
  const c32* G;      // [Nt][N2][N1]
  c32* B;            // [Nr][N1/16][N2][16]
  const c32* tw1;    // W_N1^m
  const c32* tabq;   // e^{i pi dr j}, j < Nr
  const c64* tab64;  // the same in FP64 (paired rates)
  c64 eir0;          // e^{i pi (r0 - rc)} (paired rates)
  double hrot;       // rc/Nf: the coefficients are pre-rotated by e^{2 pi i k u rc/Nf}
  double rhrot;      // rc/2 (partner bin: e^{i pi u rc})
  double hz0;        // (r0 - rc)/Nf (paired rates: z of rate 0)
  c32* yhalf;        // [Nr] c2r pre-pass value of the bin M/2, added by k3
  int Nt, Nr;
  int logN1;
  int64_t M, Nf;
  double hd2;        // dr/Nf: gamma_k = e^{2 pi i k hd2}
  double r0half;     // r0/2:  e^{i pi u r0} = e^{2 pi i u r0half}
  int wkind;
  const float* wcustom;
  float oscale;      // output scale / M  (c2r normalization)
};

// Block of T = min(N1, 512) threads; each thread owns PP = N1/T Hermitian
// pairs q = tid + i*T (i < PP).  Pair q of a regular block (columns c, N2-c)
// is (k = c + N2 q, M - k = (N2-c) + N2 (N1-1-q)); the block for p = 0 holds the
// self-paired columns 0 and N2/2: q < N1/2 pairs (N2 q, N2 (N1-q)) in column 0
// (q = 0 pairs bin 0 with the Nyquist bin M), q >= N1/2 pairs within column
// N2/2.  The one bin left over, M/2 (column 0, k1 = N1/2), enters the c2r
// pre-pass at row 0, k1 = N1/2, hence B(j, k2 = 0, n1) as C_j (-1)^{n1}: block 0
// evaluates C_j for every rate up front into yhalf[j] and k3 adds it, which
// keeps its registers out of the rate loop.  The window weight and output
// scale are folded into the Vandermonde coefficients a_t.
//
// PAIRED: the coefficients are pre-rotated by the centre rate
// rc = r0 + (Nr-1) dr/2 instead of r0, so that the remaining rates
// r_j - rc = (j - (Nr-1)/2) dr are mirrored: rates j and Nr-1-j of a bin are
// its sums at z = e^{2 pi i k (r_j - rc)/Nf} and conj(z) and come from one
// reinsch2 (partner bin: z = e^{i pi (r_j - rc)} conj(z_k)); a group of
// ROWS/2 rates holds the pairs (j, Nr-1-j), (j+1, Nr-2-j), ...
template <int N1, int N2, int NT, int ROWS, bool PAIRED = false>
__global__ __launch_bounds__(N1 < 512 ? N1 : 512, k2_waves_per_eu(N1)) void k2_czt(K2Args a) {
  constexpr int T = N1 < 512 ? N1 : 512;
  constexpr int PP = N1 / T;
  // N1 = 512: rows in layout 1 (q64, the first pass's output at lsw) with
  // per-pass twiddle tables; other N1: rows padded one in eight (lsw),
  // twiddle table swizzled alike
  constexpr int LAY = N1 == 512 ? 1 : 0;
  constexpr int P = N1 + N1 / 8;  // row pitch
  constexpr int HALF = N1 / 2;
  auto ix = [](int i) { return row_index<LAY>(i); };
  __shared__ c32 buf[ROWS * P];
  __shared__ c32 buf2[ROWS * P];
  __shared__ c32 twl[N1 + N1 / 8];
  static_assert(LAY == 0 || LAY1_TW <= N1 + N1 / 8, "twiddle table size");
  if constexpr (LAY == 1) {
    lay1_load_twiddles(twl, a.tw1, T);
  } else {
    for (int i = threadIdx.x; i < N1; i += T) twl[lsw(i)] = a.tw1[i];  // W_N1^m at lsw(m)
  }
  __shared__ c32 ahalf[NT];  // a_t of the bin M/2 (block 0 only)

  const int tid = threadIdx.x;
  const int p = blockIdx.x;
  const bool special = (p == 0);
  const int colA = special ? 0 : p;
  const int colB = special ? N2 / 2 : N2 - p;

  // G rows of the input phase, loaded in batches of RB rows (row 2u + s:
  // time step t0 + u, column A for s = 0, B for s = 1).  The first batch of
  // each row group is loaded one group ahead (the first one here, before
  // the set-up arithmetic), so its latency overlaps the previous group's FFT.
  constexpr int TG = ROWS / 2;
  constexpr int RB = 8 / PP;
  auto load_batch = [&](int t0, int nt, int r0, c32 (&v)[RB][PP]) {
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      if (r0 + r < 2 * nt) {
        const int col = ((r0 + r) & 1) ? colB : colA;
        const c32* src = a.G + (static_cast<int64_t>(t0 + ((r0 + r) >> 1)) * N2 + col) * N1;
#pragma unroll
        for (int i = 0; i < PP; ++i) v[r][i] = src[tid + i * T];
      }
    }
  };
  c32 pf[RB][PP];
  load_batch(0, min(TG, a.Nt), 0, pf);

  int rowA[PP], rowB[PP], k1A[PP], k1B[PP];
  bool nyq[PP];
  int64_t kA[PP];
  c32 WA[PP];
  const float invNf = 1.0f / static_cast<float>(a.Nf);
#pragma unroll
  for (int i = 0; i < PP; ++i) {
    const HPair h = hpair<N1, N2>(p, tid + i * T);
    rowA[i] = h.rowA;
    rowB[i] = h.rowB;
    k1A[i] = h.k1A;
    k1B[i] = h.k1B;
    nyq[i] = h.nyq;
    kA[i] = h.cA + static_cast<int64_t>(N2) * k1A[i];
    WA[i] = cis_rev(-static_cast<float>(kA[i]) * invNf);  // W_Nf^{kA}
  }
  if (special && tid < NT) ahalf[tid] = {0.0f, 0.0f};  // rows u >= Nt stay zero

  // window weight, c2r scale and the 1/2 of the c2r pre-pass, folded into
  // the coefficients
  float wA[PP], wB[PP];
#pragma unroll
  for (int i = 0; i < PP; ++i) {
    wA[i] = fast_window(a.wkind, a.wcustom, kA[i], a.Nf) * (0.5f * a.oscale);
    wB[i] = fast_window(a.wkind, a.wcustom, a.M - kA[i], a.Nf) * (0.5f * a.oscale);
  }

  // four-step twiddles W_M^{-n1 k2} of the two columns (n1 = tid + i*T),
  // applied conjugated on input (k3 applies them on output)
  const float invM = 1.0f / static_cast<float>(a.M);
  c32 twS0[PP], twS1[PP];
#pragma unroll
  for (int i = 0; i < PP; ++i) {
    const int64_t n1 = tid + i * T;
    twS0[i] = cis_rev(static_cast<float>((n1 * colA) & (a.M - 1)) * invM);
    twS1[i] = cis_rev(static_cast<float>((n1 * colB) & (a.M - 1)) * invM);
  }

  // ---- input: W_M^{n1 k2}, DFT_N1 over n1, r2c untangle, pre-rotation -> aA, aB
  // (entries u >= Nt are zeroed after the loop, so that no register is held
  // for a time step before its group is transformed)
  c32 aA[PP][NT], aB[PP][NT];
#pragma unroll
  for (int t0 = 0; t0 < NT; t0 += TG) {
    if (t0 >= a.Nt) break;  // uniform
    const int nt = min(TG, a.Nt - t0);
#pragma unroll
    for (int r0 = 0; r0 < ROWS; r0 += RB) {
      c32 v[RB][PP];
      if (r0 == 0) {
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
          for (int i = 0; i < PP; ++i) v[r][i] = pf[r][i];
      } else {
        load_batch(t0, nt, r0, v);
      }
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        if (r0 + r < 2 * nt) {
#pragma unroll
          for (int i = 0; i < PP; ++i)
            buf[(r0 + r) * P + ix(tid + i * T)] = cmulc(v[r][i], ((r0 + r) & 1) ? twS1[i] : twS0[i]);
        }
      }
    }
    if (t0 + TG < NT && t0 + TG < a.Nt) load_batch(t0 + TG, min(TG, a.Nt - t0 - TG), 0, pf);  // one group ahead
    __syncthreads();
    const c32* fz = lds_fft_pp<N1, T, false, 1, true, LAY>(buf, buf2, P, 2 * nt, twl);
#pragma unroll
    for (int uu = 0; uu < TG; ++uu) {
      const int u = t0 + uu;
      if (u < NT && uu < nt) {
        const double ud = static_cast<double>(u);
#pragma unroll
        for (int i = 0; i < PP; ++i) {
          const c32 Za = fz[(2 * uu + rowA[i]) * P + ix(k1A[i])];
          c32 XA, XB;
          if (nyq[i]) {
            XA = {Za.x + Za.y, 0.0f};
            XB = {Za.x - Za.y, 0.0f};
          } else {
            const c32 Zb = fz[(2 * uu + rowB[i]) * P + ix(k1B[i])];
            const c32 E = {0.5f * (Za.x + Zb.x), 0.5f * (Za.y - Zb.y)};
            const c32 O = {0.5f * (Za.y + Zb.y), -0.5f * (Za.x - Zb.x)};
            const c32 WO = cmul(WA[i], O);
            XA = cadd(E, WO);
            XB = {E.x - WO.x, -(E.y - WO.y)};
          }
          // pre-rotation e^{2 pi i k u rc/Nf}; partner: e^{i pi u rc} conj(.)
          c32 ra = XA, rb = XB;
          if (a.hrot != 0.0) {  // uniform (rc = 0: rates symmetric about 0)
            const c32 rq = cis_nh(ud, a.rhrot);  // e^{i pi u rc}
            const c32 rot = cis_nh(static_cast<double>(kA[i]) * ud, a.hrot);
            ra = cmul(XA, rot);
            rb = cmul(XB, cmulc(rq, rot));
          }
          aA[i][u] = {ra.x * wA[i], ra.y * wA[i]};
          aB[i][u] = {rb.x * wB[i], rb.y * wB[i]};
        }
        if (special && tid == 0) {
          // bin M/2: X = conj(Z[M/2]); rotation e^{2 pi i (M/2) u r0/Nf} = e^{i pi u r0/2}
          const float wH = fast_window(a.wkind, a.wcustom, a.M / 2, a.Nf) * a.oscale;
          const c32 Zh = fz[(2 * uu) * P + ix(HALF)];
          const c32 h = cmul(cconj(Zh), cis_nh(ud, 0.5 * a.r0half));
          ahalf[u] = {h.x * wH, h.y * wH};
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < PP; ++i)
#pragma unroll
    for (int u = 0; u < NT; ++u) {
      if (u >= a.Nt) {
        aA[i][u] = {0.0f, 0.0f};
        aB[i][u] = {0.0f, 0.0f};
      }
    }

  if (special) {
    // bin M/2 (self-paired): gamma = e^{i pi dr/2}; its c2r pre-pass value is
    // conj(Ys) for every rate
    for (int j = tid; j < a.Nr; j += T) {
      const c32 zh = cis_nh(static_cast<double>(j), 0.25 * a.hd2 * static_cast<double>(a.Nf));
      c32 acc = ahalf[NT - 1];
      for (int u = NT - 2; u >= 0; --u) acc = cfma(ahalf[u], acc, zh);
      a.yhalf[j] = cconj(acc);
    }
  }

  if constexpr (PAIRED) {
    constexpr int JG = ROWS / 2;
    constexpr int NPG = JG / 2;  // rate pairs per group
    constexpr int NW = T / 64;
    constexpr int RPW = ROWS / NW;
    static_assert(RPW * NW == ROWS && RPW * fft_max_nr<N1>() <= 64, "one butterfly per lane and pass");
    const int wv = tid >> 6, lane = tid & 63;
    const int Pm = a.Nr - 1;        // rate j pairs with Pm - j
    const int npairs = Pm / 2 + 1;  // the last is the middle rate when Pm is even
    c64 gam[PP], zp[PP], zq[PP];
#pragma unroll
    for (int i = 0; i < PP; ++i) {
      gam[i] = cis64f(static_cast<double>(kA[i]) * a.hd2);
    }
    // groups from the middle rates outwards, so that the rates k3 reads
    // first (Nr-1, 0, Nr-2, 1, ...) are the ones written last
    const int g0 = ((npairs - 1) / NPG) * NPG;
#pragma unroll
    for (int i = 0; i < PP; ++i)  // e^{2 pi i k r_g0/Nf}
      zp[i] = cis64f(static_cast<double>(kA[i]) * (a.hz0 + static_cast<double>(g0) * a.hd2));
    c32* cur = buf;
    c32* nxt = buf2;
    for (int p0 = g0; p0 >= 0; p0 -= NPG) {
      const int np = min(NPG, npairs - p0);
#pragma unroll
      for (int i = 0; i < PP; ++i) zq[i] = zp[i];
#pragma unroll
      for (int pp = 0; pp < NPG; ++pp) {
        if (pp >= np) break;  // uniform
        const int jA = p0 + pp;
        const bool twin = Pm - jA != jA;
        const c64 q = cmul(a.eir0, a.tab64[jA]);  // e^{i pi (r_jA - rc)}
#pragma unroll
        for (int i = 0; i < PP; ++i) {
          const c64 zA = zq[i];
          zq[i] = cmul(zq[i], gam[i]);
          const c64 zB = cmul(q, c64{zA.x, -zA.y});
          c32 YA[2], YB[2];  // [0]: rate jA, [1]: rate Pm - jA
          reinsch_quad<NT>([&](int t) { return aA[i][t].x; }, [&](int t) { return aA[i][t].y; }, zA,
                           [&](int t) { return aB[i][t].x; }, [&](int t) { return aB[i][t].y; }, zB, YA[0], YA[1],
                           YB[0], YB[1]);
#pragma unroll
          for (int m = 0; m < 2; ++m) {
            if (m == 1 && !twin) break;
            const int slot = 2 * pp + m;
            if (nyq[i]) {
              cur[(2 * slot + rowA[i]) * P] = {YA[m].x + YB[m].x, YA[m].x - YB[m].x};
              cur[(2 * slot + rowA[i]) * P + ix(HALF)] = {0.0f, 0.0f};
            } else {
              const c32 E = {YA[m].x + YB[m].x, YA[m].y - YB[m].y};
              const c32 D = {YA[m].x - YB[m].x, YA[m].y + YB[m].y};
              const c32 O = cmulc(D, WA[i]);
              cur[(2 * slot + rowA[i]) * P + ix(k1A[i])] = {E.x - O.y, E.y + O.x};
              cur[(2 * slot + rowB[i]) * P + ix(k1B[i])] = {E.x + O.y, -E.y + O.x};
            }
          }
        }
      }
      __syncthreads();
      // slots 2pp, 2pp+1 hold rates p0+pp, Pm-p0-pp (the middle rate once, last)
      const int nslots = 2 * np - ((Pm & 1) == 0 && p0 + np == npairs ? 1 : 0);
      const int nrw = min(RPW, max(0, 2 * nslots - wv * RPW));  // uniform per wavefront
      wave_fft<N1, true, 1, true, LAY>(cur + wv * RPW * P, P, nrw, twl, lane, [&](int f, int jx, auto& v) {
        constexpr int RL = sizeof(v) / sizeof(c32);
        const int row = wv * RPW + f;
        const int slot = row >> 1;
        const int rate = (slot & 1) ? Pm - p0 - (slot >> 1) : p0 + (slot >> 1);
        c32* dst = a.B + static_cast<int64_t>(rate) * (N1 * N2) + ((row & 1) ? colB : colA) * 16;
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          const int n1 = jx + r * (N1 / RL);
          dst[(n1 >> 4) * (N2 * 16) + (n1 & 15)] = v[r];
        }
      });
      c32* t = cur;
      cur = nxt;
      nxt = t;
#pragma unroll
      for (int i = 0; i < PP; ++i) {  // back one group: conj(gamma)^NPG, formed here to spare registers
        c64 g = {gam[i].x, -gam[i].y};
#pragma unroll
        for (int s = 1; s < NPG; ++s) g = cmul(g, c64{gam[i].x, -gam[i].y});
        zp[i] = cmul(zp[i], g);
      }
    }
  } else {
    // ---- chirp-Z for every rate, c2r pre-pass, inverse DFT_N1, store
    c64 gam[PP], z[PP];
  #pragma unroll
    for (int i = 0; i < PP; ++i) {
      gam[i] = cis64f(static_cast<double>(kA[i]) * a.hd2);
      z[i] = {1.0, 0.0};
    }

    // Rate groups of JG rates (2 JG rows: columns A and B of each rate).  The
    // rows alternate between buf and buf2 from group to group, so the one
    // barrier per group (after the c2r pre-pass writes) also orders a group's
    // writes after the previous-but-one group's FFT reads.  Each wavefront
    // then runs the inverse DFT_N1 of its own rows in place and stores them.
    constexpr int JG = ROWS / 2;
    constexpr int NW = T / 64;      // wavefronts
    constexpr int RPW = ROWS / NW;  // rows per wavefront
    static_assert(RPW * NW == ROWS && RPW * fft_max_nr<N1>() <= 64, "one butterfly per lane and pass");
    const int wv = tid >> 6, lane = tid & 63;
    c32* cur = buf;
    c32* nxt = buf2;
    for (int j0 = 0; j0 < a.Nr; j0 += JG) {
      const int nj = min(JG, a.Nr - j0);
  #pragma unroll
      for (int jj = 0; jj < JG; ++jj) {
        if (jj >= nj) break;  // uniform
        const c32 qj = a.tabq[j0 + jj];
  #pragma unroll
        for (int i = 0; i < PP; ++i) {
          const c32 zA = {static_cast<float>(z[i].x), static_cast<float>(z[i].y)};
          z[i] = cmul(z[i], gam[i]);
          const c32 zB = cmulc(qj, zA);  // e^{i pi dr j} conj(zA)
          // YA, YB: Ys of bins kA and M - kA, halved
          const c32 YA = horner<NT>(aA[i], zA);
          const c32 YB = horner<NT>(aB[i], zB);
          if (nyq[i]) {
            // pair (0, M): only the real parts enter the c2r; the bin M/2 slot
            // of the row is zero here (k3 adds it)
            cur[(2 * jj + rowA[i]) * P] = {YA.x + YB.x, YA.x - YB.x};
            cur[(2 * jj + rowA[i]) * P + ix(HALF)] = {0.0f, 0.0f};
          } else {
            const c32 E = {YA.x + YB.x, YA.y - YB.y};  // (YA + conj YB)/2
            const c32 D = {YA.x - YB.x, YA.y + YB.y};  // (YA - conj YB)/2
            const c32 O = cmulc(D, WA[i]);  // W^{-k} D
            cur[(2 * jj + rowA[i]) * P + ix(k1A[i])] = {E.x - O.y, E.y + O.x};
            cur[(2 * jj + rowB[i]) * P + ix(k1B[i])] = {E.x + O.y, -E.y + O.x};  // conj(E) + i conj(O)
          }
        }
      }
      __syncthreads();
      const int nrw = min(RPW, max(0, 2 * nj - wv * RPW));  // uniform per wavefront
      wave_fft<N1, true, 1, true, LAY>(cur + wv * RPW * P, P, nrw, twl, lane, [&](int f, int jx, auto& v) {
        constexpr int RL = sizeof(v) / sizeof(c32);
        const int row = wv * RPW + f;
        // B is blocked by 16 n1: B[j][n1/16][k2][n1%16]
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
}

// ---------------------------------------------------------------- F unpack

// Gt[t][k2][n1] = G[t][k2][n1] W_M^{n1 k2} (the four-step twiddle k2 applies
// on load), for the generic DFT_N1 that rebuilds F
__global__ __launch_bounds__(256) void twiddle_G(const c32* __restrict__ G, c32* __restrict__ Gt, int64_t total,
                                                 int logN1, int logN2) {
  const int64_t M = int64_t(1) << (logN1 + logN2);
  const float invM = 1.0f / static_cast<float>(M);
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t n1 = g & ((int64_t(1) << logN1) - 1);
    const int64_t k2 = (g >> logN1) & ((int64_t(1) << logN2) - 1);
    Gt[g] = cmul(G[g], cis_rev(-static_cast<float>((n1 * k2) & (M - 1)) * invM));
  }
}

// natural F[k, t] (k <= M) from Z[t][k2][k1] (the DFT_N1 of Gt over n1)
__global__ __launch_bounds__(256) void unpack_F(const c32* __restrict__ Z, c32* __restrict__ F, int Nt, int logN1,
                                                int logN2, int64_t M, int64_t Nf) {
  const int64_t Nk = M + 1;
  const int64_t total = Nk * Nt;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t k = g % Nk;
    const int t = static_cast<int>(g / Nk);
    const c32* Zt = Z + static_cast<int64_t>(t) * M;
    auto at = [&](int64_t kk) {  // Z[kk] for kk in [0, M)
      const int64_t k2 = kk & ((int64_t(1) << logN2) - 1), k1 = kk >> logN2;
      return Zt[(k2 << logN1) + k1];
    };
    c32 X;
    if (k == 0 || k == M) {
      const c32 z0 = at(0);
      X = {k == 0 ? z0.x + z0.y : z0.x - z0.y, 0.0f};
    } else {
      const c32 Za = at(k), Zb = at(M - k);
      const c32 E = {0.5f * (Za.x + Zb.x), 0.5f * (Za.y - Zb.y)};
      const c32 O = {0.5f * (Za.y + Zb.y), -0.5f * (Za.x - Zb.x)};
      double p = -static_cast<double>(k) / static_cast<double>(Nf);
      p -= __builtin_rint(p);
      X = cadd(E, cmul(cis_rev(static_cast<float>(p)), O));
    }
    F[k + static_cast<int64_t>(t) * Nk] = X;
  }
}

}  // namespace zdt

// ---------------------------------------------------------------- host side

namespace zdt {


int ilog2i(int64_t n) {
  int l = 0;
  while ((int64_t(1) << l) < n) ++l;
  return l;
}

bool fast_supported(int64_t Nf, int64_t Nt, int64_t Nr) {
  if (Nf < (1 << 16) || Nf > (1 << 20) || (Nf & (Nf - 1)) != 0) return false;
  if (Nt < 1 || Nt > 16) return false;
  if (Nr < 1 || Nr > 65535) return false;
  return true;
}

int64_t fast_bytes(int64_t Nf, int64_t Nt, int64_t Nr) {
  const int64_t M = Nf / 2;
  const int64_t N2 = fast_n2(Nf);
  const int64_t N1 = M / N2;
  return 8 * M * (Nt + Nr) + 8 * (N1 + N2 + 2 * Nr) + 8 * fast_stats_doubles(Nf, Nr);
}

int upload_table(c32** dst, const std::vector<c32>& h) {
  hipError_t e = hipMalloc(reinterpret_cast<void**>(dst), h.size() * sizeof(c32));
  if (e != hipSuccess) return fail(ZDT_ENOMEM, "hipMalloc(table) failed");
  ZDT_TRY_HIP(hipMemcpy(*dst, h.data(), h.size() * sizeof(c32), hipMemcpyHostToDevice));
  return ZDT_OK;
}

std::vector<c32> twiddles(int N) {
  std::vector<c32> t(N);
  for (int m = 0; m < N; ++m) {
    const long double a = -2.0L * 3.14159265358979323846264338327950288L * m / N;
    t[m] = {static_cast<float>(cosl(a)), static_cast<float>(sinl(a))};
  }
  return t;
}

void alloc_part(FastPlan* p, int& rc) {
  if (rc != ZDT_OK) return;
  const size_t n = static_cast<size_t>(fast_stats_doubles(p->Nf, p->Nr));
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&p->part), n * sizeof(double));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    p->part = nullptr;
    rc = fail(ZDT_ENOMEM, "hipMalloc(%zu bytes) failed: %s", n * sizeof(double), hipGetErrorString(e));
  }
}

void fast_destroy(FastPlan* p) {
  if (p == nullptr) return;
  for (void* q : {static_cast<void*>(p->G), static_cast<void*>(p->B), static_cast<void*>(p->tw1),
                  static_cast<void*>(p->tw2), static_cast<void*>(p->tabq), static_cast<void*>(p->tab64), static_cast<void*>(p->yhalf),
                  static_cast<void*>(p->Yc), static_cast<void*>(p->yM), static_cast<void*>(p->part)})
    if (q) (void)hipFree(q);
  delete p;
}

bool fast_n2_ok(int64_t Nf, int n2) {
  if (n2 == 0) return true;
  if (n2 != 512 && n2 != 1024) return false;
  const int64_t N1 = Nf / 2 / n2;
  // compiled (N1, N2) pairs of k2_czt
  return (n2 == 1024 && N1 >= 64 && N1 <= 512) || (n2 == 512 && N1 >= 64 && N1 <= 512);
}

bool fast_precision_ok(int64_t Nf, int n2, int precision) {
  if (precision == ZDT_PRECISION_FP32) return true;
  if (precision != ZDT_PRECISION_SPLIT_BF16 && precision != ZDT_PRECISION_SPLIT_FP16) return false;
  // the matrix-core column FFTs are the 32 x 32 split of N2 = 1024
  return (n2 ? n2 : fast_n2(Nf)) == 1024;
}

int fast_create(int64_t Nf, int64_t Nt, int64_t Nr, float dr, int n2, int precision, FastPlan** out) {
  *out = nullptr;
  if (!fast_supported(Nf, Nt, Nr)) return fail(ZDT_ENOTSUP, "fused path: unsupported shape");
  if (!fast_n2_ok(Nf, n2)) return fail(ZDT_ENOTSUP, "fused path: N2=%d not available for Nf=%lld", n2, (long long)Nf);
  if (!fast_precision_ok(Nf, n2, precision))
    return fail(ZDT_ENOTSUP, "fused path: precision %d not available for Nf=%lld", precision, (long long)Nf);
  FastPlan* p = new FastPlan();
  p->precision = precision;
  p->Nf = Nf;
  p->Nt = Nt;
  p->Nr = Nr;
  p->M = Nf / 2;
  p->N2 = n2 ? n2 : fast_n2(Nf);
  p->logN2 = ilog2i(p->N2);
  p->N1 = static_cast<int>(p->M / p->N2);
  p->logN1 = ilog2i(p->N1);
  p->NT = Nt <= 2 ? 2 : Nt <= 4 ? 4 : Nt <= 8 ? 8 : 16;
  p->dr = dr;
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
  alloc(&p->G, static_cast<size_t>(Nt * p->M));
  alloc(&p->B, static_cast<size_t>(Nr * p->M));
  alloc_part(p, rc);
  alloc(&p->yhalf, static_cast<size_t>(Nr));
  if (rc == ZDT_OK) rc = upload_table(&p->tw1, twiddles(p->N1));
  if (rc == ZDT_OK) rc = upload_table(&p->tw2, twiddles(p->N2));
  if (rc == ZDT_OK) {
    std::vector<c32> q(Nr);
    const long double pi = 3.14159265358979323846264338327950288L;
    for (int64_t j = 0; j < Nr; ++j) {
      // e^{i pi dr j}, reduced exactly in long double
      long double rev = 0.5L * static_cast<long double>(dr) * static_cast<long double>(j);
      rev -= floorl(rev);
      q[j] = {static_cast<float>(cosl(2 * pi * rev)), static_cast<float>(sinl(2 * pi * rev))};
    }
    rc = upload_table(&p->tabq, q);
    if (rc == ZDT_OK) {
      std::vector<c64> q64(Nr);
      for (int64_t j = 0; j < Nr; ++j) {
        long double rev = 0.5L * static_cast<long double>(dr) * static_cast<long double>(j);
        rev -= floorl(rev);
        q64[j] = {static_cast<double>(cosl(2 * pi * rev)), static_cast<double>(sinl(2 * pi * rev))};
      }
      hipError_t e = hipMalloc(reinterpret_cast<void**>(&p->tab64), q64.size() * sizeof(c64));
      if (e != hipSuccess) {
        (void)hipGetLastError();
        p->tab64 = nullptr;
        rc = fail(ZDT_ENOMEM, "hipMalloc(table) failed");
      } else {
        e = hipMemcpy(p->tab64, q64.data(), q64.size() * sizeof(c64), hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = fail(ZDT_EHIP, "table upload failed: %s", hipGetErrorString(e));
      }
    }
  }
  if (rc != ZDT_OK) {
    fast_destroy(p);
    return rc;
  }
  *out = p;
  return ZDT_OK;
}

int launch_k1(const FastPlan* p, const float* spec, c32* G, int64_t Nt, hipStream_t st) {
  K1Args a;
  a.spec = spec;
  a.ldspec = p->Nf;
  a.G = G;
  a.tw2 = p->tw2;
  a.logN1 = p->logN1;
  a.M = p->M;
  a.ntiles = static_cast<int>(Nt * (p->N1 / 16));
  ProfScope ps(st, "k1_input", 4.0 * p->Nf * Nt + 8.0 * p->M * Nt);
  if (p->precision != ZDT_PRECISION_FP32) return launch_k1_mfma(p, a, st);
  if (p->N2 == 1024) {
    const unsigned grid = static_cast<unsigned>(std::min<int64_t>(a.ntiles, p->resident_blocks / 2));
    k1_input<32><<<grid, CF<32>::THREADS, 0, st>>>(a);
  } else {
    const unsigned grid = static_cast<unsigned>(std::min<int64_t>(a.ntiles, p->resident_blocks));
    k1_input<16><<<grid, CF<16>::THREADS, 0, st>>>(a);
  }
  ZDT_CHECK_LAUNCH("k1_input");
  return ZDT_OK;
}

int fast_input(FastPlan* p, const float* spec, hipStream_t st) { return launch_k1(p, spec, p->G, p->Nt, st); }

template <int N1, int N2>
static int launch_k2(const K2Args& a, int NT, bool paired, hipStream_t st) {
  constexpr unsigned grid = N2 / 2;
  constexpr unsigned TB = N1 < 512 ? N1 : 512;
  constexpr int RW = k2_rows(N1);
  if (paired && NT == 16) {
    k2_czt<N1, N2, 16, RW, true><<<grid, TB, 0, st>>>(a);
    ZDT_CHECK_LAUNCH("k2_czt");
    return ZDT_OK;
  }
  switch (NT) {
    case 2:
      k2_czt<N1, N2, 2, RW><<<grid, TB, 0, st>>>(a);
      break;
    case 4:
      k2_czt<N1, N2, 4, RW><<<grid, TB, 0, st>>>(a);
      break;
    case 8:
      k2_czt<N1, N2, 8, RW><<<grid, TB, 0, st>>>(a);
      break;
    default:
      k2_czt<N1, N2, 16, RW><<<grid, TB, 0, st>>>(a);
      break;
  }
  ZDT_CHECK_LAUNCH("k2_czt");
  return ZDT_OK;
}

int fast_run(FastPlan* p, float r0, int wkind, const float* window, float* dest, double scale, hipStream_t st,
             double* colmean, double* colvar) {
  K2Args a;
  a.G = p->G;
  a.B = p->B;
  a.tw1 = p->tw1;
  a.tabq = p->tabq;
  a.tab64 = p->tab64;
  // the paired evaluation (compiled for NT = 16) around the centre rate
  const bool paired = p->Nr > 1 && p->NT == 16;
  p->paired = paired;
  const double rc = paired ? static_cast<double>(r0) + 0.5 * static_cast<double>(p->Nr - 1) * p->dr
                           : static_cast<double>(r0);
  {
    const double h = 0.5 * (static_cast<double>(r0) - rc);  // e^{i pi (r0 - rc)}
    a.eir0 = {std::cos(2.0 * M_PI * (h - std::rint(h))), std::sin(2.0 * M_PI * (h - std::rint(h)))};
  }
  a.hrot = rc / static_cast<double>(p->Nf);
  a.rhrot = 0.5 * rc;
  a.hz0 = (static_cast<double>(r0) - rc) / static_cast<double>(p->Nf);
  a.yhalf = p->yhalf;
  a.Nt = static_cast<int>(p->Nt);
  a.Nr = static_cast<int>(p->Nr);
  a.logN1 = p->logN1;
  a.M = p->M;
  a.Nf = p->Nf;
  a.hd2 = static_cast<double>(p->dr) / static_cast<double>(p->Nf);
  a.r0half = 0.5 * static_cast<double>(r0);
  a.wkind = wkind;
  a.wcustom = window;
  a.oscale = static_cast<float>(scale / static_cast<double>(p->M));
  {
    ProfScope ps(st, "k2_czt", 8.0 * p->M * (p->Nt + p->Nr));
    const int key = p->N1 * 2048 + p->N2;
    switch (key) {
      case 64 * 2048 + 512:
        ZDT_TRY((launch_k2<64, 512>(a, p->NT, paired, st)));
        break;
      case 64 * 2048 + 1024:
        ZDT_TRY((launch_k2<64, 1024>(a, p->NT, paired, st)));
        break;
      case 128 * 2048 + 512:
        ZDT_TRY((launch_k2<128, 512>(a, p->NT, paired, st)));
        break;
      case 256 * 2048 + 512:
        ZDT_TRY((launch_k2<256, 512>(a, p->NT, paired, st)));
        break;
      case 512 * 2048 + 512:
        ZDT_TRY((launch_k2<512, 512>(a, p->NT, paired, st)));
        break;
      case 128 * 2048 + 1024:
        ZDT_TRY((launch_k2<128, 1024>(a, p->NT, paired, st)));
        break;
      case 256 * 2048 + 1024:
        ZDT_TRY((launch_k2<256, 1024>(a, p->NT, paired, st)));
        break;
      case 512 * 2048 + 1024:
        ZDT_TRY((launch_k2<512, 1024>(a, p->NT, paired, st)));
        break;
      default:
        return fail(ZDT_ENOTSUP, "fused path: N1=%d N2=%d", p->N1, p->N2);
    }
  }
  p->have_B = true;
  if (dest != nullptr) ZDT_TRY(fast_output(p, dest, scale, st, colmean, colvar));
  return ZDT_OK;
}

// Column mean and corrected variance from k3's per-wavefront partials, one
// block per column, merged in a fixed order around the column's first
// partial shift (the same statistics as colstats, fdrstats of
// src/fdrutils.jl:39-50).
__global__ __launch_bounds__(256) void k3_stats_merge(const double* __restrict__ part, int np, int64_t Nf,
                                                      double* __restrict__ colmean, double* __restrict__ colvar) {
  const int64_t j = blockIdx.x;
  const double* pj = part + j * np * 3;
  const double C0 = pj[2];
  const double n = static_cast<double>(Nf) / np;  // values per partial
  double S1 = 0.0, S2 = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) {
    const double s1 = pj[3 * i], s2 = pj[3 * i + 1], d = pj[3 * i + 2] - C0;
    S1 += s1 + n * d;
    S2 += s2 + d * (2.0 * s1 + n * d);
  }
  for (int off = 32; off > 0; off >>= 1) {
    S1 += __shfl_xor(S1, off, 64);
    S2 += __shfl_xor(S2, off, 64);
  }
  __shared__ double sh[2][4];
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = S1;
    sh[1][threadIdx.x >> 6] = S2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double T1 = (sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]);
    const double T2 = (sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3]);
    const double N = static_cast<double>(Nf);
    colmean[j] = C0 + T1 / N;
    colvar[j] = Nf > 1 ? (T2 - T1 * T1 / N) / (N - 1.0) : 0.0 / 0.0;
  }
}

static int k3_waves(const FastPlan* p) { return (p->N2 == 1024 ? CF<32>::THREADS : CF<16>::THREADS) / 64; }

// (N1/16) tiles x k3 wavefronts per rate is M/2048 for both N2 = 1024
// (8 wavefronts) and N2 = 512 (4 wavefronts)
int64_t fast_stats_doubles(int64_t Nf, int64_t Nr) { return Nr * (Nf / 2 / 2048) * 3; }

void fast_set_smax(FastPlan* p, float* smax) {
  if (p != nullptr) p->smax = smax;
}

int fast_plan_n2(const FastPlan* p) { return p != nullptr ? p->N2 : 0; }

int fast_output(FastPlan* p, float* dest, double scale, hipStream_t st, double* colmean, double* colvar) {
  (void)scale;  // folded into k2
  if (!p->have_B) return fail(ZDT_ESTATE, "output!: no post-processed spectra");
  if (colmean != nullptr && p->part == nullptr) return fail(ZDT_ESTATE, "fused statistics: no partials buffer");
  K3Args a;
  a.B = p->B;
  a.yhalf = p->yhalf;
  a.dest = dest;
  a.ld = p->Nf;
  a.tw2 = p->tw2;
  a.logN1 = p->logN1;
  a.Nr = static_cast<int>(p->Nr);
  a.ntiles = static_cast<int>(p->Nr * (p->N1 / 16));
  a.part = colmean != nullptr ? p->part : nullptr;
  a.paired = p->paired ? 1 : 0;
  a.smax = p->smax;
  {
    ProfScope ps(st, "k3_output", 8.0 * p->M * p->Nr + 4.0 * p->Nf * p->Nr);
    if (p->precision != ZDT_PRECISION_FP32) {
      ZDT_TRY(launch_k3_mfma(p, a, st));
    } else if (p->N2 == 1024) {
      const unsigned grid = static_cast<unsigned>(std::min<int64_t>(a.ntiles, p->resident_blocks / 2));
      if (a.smax) k3_output<32, true><<<grid, CF<32>::THREADS, 0, st>>>(a);
      else k3_output<32, false><<<grid, CF<32>::THREADS, 0, st>>>(a);
    } else {
      const unsigned grid = static_cast<unsigned>(std::min<int64_t>(a.ntiles, p->resident_blocks));
      if (a.smax) k3_output<16, true><<<grid, CF<16>::THREADS, 0, st>>>(a);
      else k3_output<16, false><<<grid, CF<16>::THREADS, 0, st>>>(a);
    }
    ZDT_CHECK_LAUNCH("k3_output");
  }
  if (colmean != nullptr) {
    const int np = (p->N1 / 16) * k3_waves(p);
    k3_stats_merge<<<static_cast<unsigned>(p->Nr), 256, 0, st>>>(p->part, np, p->Nf, colmean, colvar);
    ZDT_CHECK_LAUNCH("k3_stats_merge");
  }
  return ZDT_OK;
}

int fast_unpack_F(FastPlan* p, c32* F, hipStream_t st) {
  const FftPlan* fp = nullptr;
  ZDT_TRY(fft_plan_get(p->N1, false, &fp));
  const size_t n = static_cast<size_t>(p->Nt * p->M);
  void *Z = nullptr, *s0 = nullptr, *s1 = nullptr, *Gt = nullptr;
  int rc = ZDT_OK;
  if (hipMallocAsync(&Z, n * sizeof(c32), st) != hipSuccess || hipMallocAsync(&s0, n * sizeof(c32), st) != hipSuccess ||
      hipMallocAsync(&s1, n * sizeof(c32), st) != hipSuccess || hipMallocAsync(&Gt, n * sizeof(c32), st) != hipSuccess) {
    (void)hipGetLastError();
    rc = fail(ZDT_ENOMEM, "F unpack: allocation failed");
  }
  if (rc == ZDT_OK) {
    int64_t g = (static_cast<int64_t>(n) + 255) / 256;
    if (g > (1 << 20)) g = 1 << 20;
    twiddle_G<<<static_cast<unsigned>(g), 256, 0, st>>>(p->G, static_cast<c32*>(Gt), static_cast<int64_t>(n), p->logN1,
                                                         p->logN2);
    if (hipGetLastError() != hipSuccess) rc = fail(ZDT_EHIP, "twiddle_G launch failed");
  }
  // DFT_N1 over n1 (contiguous rows of Gt): Z[t][k2][k1]
  if (rc == ZDT_OK)
    rc = fft_c2c(fp, static_cast<c32*>(Gt), 1, p->N1, Z, 1, p->N1, static_cast<int64_t>(p->Nt) * p->N2, false, 1.0, s0,
                 s1, st);
  if (rc == ZDT_OK) {
    const int64_t total = (p->M + 1) * p->Nt;
    int64_t g = (total + 255) / 256;
    if (g > (1 << 20)) g = 1 << 20;
    unpack_F<<<static_cast<unsigned>(g), 256, 0, st>>>(static_cast<const c32*>(Z), F, static_cast<int>(p->Nt),
                                                        p->logN1, p->logN2, p->M, p->Nf);
    if (hipGetLastError() != hipSuccess) rc = fail(ZDT_EHIP, "unpack_F launch failed");
  }
  if (Z) (void)hipFreeAsync(Z, st);
  if (s0) (void)hipFreeAsync(s0, st);
  if (s1) (void)hipFreeAsync(s1, st);
  if (Gt) (void)hipFreeAsync(Gt, st);
  return rc;
}

}  // namespace zdt
