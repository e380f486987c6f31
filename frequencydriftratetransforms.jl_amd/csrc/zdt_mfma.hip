// Split-precision matrix-core variants of the column FFTs of k1 (input!) and
// k3 (output!) for N2 = 1024 = 32 x 32 (the frequency-axis transforms of
// src/zdtfdr.jl:119-120,211,330, four-step, see zdt_fast.hip).
//
// Both radix-32 stages of the 1024-point column DFT run as dense matrix
// products on the MFMA units:
//   Y (32 x 64 columns per wavefront) = E32 . X,  E32[m'][m] = e^{s 2 pi i m' m/32}
// in the real form  Yr = Er Xr - Ei Xi,  Yi = Ei Xr + Er Xi,  with every
// operand split into a high and a low part of 16 bits (x = hi + lo, hi the
// rounded x, lo the rounded x - hi) and the product taken as
// hi.hi + hi.lo + lo.hi (the lo.lo term, ~2^-16 relative, is dropped),
// accumulated in FP32 by v_mfma_f32_32x32x16_{bf16,f16}.  BF16 keeps the FP32
// exponent range; FP16 holds 3 more mantissa bits per part but needs range
// control, so each column is scaled by a power of two that brings its largest
// component to [2^13, 2^14) before the split, and unscaled (exactly) after.
//
// Operand layouts of the 32x32x16 MFMA (lane l, h = l / 32):
//   A (32 x 16): row l % 32, k = 8h + v (v < 8 elements per lane)
//   B (16 x 32): column l % 32, k = 8h + v
//   D (32 x 32): column l % 32, row 8 (v / 4) + 4h + v % 4 (v < 16)
// A wavefront owns 64 columns of a stage as two 32-column tiles; lane l holds
// column l % 32 of each tile at the 16 rows m = 8h + q (q < 8) and
// 16 + 8h + q - 8 (q >= 8), i.e. B fragments for K = rows 0..15 and 16..31.
// The twiddles between the stages, the bin M/2 term, the four-step twiddle
// of k3, the statistics and every load and store stay FP32 as in k1_input /
// k3_output; only the DFT-32 butterflies change arithmetic.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "zdt_fast_dev.h"
#include "zdt_fast_plan.h"

namespace zdt {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

template <int PREC>
struct Mfma;

template <>
struct Mfma<ZDT_PRECISION_SPLIT_BF16> {
  using V = bf16x8;
  using E = __bf16;
  static constexpr bool SCALED = false;
  __device__ static __forceinline__ f32x16 mma(V a, V b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
};

template <>
struct Mfma<ZDT_PRECISION_SPLIT_FP16> {
  using V = f16x8;
  using E = _Float16;
  static constexpr bool SCALED = true;
  __device__ static __forceinline__ f32x16 mma(V a, V b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  }
};

// x = hi + lo, each part rounded to nearest
template <int PREC>
__device__ __forceinline__ void split8(const float (&x)[8], typename Mfma<PREC>::V& hi, typename Mfma<PREC>::V& lo) {
  using E = typename Mfma<PREC>::E;
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    const E h = static_cast<E>(x[v]);
    hi[v] = h;
    lo[v] = static_cast<E>(x[v] - static_cast<float>(h));
  }
}

// sign flip of eight 16-bit values
template <typename V>
__device__ __forceinline__ V vneg(V v) {
  u32x4 u = __builtin_bit_cast(u32x4, v);
  u ^= 0x80008000u;
  return __builtin_bit_cast(V, u);
}

// E32 as A fragments in LDS, [fragment][lane]: fragment 4 f + {0, 1, 2, 3} =
// Er hi, Er lo, Ei hi, Ei lo for k = m in [16 f, 16 f + 16).  Wavefront 0
// writes the table; every wavefront reads lane l's entries in lane l, after
// the barrier that precedes the kernels' tile loop.
template <int PREC>
struct Dft32A {
  const typename Mfma<PREC>::V* t;  // LDS, 8 x 64 fragments
  int lane;
  __device__ __forceinline__ typename Mfma<PREC>::V get(int frag) const { return t[frag * 64 + lane]; }
};

template <int PREC, bool INV>
__device__ __forceinline__ Dft32A<PREC> dft32_a(typename Mfma<PREC>::V* tab, int lane, bool writer) {
  if (writer) {
    const int i = lane & 31, h = lane >> 5;
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      float cr[8], ci[8];
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int m = 16 * f + 8 * h + v;
        double s, c;
        sincospi(static_cast<double>((i * m) & 31) / 16.0, &s, &c);
        cr[v] = static_cast<float>(c);
        ci[v] = static_cast<float>(INV ? s : -s);
      }
      typename Mfma<PREC>::V a, b;
      split8<PREC>(cr, a, b);
      tab[(4 * f + 0) * 64 + lane] = a;
      tab[(4 * f + 1) * 64 + lane] = b;
      split8<PREC>(ci, a, b);
      tab[(4 * f + 2) * 64 + lane] = a;
      tab[(4 * f + 3) * 64 + lane] = b;
    }
  }
  return Dft32A<PREC>{tab, lane};
}

// Power of two that brings the column's largest component (lanes l and
// l ^ 32 hold the column) to [2^13, 2^14); 1 for a zero or non-finite column.
__device__ __forceinline__ float col_scale(const c32 (&x)[16]) {
  float m = 0.0f;
#pragma unroll
  for (int q = 0; q < 16; ++q) m = fmaxf(m, fmaxf(fabsf(x[q].x), fabsf(x[q].y)));
  m = fmaxf(m, __shfl_xor(m, 32, 64));
  if (!(m > 0.0f) || !(m <= 3.0e38f)) return 1.0f;
  int e;
  (void)frexpf(m, &e);  // m = f 2^e, f in [1/2, 1)
  return ldexpf(1.0f, std::min(14 - e, 100));
}

// One DFT-32 of a 32-column tile: x in the B layout (see the file comment),
// scaled by sc before the split; Y = E32 X in the D layout (unscaled by the
// caller).  24 MFMAs: 2 K-halves x (Yr: Er Xr, Ei (-Xi); Yi: Ei Xr, Er Xi) x 3
// split products, interleaved over the two accumulators.
template <int PREC>
__device__ __forceinline__ void dft32_tile(const c32 (&x)[16], const Dft32A<PREC>& A, float sc, f32x16& yr,
                                           f32x16& yi) {
  using M = Mfma<PREC>;
  using V = typename M::V;
  V xrh[2], xrl[2], xih[2], xil[2];
#pragma unroll
  for (int f = 0; f < 2; ++f) {
    float re[8], im[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      re[v] = M::SCALED ? x[8 * f + v].x * sc : x[8 * f + v].x;
      im[v] = M::SCALED ? x[8 * f + v].y * sc : x[8 * f + v].y;
    }
    split8<PREC>(re, xrh[f], xrl[f]);
    split8<PREC>(im, xih[f], xil[f]);
  }
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    yr[v] = 0.0f;
    yi[v] = 0.0f;
  }
#pragma unroll
  for (int f = 0; f < 2; ++f) {
    const V nih = vneg(xih[f]), nil = vneg(xil[f]);
    const V arh = A.get(4 * f), arl = A.get(4 * f + 1), aih = A.get(4 * f + 2), ail = A.get(4 * f + 3);
    yr = M::mma(arh, xrh[f], yr);
    yi = M::mma(aih, xrh[f], yi);
    yr = M::mma(arh, xrl[f], yr);
    yi = M::mma(aih, xrl[f], yi);
    yr = M::mma(arl, xrh[f], yr);
    yi = M::mma(ail, xrh[f], yi);
    yr = M::mma(aih, nih, yr);
    yi = M::mma(arh, xih[f], yi);
    yr = M::mma(aih, nil, yr);
    yi = M::mma(arh, xil[f], yi);
    yr = M::mma(ail, nih, yr);
    yi = M::mma(arl, xih[f], yi);
  }
}

// exchange buffer pitch (complex per m' slab of [r][c] = 512 entries): 528
// puts the two 16-lane halves of a stage-2 read (m' and m'+1) on disjoint banks
constexpr int MF_PITCH = 528;

__device__ __forceinline__ int mf_row(int q, int h) { return q < 8 ? 8 * h + q : 16 + 8 * h + (q - 8); }
__device__ __forceinline__ int mf_drow(int v, int h) { return 8 * (v >> 2) + 4 * h + (v & 3); }

// ---------------------------------------------------------------- k1 (input!)

template <int PREC>
__global__ __launch_bounds__(512, 1) void k1_input_mfma(K1Args a) {
  constexpr int N2 = 1024;
  __shared__ c32 ex[32 * MF_PITCH];
  __shared__ c32 twl[N2];
  __shared__ typename Mfma<PREC>::V atab[8 * 64];
  cf_load_twiddles<N2>(twl, a.tw2, 512);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  int h = lane >> 5, c = lane & 15, rr = (lane >> 4) & 1;
  const int N1 = 1 << a.logN1;
  const int nblk = N1 >> 4;
  int tile = blockIdx.x;
  if (tile >= a.ntiles) return;
  const Dft32A<PREC> A = dft32_a<PREC, false>(atab, lane, w == 0);
  auto load = [&](int tl, c32(&x)[2][16]) {
    const int t = tl / nblk, nb = tl - t * nblk;
    const c32* z = reinterpret_cast<const c32*>(a.spec + static_cast<int64_t>(t) * a.ldspec);
    const unsigned o = (nb << 4) + c;
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
      const unsigned r = 4 * w + 2 * tt + rr;
#pragma unroll
      for (int q = 0; q < 16; ++q) x[tt][q] = z[((r + 32u * mf_row(q, h)) << a.logN1) + o];
    }
  };
  c32 x[2][16];
  load(tile, x);
  __syncthreads();  // twiddle table
  for (;;) {
    asm volatile("" : "+v"(lane));
    h = lane >> 5;
    c = lane & 15;
    rr = (lane >> 4) & 1;
    const int t = tile / nblk, nb = tile - t * nblk;
    // stage 1: DFT-32 over m for the columns (r, c), times W_N2^{r m'}
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
      const float sc = Mfma<PREC>::SCALED ? col_scale(x[tt]) : 1.0f;
      const float isc = 1.0f / sc;
      __builtin_amdgcn_sched_barrier(0);
      f32x16 yr, yi;
      dft32_tile<PREC>(x[tt], A, sc, yr, yi);
      const int r = 4 * w + 2 * tt + rr;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int mp = mf_drow(v, h);
        const c32 z = {yr[v] * isc, yi[v] * isc};
        ex[mp * MF_PITCH + r * 16 + c] = cmul(z, twl[r * mp]);
      }
    }
    __syncthreads();
    // the next tile's loads fly while this one finishes
    const int next = tile + gridDim.x;
    if (next < a.ntiles) load(next, x);
    const unsigned n1 = (nb << 4) + c;
    c32* G = a.G + static_cast<int64_t>(t) * N2 * N1;
    // stage 2: DFT-32 over r for the columns (m', c) -> rows m' + 32 r'
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
      const int mp = 4 * w + 2 * tt + rr;
      c32 y[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) y[q] = ex[mp * MF_PITCH + mf_row(q, h) * 16 + c];
      const float sc = Mfma<PREC>::SCALED ? col_scale(y) : 1.0f;
      const float isc = 1.0f / sc;
      __builtin_amdgcn_sched_barrier(0);
      f32x16 yr, yi;
      dft32_tile<PREC>(y, A, sc, yr, yi);
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const unsigned rp = mf_drow(v, h);
        G[((mp + 32u * rp) << a.logN1) + n1] = {yr[v] * isc, yi[v] * isc};
      }
    }
    __syncthreads();
    tile = next;
    if (tile >= a.ntiles) break;
  }
}

// ---------------------------------------------------------------- k3 (output!)

// TMAX: also write the FDR's run maxima (K3Args::smax; a separate instance,
// so that the plain one keeps its registers)
template <int PREC, bool TMAX>
__global__ __launch_bounds__(512, 1) void k3_output_mfma(K3Args a) {
  constexpr int N2 = 1024;
  __shared__ c32 ex[32 * MF_PITCH];
  __shared__ c32 twl[N2];
  __shared__ typename Mfma<PREC>::V atab[8 * 64];
  cf_load_twiddles<N2>(twl, a.tw2, 512);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  int h = lane >> 5, c = lane & 15, rr = (lane >> 4) & 1;
  const int N1 = 1 << a.logN1;
  const int nblk = N1 >> 4;
  int tile = blockIdx.x;
  if (tile >= a.ntiles) return;
  const Dft32A<PREC> A = dft32_a<PREC, true>(atab, lane, w == 0);
  // rates last to first: the most recently written spectra are read first
  auto load = [&](int tl, c32(&x)[2][16], c32& hv) {
    const int jr = tl / nblk, nb = tl - jr * nblk;
    const int jj = k3_rate(jr, a.Nr, a.paired);
    const c32* b = a.B + static_cast<int64_t>(jj) * N2 * N1 + static_cast<int64_t>(nb) * N2 * 16;
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
      const unsigned r = 4 * w + 2 * tt + rr;
#pragma unroll
      for (int q = 0; q < 16; ++q) x[tt][q] = b[(r + 32u * mf_row(q, h)) * 16 + c];
    }
    hv = a.yhalf[jj];
  };
  c32 x[2][16], hv;
  load(tile, x, hv);
  __syncthreads();
  for (;;) {
    asm volatile("" : "+v"(lane));
    h = lane >> 5;
    c = lane & 15;
    rr = (lane >> 4) & 1;
    const int jr = tile / nblk, nb = tile - jr * nblk;
    const int j = k3_rate(jr, a.Nr, a.paired);
    const unsigned n1 = (nb << 4) + c;
    // the bin M/2 term enters row k2 = 0 (r = 0, m = 0)
    if (w == 0 && rr == 0 && h == 0) x[0][0] = (c & 1) ? csub(x[0][0], hv) : cadd(x[0][0], hv);
    {
      // four-step twiddle e^{2 pi i n1 k2/M}, k2 = r + 32 m: T[q] for
      // m = 8h + q, times b16 = e^{2 pi i n1 512/M} for m = 16 + 8h + q
      // (n1 k2 < M: every phase is an exact float)
      const float invM = 1.0f / static_cast<float>(N1 * N2);
      const c32 s1 = cis_rev(static_cast<float>(n1 * 32u) * invM);
      const c32 s2 = cmul(s1, s1), s3 = cmul(s2, s1);
      const c32 b16 = cis_rev(static_cast<float>(n1 * 512u) * invM);
#pragma unroll
      for (int tt = 0; tt < 2; ++tt) {
        const unsigned r = 4 * w + 2 * tt + rr;
        c32 T[8];
        T[0] = cis_rev(static_cast<float>(n1 * (r + 256u * h)) * invM);
        T[4] = cis_rev(static_cast<float>(n1 * (r + 256u * h + 128u)) * invM);
        T[1] = cmul(T[0], s1);
        T[2] = cmul(T[0], s2);
        T[3] = cmul(T[0], s3);
        T[5] = cmul(T[4], s1);
        T[6] = cmul(T[4], s2);
        T[7] = cmul(T[4], s3);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          x[tt][q] = cmul(x[tt][q], T[q]);
          x[tt][q + 8] = cmul(x[tt][q + 8], cmul(T[q], b16));
        }
      }
    }
    // stage 1: inverse DFT-32 over m for the columns (r, c), times W_N2^{-r m'}
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
      const float sc = Mfma<PREC>::SCALED ? col_scale(x[tt]) : 1.0f;
      const float isc = 1.0f / sc;
      __builtin_amdgcn_sched_barrier(0);
      f32x16 yr, yi;
      dft32_tile<PREC>(x[tt], A, sc, yr, yi);
      const int r = 4 * w + 2 * tt + rr;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int mp = mf_drow(v, h);
        const c32 z = {yr[v] * isc, yi[v] * isc};
        ex[mp * MF_PITCH + r * 16 + c] = cmulc(z, twl[r * mp]);
      }
    }
    __syncthreads();
    const int next = tile + gridDim.x;
    if (next < a.ntiles) load(next, x, hv);
    c32* out = reinterpret_cast<c32*>(a.dest + static_cast<int64_t>(j) * a.ld);
    double s1 = 0.0, s2 = 0.0, ct = 0.0;
    // stage 2: inverse DFT-32 over r for the columns (m', c) -> rows m' + 32 r'
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
      const int mp = 4 * w + 2 * tt + rr;
      c32 y[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) y[q] = ex[mp * MF_PITCH + mf_row(q, h) * 16 + c];
      const float sc = Mfma<PREC>::SCALED ? col_scale(y) : 1.0f;
      const float isc = 1.0f / sc;
      __builtin_amdgcn_sched_barrier(0);
      f32x16 yr, yi;
      dft32_tile<PREC>(y, A, sc, yr, yi);
      typedef float f2v __attribute__((ext_vector_type(2)));
      float keep = 0.0f;  // TMAX: run maximum of step v = c (lanes of a row share mp and h)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const float fr = yr[v] * isc, fi = yi[v] * isc;
        if (a.part != nullptr) {  // uniform
          if (tt == 0 && v == 0) ct = static_cast<double>(fr);
          const double d0 = static_cast<double>(fr) - ct, d1 = static_cast<double>(fi) - ct;
          s1 += d0 + d1;
          s2 += d0 * d0 + d1 * d1;
        }
        const unsigned rp = mf_drow(v, h);
        // the FDR is not read again (see k3_output)
        const f2v val = {fr, fi};
        __builtin_nontemporal_store(val, reinterpret_cast<f2v*>(out + ((mp + 32u * rp) << a.logN1) + n1));
        if constexpr (TMAX) {
          const float m = row16_max(fmaxf(fr, fi));
          keep = c == v ? m : keep;
        }
      }
      if constexpr (TMAX)  // the run n1 = 16 nb + c at n2 = mp + 32 rp: smax[j][nb][n2]
        a.smax[(static_cast<int64_t>(j) * (int64_t(1) << (a.logN1 - 4)) + nb) * N2 + mp + 32u * mf_drow(c, h)] = keep;
    }
    if (a.part != nullptr) {
      k3_stats(a.part + ((static_cast<int64_t>(j) * nblk + nb) * 8 + w) * 3, s1, s2, ct);
    }
    __syncthreads();
    tile = next;
    if (tile >= a.ntiles) break;
  }
}

// ---------------------------------------------------------------- launchers

bool mfma_supported(const FastPlan* p) { return p->N2 == 1024; }

int launch_k1_mfma(const FastPlan* p, const K1Args& a, hipStream_t st) {
  if (!mfma_supported(p)) return fail(ZDT_ENOTSUP, "split-precision column FFTs need N2 = 1024");
  const unsigned grid = static_cast<unsigned>(std::min<int64_t>(a.ntiles, p->resident_blocks / 2));
  if (p->precision == ZDT_PRECISION_SPLIT_BF16)
    k1_input_mfma<ZDT_PRECISION_SPLIT_BF16><<<grid, 512, 0, st>>>(a);
  else
    k1_input_mfma<ZDT_PRECISION_SPLIT_FP16><<<grid, 512, 0, st>>>(a);
  ZDT_CHECK_LAUNCH("k1_input_mfma");
  return ZDT_OK;
}

int launch_k3_mfma(const FastPlan* p, const K3Args& a, hipStream_t st) {
  if (!mfma_supported(p)) return fail(ZDT_ENOTSUP, "split-precision column FFTs need N2 = 1024");
  const unsigned grid = static_cast<unsigned>(std::min<int64_t>(a.ntiles, p->resident_blocks / 2));
  const bool bf16 = p->precision == ZDT_PRECISION_SPLIT_BF16;
  if (a.smax != nullptr) {
    if (bf16) k3_output_mfma<ZDT_PRECISION_SPLIT_BF16, true><<<grid, 512, 0, st>>>(a);
    else k3_output_mfma<ZDT_PRECISION_SPLIT_FP16, true><<<grid, 512, 0, st>>>(a);
  } else {
    if (bf16) k3_output_mfma<ZDT_PRECISION_SPLIT_BF16, false><<<grid, 512, 0, st>>>(a);
    else k3_output_mfma<ZDT_PRECISION_SPLIT_FP16, false><<<grid, 512, 0, st>>>(a);
  }
  ZDT_CHECK_LAUNCH("k3_output_mfma");
  return ZDT_OK;
}

}  // namespace zdt
