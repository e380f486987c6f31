// Generic batched mixed-radix FFT for gfx950 (FP32 and FP64).
//
// This engine serves every transform whose size or layout the fused ZDT path
// does not cover: the step functions' length-Nl FFT along dim 2 of the Nk x Nl
// work matrix (src/zdtfdr.jl:122-123,186,250-252), the chirp-kernel spectrum V
// (computed in FP64, src/zdtfdr.jl:173-187), and the frequency-axis r2c/c2r of
// the generic path (src/zdtfdr.jl:119-120,211,330) for any Nf whose prime
// factors are in {2,3,5,7,11,13}.
//
// Algorithm: Stockham auto-sort, one kernel per radix pass, in the
// formulation of Govindaraju et al., "High performance discrete Fourier
// transforms on graphics processors" (SC'08): pass with radix R over the
// current sub-transform length Ns reads x[j + r*N/R], applies the twiddles
// W_{Ns*R}^{r*(j mod Ns)}, does a length-R DFT and writes
// y[(j/Ns)*Ns*R + (j mod Ns) + r*Ns].  Twiddles come from a per-N FP64-derived
// table W_N^m.  The inverse is conj(forward(conj(x))).  The thread order is
// chosen so that consecutive lanes touch consecutive addresses: along j when
// the transform axis is contiguous, along the batch otherwise.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

#include "fft_butterflies.h"
#include "fft_generic.h"

namespace zdt {

// ---------------------------------------------------------------- pass kernel

struct PassArgs {
  int64_t N, NR, Ns, twstep;  // twstep = N/(Ns*R)
  int64_t ies, ibs, oes, obs, batch;
  double scale;
  int inverse, jmajor;
};

template <typename T, int R>
__global__ __launch_bounds__(256) void stockham_pass(const cplx<T>* __restrict__ in,
                                                     cplx<T>* __restrict__ out,
                                                     const cplx<T>* __restrict__ tw, PassArgs a) {
  const int64_t total = a.NR * a.batch;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    int64_t j, b;
    if (a.jmajor) {
      j = g % a.NR;
      b = g / a.NR;
    } else {
      b = g % a.batch;
      j = g / a.batch;
    }
    const int64_t k = j % a.Ns;
    cplx<T> v[R];
    const cplx<T>* src = in + b * a.ibs;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      v[r] = src[(j + r * a.NR) * a.ies];
      if (a.inverse) v[r].y = -v[r].y;
    }
    if (a.Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[r * k * a.twstep]);
    }
    dft<T, R>(v);
    const T sc = static_cast<T>(a.scale);
    const int64_t idxD = (j / a.Ns) * a.Ns * R + k;
    cplx<T>* dst = out + b * a.obs;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      cplx<T> y = v[r];
      if (a.inverse) y.y = -y.y;
      if (a.scale != 1.0) y = {y.x * sc, y.y * sc};
      dst[(idxD + r * a.Ns) * a.oes] = y;
    }
  }
}

static unsigned grid_for(int64_t total, int threads) {
  int64_t g = (total + threads - 1) / threads;
  if (g > (1 << 20)) g = 1 << 20;
  if (g < 1) g = 1;
  return static_cast<unsigned>(g);
}

template <typename T>
static int launch_pass(int R, const cplx<T>* in, cplx<T>* out, const cplx<T>* tw, const PassArgs& a,
                       hipStream_t st) {
  const unsigned grid = grid_for(a.NR * a.batch, 256);
  static const char* names[17] = {nullptr, nullptr, "fft_pass_r2", "fft_pass_r3", "fft_pass_r4", "fft_pass_r5",
                                  nullptr, "fft_pass_r7", "fft_pass_r8", nullptr, nullptr, "fft_pass_r11",
                                  nullptr, "fft_pass_r13", nullptr, nullptr, "fft_pass_r16"};
  ProfScope ps(st, (R >= 2 && R <= 16 && names[R]) ? names[R] : "fft_pass",
               2.0 * static_cast<double>(a.N * a.batch) * sizeof(cplx<T>));
#define ZDT_PASS_CASE(RR)                                                  \
  case RR:                                                                 \
    stockham_pass<T, RR><<<grid, 256, 0, st>>>(in, out, tw, a);            \
    break;
  switch (R) {
    ZDT_PASS_CASE(2)
    ZDT_PASS_CASE(3)
    ZDT_PASS_CASE(4)
    ZDT_PASS_CASE(5)
    ZDT_PASS_CASE(7)
    ZDT_PASS_CASE(8)
    ZDT_PASS_CASE(11)
    ZDT_PASS_CASE(13)
    ZDT_PASS_CASE(16)
    default:
      return fail(ZDT_ENOTSUP, "internal: radix %d not implemented", R);
  }
#undef ZDT_PASS_CASE
  ZDT_CHECK_LAUNCH("stockham_pass");
  return ZDT_OK;
}

template <typename T>
__global__ __launch_bounds__(256) void scale_copy(const cplx<T>* __restrict__ in, int64_t ibs,
                                                  cplx<T>* __restrict__ out, int64_t obs, int64_t batch,
                                                  T scale) {
  for (int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; b < batch;
       b += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const cplx<T> v = in[b * ibs];
    out[b * obs] = {v.x * scale, v.y * scale};
  }
}

// ---------------------------------------------------------------- four-step transforms
//
// For FP32 transforms along a contiguous axis with N = N1 * N2 (N1 <= 1024,
// N2 <= 2048, N <= 2^20) the passes are grouped into two kernels that each
// move the data through HBM once (the "four-step" FFT, D. H. Bailey, "FFTs in
// external or hierarchical memory", 1990):
//   fs_cols  X1[n2 + N2 k1] = W_N^{n2 k1} sum_{n1} x[n2 + N2 n1] W_N1^{n1 k1}
//            (TC adjacent columns n2 per block, so loads and stores are
//            TC-element runs)
//   fs_rows  X[k1 + N1 k2] = sum_{n2} X1[n2 + N2 k1] W_N2^{n2 k2}
//            (TR adjacent rows k1 per block; the stores are TR-element runs)
// Inside a block the sub-transforms run as Stockham passes on one LDS buffer
// of up to FS_E elements: each pass reads its butterflies' inputs into
// registers, synchronizes, and writes the outputs in place.  Twiddles are
// W_N^m = thi[m >> 10] * tlo[m & 1023] from two 1024-entry tables held in LDS
// (the sub-transform twiddle W_L^e is W_N^{e N/L}).  The inverse is
// conj(forward(conj(x))): fs_cols conjugates on load, fs_rows on store.

constexpr int FS_T = 512;    // threads per block
constexpr int FS_E = 8192;   // elements per block (TC * N1 or TR * N2)
constexpr int FS_LGB = 10;   // table split
constexpr int FS_TAB = 1 << FS_LGB;
constexpr int FS_MAXP = 12;  // passes per sub-transform

struct FsArgs {
  const c32* in;
  int64_t ibs;
  c32* out;
  int64_t obs;
  const c32* tlo;
  const c32* thi;
  int N, N1, N2, T;  // T = TC (fs_cols) or TR (fs_rows)
  int nhi;           // entries of thi
  int np;            // passes of the sub-transform
  int rad[FS_MAXP];
  int inverse;
  float scale;
  float pscale;  // fs_cols<true>: scale of the c2r pre-pass
  c32 w2;        // fs_cols<true>: W_{2N}^1
};

__device__ __forceinline__ c32 fs_tw(const c32* tl, const c32* th, int m) {
  return cmul(th[m >> FS_LGB], tl[m & (FS_TAB - 1)]);
}

// u / d for u * d < 2^32 by one high multiply with m = ceil(2^32 / d)
// (m = 0 marks d = 1)
struct FsDiv {
  uint32_t d, m;
};
__device__ __forceinline__ FsDiv fs_div(int d) {
  const uint32_t ud = static_cast<uint32_t>(d);
  return {ud, ud == 1 ? 0u : static_cast<uint32_t>(0xFFFFFFFFull / ud + 1)};
}
__device__ __forceinline__ int fs_quo(int u, FsDiv q) {
  return q.m == 0 ? u : static_cast<int>(__umulhi(static_cast<uint32_t>(u), q.m));
}

// One Stockham pass of radix R over T interleaved transforms of length L
// held in buf (element n of transform c at n * T + c, T a power of two).
// Butterfly u covers transform c = u mod T and index j = u / T < L/R.  Every
// index is below 2^13, so j mod Ns is an exact high multiply.
template <int R>
__device__ __forceinline__ void fs_pass(c32* buf, const c32* tl, const c32* th, int L, int T, int Ns, int mstep) {
  constexpr int NB = (FS_E / R + FS_T - 1) / FS_T;
  const int Lr = L / R;
  const int nbf = Lr * T;
  const int lgT = __ffs(T) - 1;
  const FsDiv dNs = fs_div(Ns);
  // opaque to the optimizer, so that no index below is hoisted out of the
  // pass loop (which would keep every pass's indices live and spill)
  int tid = static_cast<int>(threadIdx.x);
  asm volatile("" : "+v"(tid));
  c32 v[NB][R];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int u = tid + i * FS_T;
    if (u < nbf) {
      const int c = u & (T - 1), j = u >> lgT;
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = buf[(j + r * Lr) * T + c];
      const int k = j - fs_quo(j, dNs) * Ns;
      if (k != 0) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[i][r] = cmul(v[i][r], fs_tw(tl, th, r * k * mstep));
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int u = tid + i * FS_T;
    if (u < nbf) {
      const int c = u & (T - 1), j = u >> lgT;
      const int k = j - fs_quo(j, dNs) * Ns;
      dft<float, R>(v[i]);
      const int ob = ((j - k) * R + k) * T + c;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[ob + r * Ns * T] = v[i][r];
    }
  }
  __syncthreads();
}

// all passes of the T interleaved sub-transforms of length L (natural order in
// and out)
__device__ __forceinline__ void fs_dft(c32* buf, const c32* tl, const c32* th, const FsArgs& a, int L) {
  int Ns = 1;
  for (int p = 0; p < a.np; ++p) {
    const int R = a.rad[p];
    const int mstep = a.N / (Ns * R);  // W_{Ns R}^{r k} = W_N^{r k mstep}
    switch (R) {
      case 2: fs_pass<2>(buf, tl, th, L, a.T, Ns, mstep); break;
      case 3: fs_pass<3>(buf, tl, th, L, a.T, Ns, mstep); break;
      case 4: fs_pass<4>(buf, tl, th, L, a.T, Ns, mstep); break;
      case 5: fs_pass<5>(buf, tl, th, L, a.T, Ns, mstep); break;
      case 7: fs_pass<7>(buf, tl, th, L, a.T, Ns, mstep); break;
      case 8: fs_pass<8>(buf, tl, th, L, a.T, Ns, mstep); break;
      case 11: fs_pass<11>(buf, tl, th, L, a.T, Ns, mstep); break;
      case 13: fs_pass<13>(buf, tl, th, L, a.T, Ns, mstep); break;
      default: fs_pass<16>(buf, tl, th, L, a.T, Ns, mstep); break;
    }
    Ns *= R;
  }
}

__device__ __forceinline__ void fs_tables(c32* tl, c32* th, const FsArgs& a) {
  for (int i = threadIdx.x; i < FS_TAB; i += FS_T) tl[i] = a.tlo[i];
  for (int i = threadIdx.x; i < a.nhi; i += FS_T) th[i] = a.thi[i];
}

// columns: block = (batch b, columns n2 in [c0, c0 + TC)); TC is a power of two.
// C2R: the input is the half spectrum X[0..N] of a real sequence of length 2N
// (column stride ibs) and the kernel forms the packed sequence of the c2r
// transform on load (as c2r_pre): Z[k] = (E + i O) pscale with
// E = (X[k] + conj X[N-k])/2, O = W_{2N}^{-k} (X[k] - conj X[N-k])/2,
// W_{2N}^k = W_N^{k/2} (W_{2N}^1 for odd k); Im X[0] and Im X[N] are ignored.
template <bool C2R>
__global__ __launch_bounds__(FS_T, 4) void fs_cols(FsArgs a) {
  __shared__ c32 buf[FS_E];
  __shared__ c32 tl[FS_TAB], th[FS_TAB];
  fs_tables(tl, th, a);
  const int TC = a.T;
  const int lgT = __ffs(TC) - 1;
  const int tiles = (a.N2 + TC - 1) / TC;
  const int64_t b = blockIdx.x / tiles;
  int q = static_cast<int>(blockIdx.x % tiles);
  if (C2R && (tiles & 15) == 0) {
    // tile i and its mirror tiles-1-i (which reads the same columns of X)
    // eight blocks apart: same XCD (blocks go round-robin over the eight),
    // in flight together, so the second read of a column hits in L2
    const int g = q >> 4, x = q & 15;
    q = x < 8 ? g * 8 + x : tiles - 1 - (g * 8 + x - 8);
  }
  const int c0 = q * TC;
  const int E = a.N1 * TC;
  if constexpr (!C2R) {
    const c32* src = a.in + b * a.ibs + c0;
    // all loads issued before the first LDS store
    c32 v[FS_E / FS_T];
#pragma unroll
    for (int i = 0; i < FS_E / FS_T; ++i) {
      const int e = static_cast<int>(threadIdx.x) + i * FS_T;
      const int c = e & (TC - 1), n = e >> lgT;
      v[i] = {0.0f, 0.0f};
      if (e < E && c0 + c < a.N2) v[i] = src[c + static_cast<int64_t>(n) * a.N2];
    }
#pragma unroll
    for (int i = 0; i < FS_E / FS_T; ++i) {
      const int e = static_cast<int>(threadIdx.x) + i * FS_T;
      if (e < E) buf[e] = {v[i].x, a.inverse ? -v[i].y : v[i].y};  // element n of column c at n * TC + c
    }
  } else {
    const c32* X = a.in + b * a.ibs;
    c32 p[FS_E / FS_T], q[FS_E / FS_T];
#pragma unroll
    for (int i = 0; i < FS_E / FS_T; ++i) {
      const int e = static_cast<int>(threadIdx.x) + i * FS_T;
      const int c = e & (TC - 1), n = e >> lgT;
      p[i] = q[i] = {0.0f, 0.0f};
      if (e < E && c0 + c < a.N2) {
        const int k = c0 + c + n * a.N2;
        p[i] = X[k];
        q[i] = X[a.N - k];  // X[N] for k = 0
      }
    }
    __syncthreads();  // tables
#pragma unroll
    for (int i = 0; i < FS_E / FS_T; ++i) {
      const int e = static_cast<int>(threadIdx.x) + i * FS_T;
      const int c = e & (TC - 1), n = e >> lgT;
      const int k = c0 + c + n * a.N2;
      c32 z;
      if (k == 0) {
        z = {0.5f * (p[i].x + q[i].x), 0.5f * (p[i].x - q[i].x)};
      } else {
        const c32 Ev = {0.5f * (p[i].x + q[i].x), 0.5f * (p[i].y - q[i].y)};
        const c32 D = {0.5f * (p[i].x - q[i].x), 0.5f * (p[i].y + q[i].y)};
        c32 w = fs_tw(tl, th, k >> 1);
        if (k & 1) w = cmul(w, a.w2);
        const c32 O = cmulc(D, w);
        z = {Ev.x - O.y, Ev.y + O.x};
      }
      if (e < E) buf[e] = {z.x * a.pscale, -z.y * a.pscale};  // the inverse: conjugated
    }
  }
  __syncthreads();
  fs_dft(buf, tl, th, a, a.N1);
  c32* dst = a.out + b * static_cast<int64_t>(a.N) + c0;
  int tid = static_cast<int>(threadIdx.x);
  asm volatile("" : "+v"(tid));  // keeps the store indices from being computed before the passes
#pragma unroll
  for (int i = 0; i < FS_E / FS_T; ++i) {
    const int e = tid + i * FS_T;
    const int c = e & (TC - 1), k1 = e >> lgT;
    if (e < E && c0 + c < a.N2) dst[c + static_cast<int64_t>(k1) * a.N2] = cmul(buf[e], fs_tw(tl, th, (c0 + c) * k1));
  }
}

// rows: block = (batch b, rows k1 in [r0, r0 + TR)); TR is a power of two
__global__ __launch_bounds__(FS_T, 4) void fs_rows(FsArgs a) {
  __shared__ c32 buf[FS_E];
  __shared__ c32 tl[FS_TAB], th[FS_TAB];
  fs_tables(tl, th, a);
  const int TR = a.T;
  const int tiles = (a.N1 + TR - 1) / TR;
  const int64_t b = blockIdx.x / tiles;
  const int r0 = static_cast<int>(blockIdx.x % tiles) * TR;
  const int nrow = min(TR, a.N1 - r0);
  const c32* src = a.in + b * static_cast<int64_t>(a.N) + static_cast<int64_t>(r0) * a.N2;
  const int E = a.N2 * TR;
  const int lgT = __ffs(TR) - 1;
  {
    // rows interleaved as in fs_cols (element n of row c at n * TR + c), which
    // keeps the passes' LDS accesses as conflict-free as fs_cols's
    c32 v[FS_E / FS_T];
#pragma unroll
    for (int i = 0; i < FS_E / FS_T; ++i) {
      const int e = static_cast<int>(threadIdx.x) + i * FS_T;
      const int c = e & (TR - 1), n = e >> lgT;
      v[i] = {0.0f, 0.0f};
      if (e < E && c < nrow) v[i] = src[static_cast<int64_t>(c) * a.N2 + n];
    }
#pragma unroll
    for (int i = 0; i < FS_E / FS_T; ++i) {
      const int e = static_cast<int>(threadIdx.x) + i * FS_T;
      if (e < E) buf[e] = v[i];
    }
  }
  __syncthreads();
  fs_dft(buf, tl, th, a, a.N2);
  c32* dst = a.out + b * a.obs + r0;
  int tid = static_cast<int>(threadIdx.x);
  asm volatile("" : "+v"(tid));  // keeps the store indices from being computed before the passes
#pragma unroll
  for (int i = 0; i < FS_E / FS_T; ++i) {
    const int e = tid + i * FS_T;
    const int c = e & (TR - 1), k2 = e >> lgT;
    if (e < E && c < nrow) {
      c32 v = buf[e];
      v = {v.x * a.scale, (a.inverse ? -v.y : v.y) * a.scale};
      dst[c + static_cast<int64_t>(k2) * a.N1] = v;
    }
  }
}

// c2r (c2r = true): in is the half spectrum X[0..N] (column stride ibs) of real
// sequences of length 2N, out their values as N complex pairs, scaled by
// `scale` / N (the normalized inverse times scale).
static int fs_c2c(const FftPlan* p, const c32* in, int64_t ibs, c32* out, int64_t obs, int64_t batch, bool inverse,
                  double scale, c32* s0, hipStream_t st, bool c2r = false) {
  FsArgs a;
  a.tlo = p->tlo;
  a.thi = p->thi;
  a.N = static_cast<int>(p->N);
  a.N1 = p->N1;
  a.N2 = p->N2;
  a.nhi = static_cast<int>((p->N + FS_TAB - 1) / FS_TAB);
  a.inverse = inverse ? 1 : 0;
  const double bytes = 16.0 * static_cast<double>(p->N) * static_cast<double>(batch);
  {
    a.in = in;
    a.ibs = ibs;
    a.out = s0;
    a.obs = p->N;
    a.T = p->TC;
    a.np = static_cast<int>(p->rad1.size());
    for (int i = 0; i < a.np; ++i) a.rad[i] = p->rad1[i];
    a.scale = 1.0f;
    a.pscale = static_cast<float>(scale / static_cast<double>(p->N));
    const double ph = -3.14159265358979323846 / static_cast<double>(p->N);
    a.w2 = {static_cast<float>(std::cos(ph)), static_cast<float>(std::sin(ph))};
    const int64_t grid = (p->N2 + p->TC - 1) / p->TC * batch;
    if (c2r) {
      ProfScope ps(st, "fft_fs_cols_c2r", bytes + 8.0 * static_cast<double>(p->N) * static_cast<double>(batch));
      fs_cols<true><<<static_cast<unsigned>(grid), FS_T, 0, st>>>(a);
    } else {
      ProfScope ps(st, "fft_fs_cols", bytes);
      fs_cols<false><<<static_cast<unsigned>(grid), FS_T, 0, st>>>(a);
    }
    ZDT_CHECK_LAUNCH("fs_cols");
  }
  {
    a.in = s0;
    a.ibs = p->N;
    a.out = out;
    a.obs = obs;
    a.T = p->TR;
    a.np = static_cast<int>(p->rad2.size());
    for (int i = 0; i < a.np; ++i) a.rad[i] = p->rad2[i];
    a.scale = c2r ? 1.0f : static_cast<float>(scale);
    const int64_t grid = (p->N1 + p->TR - 1) / p->TR * batch;
    ProfScope ps(st, "fft_fs_rows", bytes);
    fs_rows<<<static_cast<unsigned>(grid), FS_T, 0, st>>>(a);
    ZDT_CHECK_LAUNCH("fs_rows");
  }
  return ZDT_OK;
}

static int pow2_floor(int64_t x) {
  int r = 1;
  while (static_cast<int64_t>(r) * 2 <= x) r *= 2;
  return r;
}

// Chooses the four-step split of an FP32 plan (sets p->N1 = 0 if none applies):
// the one whose column and row blocks both move the longest contiguous runs.
static void fs_choose(FftPlan* p, int maxr) {
  p->N1 = 0;
  const int64_t N = p->N;
  if (p->dbl || N < 4096 || N > (int64_t{1} << 20)) return;
  int best = -1;
  for (int64_t n1 = 2; n1 <= 1024; ++n1) {
    if (N % n1 != 0) continue;
    const int64_t n2 = N / n1;
    if (n2 < 2 || n2 > 2048) continue;
    std::vector<int> r1, r2;
    if (!fft_factorize(n1, &r1, maxr) || !fft_factorize(n2, &r2, maxr)) continue;
    if (static_cast<int>(r1.size()) > FS_MAXP || static_cast<int>(r2.size()) > FS_MAXP) continue;
    const int tc = std::min(16, pow2_floor(FS_E / n1));
    const int tr = std::min(16, pow2_floor(FS_E / n2));
    const int score = std::min(tc, tr) * 100 + tc;
    if (score > best) {
      best = score;
      p->N1 = static_cast<int>(n1);
      p->N2 = static_cast<int>(n2);
      p->TC = tc;
      p->TR = tr;
      p->rad1 = r1;
      p->rad2 = r2;
    }
  }
}

// ---------------------------------------------------------------- plans

bool fft_factorize(int64_t N, std::vector<int>* radices, int maxr) {
  radices->clear();
  if (N < 1) return false;
  int64_t n = N;
  for (int r : {16, 8, 4, 2, 3, 5, 7, 11, 13}) {
    if ((r & (r - 1)) == 0 && r > maxr) continue;  // power-of-two radices above the cap
    while (n % r == 0) {
      radices->push_back(r);
      n /= r;
    }
  }
  return n == 1;
}

namespace {
struct PlanKey {
  int device;
  int64_t N;
  bool dbl;
  int maxr;
  bool operator<(const PlanKey& o) const {
    return std::tie(device, N, dbl, maxr) < std::tie(o.device, o.N, o.dbl, o.maxr);
  }
};
std::mutex g_plan_mu;
std::map<PlanKey, FftPlan*> g_plans;
}  // namespace

int fft_plan_get(int64_t N, bool dbl, const FftPlan** out, int maxr) {
  std::vector<int> radices;
  if (maxr != 16 && maxr != 8 && maxr != 4) return fail(ZDT_EINVAL, "FFT radix cap must be 4, 8 or 16");
  if (!fft_factorize(N, &radices, maxr))
    return fail(ZDT_ENOTSUP, "FFT length %lld has a prime factor outside {2,3,5,7,11,13}",
                static_cast<long long>(N));
  int dev = 0;
  ZDT_TRY_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_plan_mu);
  PlanKey key{dev, N, dbl, maxr};
  auto it = g_plans.find(key);
  if (it != g_plans.end()) {
    *out = it->second;
    return ZDT_OK;
  }
  FftPlan* p = new FftPlan();
  p->N = N;
  p->dbl = dbl;
  p->radices = radices;
  // twiddle table W_N^m = exp(-2 pi i m / N), m < N, from an exact reduction
  const size_t esz = dbl ? sizeof(c64) : sizeof(c32);
  std::vector<unsigned char> host(static_cast<size_t>(N) * esz);
  for (int64_t m = 0; m < N; ++m) {
    // angle = 2 pi m/N; use the symmetric quarter to keep cos/sin exact at
    // multiples of pi/2
    const long double a = 2.0L * 3.14159265358979323846264338327950288L * static_cast<long double>(m) /
                          static_cast<long double>(N);
    double c = static_cast<double>(cosl(a));
    double s = -static_cast<double>(sinl(a));
    if (4 * m == N || 4 * m == 3 * N) c = 0.0;
    if (2 * m == N) s = 0.0;
    if (dbl) {
      reinterpret_cast<c64*>(host.data())[m] = {c, s};
    } else {
      reinterpret_cast<c32*>(host.data())[m] = {static_cast<float>(c), static_cast<float>(s)};
    }
  }
  hipError_t e = hipMalloc(&p->tw, host.size());
  if (e != hipSuccess) {
    delete p;
    return fail(ZDT_ENOMEM, "hipMalloc of FFT twiddles (%zu bytes) failed", host.size());
  }
  e = hipMemcpy(p->tw, host.data(), host.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(p->tw);
    delete p;
    return hip_fail(e, "hipMemcpy(twiddles)");
  }
  fs_choose(p, maxr);
  if (p->N1 > 0) {
    // W_N^m for m < 1024 and W_N^{1024 m} for m < ceil(N/1024)
    const int64_t nhi = (N + FS_TAB - 1) / FS_TAB;
    std::vector<c32> lo(FS_TAB), hi(static_cast<size_t>(nhi));
    const auto wn = [N](int64_t m) -> c32 {
      m %= N;
      const long double a = 2.0L * 3.14159265358979323846264338327950288L * static_cast<long double>(m) /
                            static_cast<long double>(N);
      double c = static_cast<double>(cosl(a)), sn = -static_cast<double>(sinl(a));
      if (4 * m == N || 4 * m == 3 * N) c = 0.0;
      if (2 * m == N) sn = 0.0;
      return {static_cast<float>(c), static_cast<float>(sn)};
    };
    for (int64_t m = 0; m < FS_TAB; ++m) lo[static_cast<size_t>(m)] = wn(m);
    for (int64_t m = 0; m < nhi; ++m) hi[static_cast<size_t>(m)] = wn(m * FS_TAB);
    e = hipMalloc(&p->tlo, (FS_TAB + nhi) * sizeof(c32));
    if (e != hipSuccess) {
      (void)hipFree(p->tw);
      delete p;
      return fail(ZDT_ENOMEM, "hipMalloc of FFT twiddles failed");
    }
    p->thi = p->tlo + FS_TAB;
    e = hipMemcpy(p->tlo, lo.data(), FS_TAB * sizeof(c32), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->thi, hi.data(), nhi * sizeof(c32), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(p->tlo);
      (void)hipFree(p->tw);
      delete p;
      return hip_fail(e, "hipMemcpy(twiddles)");
    }
  }
  g_plans[key] = p;
  *out = p;
  return ZDT_OK;
}

template <typename T>
static int c2c_impl(const FftPlan* p, const void* in, int64_t ies, int64_t ibs, void* out, int64_t oes,
                    int64_t obs, int64_t batch, bool inverse, double scale, void* s0, void* s1,
                    hipStream_t st) {
  const int np = static_cast<int>(p->radices.size());
  const int64_t N = p->N;
  const cplx<T>* tw = static_cast<const cplx<T>*>(p->tw);
  if constexpr (std::is_same<T, float>::value) {
    if (p->N1 > 0 && ies == 1 && oes == 1)
      return fs_c2c(p, static_cast<const c32*>(in), ibs, static_cast<c32*>(out), obs, batch, inverse, scale,
                    static_cast<c32*>(s0), st);
  }
  if (N == 1) {
    // the length-1 DFT is the identity (times scale)
    scale_copy<T><<<grid_for(batch, 256), 256, 0, st>>>(static_cast<const cplx<T>*>(in), ibs,
                                                        static_cast<cplx<T>*>(out), obs, batch,
                                                        static_cast<T>(scale));
    ZDT_CHECK_LAUNCH("scale_copy");
    return ZDT_OK;
  }
  int64_t Ns = 1;
  const void* src = in;
  int64_t ses = ies, sbs = ibs;
  for (int pi = 0; pi < np; ++pi) {
    const int R = p->radices[pi];
    const bool last = (pi == np - 1);
    void* dst = last ? out : ((pi & 1) ? s1 : s0);
    const int64_t des = last ? oes : 1;
    const int64_t dbs = last ? obs : N;
    PassArgs a;
    a.N = N;
    a.NR = N / R;
    a.Ns = Ns;
    a.twstep = N / (Ns * R);
    a.ies = ses;
    a.ibs = sbs;
    a.oes = des;
    a.obs = dbs;
    a.batch = batch;
    a.scale = last ? scale : 1.0;
    a.inverse = inverse ? 1 : 0;
    // consecutive lanes along the contiguous index
    a.jmajor = (ses == 1 || batch == 1) ? 1 : 0;
    ZDT_TRY(launch_pass<T>(R, static_cast<const cplx<T>*>(src), static_cast<cplx<T>*>(dst), tw, a, st));
    Ns *= R;
    src = dst;
    ses = des;
    sbs = dbs;
  }
  return ZDT_OK;
}

int fft_c2c(const FftPlan* p, const void* in, int64_t ies, int64_t ibs, void* out, int64_t oes,
            int64_t obs, int64_t batch, bool inverse, double scale, void* s0, void* s1, hipStream_t st) {
  if (p->dbl)
    return c2c_impl<double>(p, in, ies, ibs, out, oes, obs, batch, inverse, scale, s0, s1, st);
  return c2c_impl<float>(p, in, ies, ibs, out, oes, obs, batch, inverse, scale, s0, s1, st);
}

// ---------------------------------------------------------------- real transforms

// r2c post-processing for even N = 2M.  Z = FFT_M(x[2n] + i x[2n+1]).
// X[k] = (Z[k] + conj Z[M-k])/2 - i W_N^k (Z[k] - conj Z[M-k])/2, k in [0, M].
__global__ __launch_bounds__(256) void r2c_post(const c32* __restrict__ Z, c32* __restrict__ X,
                                                const c32* __restrict__ twN, int64_t M, int64_t ldx,
                                                int64_t batch) {
  const int64_t half = M / 2 + 1;  // k in [0, M/2]
  const int64_t total = half * batch;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t k = g % half, b = g / half;
    const c32* z = Z + b * M;
    c32* x = X + b * ldx;
    if (k == 0) {
      const c32 z0 = z[0];
      x[0] = {z0.x + z0.y, 0.0f};
      x[M] = {z0.x - z0.y, 0.0f};
      continue;
    }
    const int64_t km = M - k;
    const c32 a = z[k], bb = z[km];
    {
      // E = (a + conj bb)/2, O = (a - conj bb)/(2i)
      const c32 E = {0.5f * (a.x + bb.x), 0.5f * (a.y - bb.y)};
      const c32 O = {0.5f * (a.y + bb.y), -0.5f * (a.x - bb.x)};
      x[k] = cadd(E, cmul(twN[k], O));
    }
    if (km != k) {
      const c32 E = {0.5f * (bb.x + a.x), 0.5f * (bb.y - a.y)};
      const c32 O = {0.5f * (bb.y + a.y), -0.5f * (bb.x - a.x)};
      x[km] = cadd(E, cmul(twN[km], O));
    }
  }
}

// c2r pre-processing for even N = 2M.  The imaginary parts of X[0] and X[M]
// are ignored (the c2r convention of FFTW/cuFFT).  Z'[k] = (E + i O) * scale
// with E = (X[k] + conj X[M-k])/2, O = W_N^{-k} (X[k] - conj X[M-k])/2.
__global__ __launch_bounds__(256) void c2r_pre(const c32* __restrict__ X, c32* __restrict__ Z,
                                               const c32* __restrict__ twN, int64_t M, int64_t ldx,
                                               int64_t batch, float scale) {
  const int64_t half = M / 2 + 1;
  const int64_t total = half * batch;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t k = g % half, b = g / half;
    const c32* x = X + b * ldx;
    c32* z = Z + b * M;
    if (k == 0) {
      const float a = x[0].x, c = x[M].x;
      z[0] = {0.5f * (a + c) * scale, 0.5f * (a - c) * scale};
      continue;
    }
    const int64_t km = M - k;
    const c32 p = x[k], q = x[km];
    {
      const c32 E = {0.5f * (p.x + q.x), 0.5f * (p.y - q.y)};
      const c32 D = {0.5f * (p.x - q.x), 0.5f * (p.y + q.y)};
      const c32 O = cmulc(D, twN[k]);  // D * conj(W^k) = W^{-k} D
      z[k] = {(E.x - O.y) * scale, (E.y + O.x) * scale};
    }
    if (km != k) {
      const c32 E = {0.5f * (q.x + p.x), 0.5f * (q.y - p.y)};
      const c32 D = {0.5f * (q.x - p.x), 0.5f * (q.y + p.y)};
      const c32 O = cmulc(D, twN[km]);
      z[km] = {(E.x - O.y) * scale, (E.y + O.x) * scale};
    }
  }
}

// odd N helpers
__global__ __launch_bounds__(256) void real_to_complex(const float* __restrict__ in, int64_t ldi,
                                                       c32* __restrict__ out, int64_t N, int64_t batch) {
  const int64_t total = N * batch;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t n = g % N, b = g / N;
    out[b * N + n] = {in[b * ldi + n], 0.0f};
  }
}

__global__ __launch_bounds__(256) void copy_rows(const c32* __restrict__ in, int64_t ldi,
                                                 c32* __restrict__ out, int64_t ldo, int64_t rows,
                                                 int64_t batch) {
  const int64_t total = rows * batch;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t n = g % rows, b = g / rows;
    out[b * ldo + n] = in[b * ldi + n];
  }
}

// full Hermitian spectrum of length N (odd) from Nk = (N+1)/2 bins; Im X[0] ignored
__global__ __launch_bounds__(256) void hermitian_extend(const c32* __restrict__ X, int64_t ldx,
                                                        c32* __restrict__ full, int64_t N,
                                                        int64_t batch) {
  const int64_t total = N * batch;
  const int64_t Nk = N / 2 + 1;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t n = g % N, b = g / N;
    c32 v;
    if (n == 0) {
      v = {X[b * ldx].x, 0.0f};
    } else if (n < Nk) {
      v = X[b * ldx + n];
    } else {
      v = cconj(X[b * ldx + (N - n)]);
    }
    full[b * N + n] = v;
  }
}

__global__ __launch_bounds__(256) void complex_real_part(const c32* __restrict__ in, float* __restrict__ out,
                                                         int64_t ldo, int64_t N, int64_t batch) {
  const int64_t total = N * batch;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t n = g % N, b = g / N;
    out[b * ldo + n] = in[b * N + n].x;
  }
}

size_t fft_real_scratch_elems(int64_t N, int64_t batch) {
  const int64_t n = (N % 2 == 0) ? N / 2 : N;
  return static_cast<size_t>(n) * static_cast<size_t>(batch);
}

int fft_r2c(int64_t N, const float* in, int64_t ldi, c32* out, int64_t ldo, int64_t batch, void* s0,
            void* s1, void* s2, hipStream_t st, int maxr) {
  if (N % 2 == 0) {
    const int64_t M = N / 2;
    if (ldi % 2 != 0) return fail(ZDT_EINVAL, "r2c: column stride must be even for even N");
    const FftPlan *pM = nullptr, *pN = nullptr;
    ZDT_TRY(fft_plan_get(N, false, &pN, maxr));
    c32* Z = static_cast<c32*>(s2);
    if (M == 1) {
      // length-1 complex FFT of the packed pair is the pair itself
      ZDT_TRY_HIP(hipMemcpy2DAsync(Z, sizeof(c32), in, ldi * sizeof(float), sizeof(c32), batch,
                                   hipMemcpyDeviceToDevice, st));
    } else {
      ZDT_TRY(fft_plan_get(M, false, &pM, maxr));
      ZDT_TRY(fft_c2c(pM, in, 1, ldi / 2, Z, 1, M, batch, false, 1.0, s0, s1, st));
    }
    ProfScope ps(st, "r2c_post", 16.0 * static_cast<double>(M * batch));
    r2c_post<<<grid_for((M / 2 + 1) * batch, 256), 256, 0, st>>>(Z, out, static_cast<const c32*>(pN->tw),
                                                                   M, ldo, batch);
    ZDT_CHECK_LAUNCH("r2c_post");
    return ZDT_OK;
  }
  const FftPlan* p = nullptr;
  ZDT_TRY(fft_plan_get(N, false, &p, maxr));
  c32* full = static_cast<c32*>(s2);
  real_to_complex<<<grid_for(N * batch, 256), 256, 0, st>>>(in, ldi, full, N, batch);
  ZDT_CHECK_LAUNCH("real_to_complex");
  // N odd and > 1: FFT into s0/s1 chain, final result into s0 region is not
  // allowed (scratch), so transform in place on `full`.
  ZDT_TRY(fft_c2c(p, full, 1, N, full, 1, N, batch, false, 1.0, s0, s1, st));
  copy_rows<<<grid_for((N / 2 + 1) * batch, 256), 256, 0, st>>>(full, N, out, ldo, N / 2 + 1, batch);
  ZDT_CHECK_LAUNCH("copy_rows");
  return ZDT_OK;
}

int fft_c2r(int64_t N, const c32* in, int64_t ldi, float* out, int64_t ldo, int64_t batch, double scale,
            void* s0, void* s1, void* s2, hipStream_t st, int maxr) {
  if (N % 2 == 0) {
    const int64_t M = N / 2;
    if (ldo % 2 != 0) return fail(ZDT_EINVAL, "c2r: column stride must be even for even N");
    const FftPlan *pM = nullptr, *pN = nullptr;
    if (M > 1) {
      ZDT_TRY(fft_plan_get(M, false, &pM, maxr));
      if (pM->N1 > 0)  // the pre-pass fused into the four-step columns kernel
        return fs_c2c(pM, in, ldi, reinterpret_cast<c32*>(out), ldo / 2, batch, true, scale, static_cast<c32*>(s0),
                      st, true);
    }
    ZDT_TRY(fft_plan_get(N, false, &pN, maxr));
    c32* Z = static_cast<c32*>(s2);
    {
      ProfScope ps(st, "c2r_pre", 16.0 * static_cast<double>(M * batch));
      c2r_pre<<<grid_for((M / 2 + 1) * batch, 256), 256, 0, st>>>(
          in, Z, static_cast<const c32*>(pN->tw), M, ldi, batch, static_cast<float>(scale / M));
      ZDT_CHECK_LAUNCH("c2r_pre");
    }
    if (M == 1) {
      ZDT_TRY_HIP(hipMemcpy2DAsync(out, ldo * sizeof(float), Z, sizeof(c32), sizeof(c32), batch,
                                   hipMemcpyDeviceToDevice, st));
    } else {
      ZDT_TRY(fft_plan_get(M, false, &pM, maxr));
      ZDT_TRY(fft_c2c(pM, Z, 1, M, out, 1, ldo / 2, batch, true, 1.0, s0, s1, st));
    }
    return ZDT_OK;
  }
  const FftPlan* p = nullptr;
  ZDT_TRY(fft_plan_get(N, false, &p, maxr));
  c32* full = static_cast<c32*>(s2);
  hermitian_extend<<<grid_for(N * batch, 256), 256, 0, st>>>(in, ldi, full, N, batch);
  ZDT_CHECK_LAUNCH("hermitian_extend");
  ZDT_TRY(fft_c2c(p, full, 1, N, full, 1, N, batch, true, scale / static_cast<double>(N), s0, s1, st));
  complex_real_part<<<grid_for(N * batch, 256), 256, 0, st>>>(full, out, ldo, N, batch);
  ZDT_CHECK_LAUNCH("complex_real_part");
  return ZDT_OK;
}

}  // namespace zdt
