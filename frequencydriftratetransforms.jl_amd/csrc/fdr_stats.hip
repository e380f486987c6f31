// FDR statistics, normalization and proto-hit extraction (src/fdrutils.jl).
//
// fdrstats (:39-50) needs the mean of column 1 and the minimum over columns
// of the corrected standard deviation: colstats produces per-column shifted
// FP64 sums S1 = sum(x - c), S2 = sum((x - c)^2) with c = the column's first
// element (the shift keeps S2 - S1^2/N free of cancellation), reduced in a
// fixed order so the threshold is bit-reproducible run to run.
//
// findprotohits (:148-169) is findall(>=(thr)) in column-major order.  The
// matrix is cut into tiles of HIT_TILE consecutive linear elements; a count
// pass writes the number of hits per tile, one block scans the counts, and the
// write pass re-reads only the tiles that hold hits, compacting them in order
// with a block-level scan.  Hits are sparse at useful thresholds, so the data
// is read about once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "zdt_fast_dev.h"
#include "zdt_kernels.h"

namespace zdt {

constexpr int STAT_THREADS = 256;
constexpr int STAT_ROWS = 16384;  // rows per partial-sum block

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// one block per (chunk of STAT_ROWS rows, column); grid.y is capped at the
// launch limit, so a block walks the columns j = blockIdx.y + q*gridDim.y;
// vec: the columns are 16-byte aligned (float4 loads)
__global__ __launch_bounds__(STAT_THREADS) void colstats_partial(const float* __restrict__ fdr, int64_t Nf,
                                                                 int64_t Nr, int64_t ld, double* __restrict__ part,
                                                                 int nchunk, int vec) {
  const int chunk = blockIdx.x;
  constexpr int NW = STAT_THREADS / 64;
  __shared__ double sh1[NW], sh2[NW];
  const int wv = threadIdx.x >> 6;
  for (int64_t j = blockIdx.y; j < Nr; j += gridDim.y) {  // uniform per block
    const float* col = fdr + j * ld;
    const double c = static_cast<double>(col[0]);
    const int64_t r0 = static_cast<int64_t>(chunk) * STAT_ROWS;
    int64_t r1 = r0 + STAT_ROWS;
    if (r1 > Nf) r1 = Nf;
    double s1 = 0.0, s2 = 0.0;
    if (vec && r1 - r0 == STAT_ROWS) {
      const float4* p = reinterpret_cast<const float4*>(col + r0);
      constexpr int NV = STAT_ROWS / 4 / STAT_THREADS;
#pragma unroll 8
      for (int k = 0; k < NV; ++k) {
        const float4 v = p[threadIdx.x + k * STAT_THREADS];
        const double d0 = static_cast<double>(v.x) - c, d1 = static_cast<double>(v.y) - c;
        const double d2 = static_cast<double>(v.z) - c, d3 = static_cast<double>(v.w) - c;
        s1 += (d0 + d1) + (d2 + d3);
        s2 += (d0 * d0 + d1 * d1) + (d2 * d2 + d3 * d3);
      }
    } else {
      for (int64_t i = r0 + threadIdx.x; i < r1; i += STAT_THREADS) {
        const double d = static_cast<double>(col[i]) - c;
        s1 += d;
        s2 += d * d;
      }
    }
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    if ((threadIdx.x & 63) == 0) {
      sh1[wv] = s1;
      sh2[wv] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double t1 = 0.0, t2 = 0.0;
      for (int w = 0; w < NW; ++w) {
        t1 += sh1[w];
        t2 += sh2[w];
      }
      part[(j * nchunk + chunk) * 2 + 0] = t1;
      part[(j * nchunk + chunk) * 2 + 1] = t2;
    }
    __syncthreads();  // sh1/sh2 are rewritten for the next column
  }
}

// per column: mean and corrected variance from the partial sums (fixed order)
__global__ __launch_bounds__(256) void colstats_final(const float* __restrict__ fdr, int64_t Nf, int64_t Nr,
                                                      int64_t ld, const double* __restrict__ part, int nchunk,
                                                      double* __restrict__ colmean,
                                                      double* __restrict__ colvar) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= Nr) return;
  double s1 = 0.0, s2 = 0.0;
  for (int c = 0; c < nchunk; ++c) {
    s1 += part[(j * nchunk + c) * 2 + 0];
    s2 += part[(j * nchunk + c) * 2 + 1];
  }
  const double n = static_cast<double>(Nf);
  const double shift = static_cast<double>(fdr[j * ld]);
  colmean[j] = shift + s1 / n;
  colvar[j] = Nf > 1 ? (s2 - s1 * s1 / n) / (n - 1.0) : 0.0 / 0.0;
}

int colstats_nchunk(int64_t Nf) { return static_cast<int>((Nf + STAT_ROWS - 1) / STAT_ROWS); }

int launch_colstats2(const float* fdr, int64_t Nf, int64_t Nr, int64_t ld, double* part, double* colmean,
                     double* colvar, hipStream_t st) {
  const int nchunk = colstats_nchunk(Nf);
  dim3 grid(nchunk, static_cast<unsigned>(std::min<int64_t>(Nr, 65535)));
  const int vec = (ld % 4 == 0 && (reinterpret_cast<uintptr_t>(fdr) & 15) == 0) ? 1 : 0;
  colstats_partial<<<grid, STAT_THREADS, 0, st>>>(fdr, Nf, Nr, ld, part, nchunk, vec);
  ZDT_CHECK_LAUNCH("colstats_partial");
  colstats_final<<<static_cast<unsigned>((Nr + 255) / 256), 256, 0, st>>>(fdr, Nf, Nr, ld, part, nchunk,
                                                                          colmean, colvar);
  ZDT_CHECK_LAUNCH("colstats_final");
  return ZDT_OK;
}

// ---------------------------------------------------------------- normalize

__global__ __launch_bounds__(256) void normalize_kernel(float* __restrict__ fdr, int64_t Nf, int64_t Nr,
                                                        int64_t ld, double m, double inv_s) {
  const int64_t total = Nf * Nr;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t i = g % Nf, j = g / Nf;
    float* p = fdr + i + j * ld;
    *p = static_cast<float>((static_cast<double>(*p) - m) * inv_s);
  }
}

int launch_normalize(float* fdr, int64_t Nf, int64_t Nr, int64_t ld, double m, double s, hipStream_t st) {
  int64_t g = (Nf * Nr + 255) / 256;
  if (g > (1 << 20)) g = 1 << 20;
  if (g < 1) g = 1;
  normalize_kernel<<<static_cast<unsigned>(g), 256, 0, st>>>(fdr, Nf, Nr, ld, m, 1.0 / s);
  ZDT_CHECK_LAUNCH("normalize_kernel");
  return ZDT_OK;
}

// ---------------------------------------------------------------- proto-hits

constexpr int HIT_THREADS = 256;
constexpr int HIT_PER_THREAD = 16;
constexpr int64_t HIT_TILE = HIT_THREADS * HIT_PER_THREAD;
static_assert(HIT_TILE == FDR_TILE && HIT_TILE % FDR_RUN == 0, "k3's run maxima tile the proto-hit tiles");

int64_t hits_tiles(int64_t Nf, int64_t Nr) { return (Nf * Nr + HIT_TILE - 1) / HIT_TILE; }

__device__ __forceinline__ float load_lin(const float* fdr, int64_t Nf, int64_t ld, int64_t g) {
  if (ld == Nf) return fdr[g];
  return fdr[(g % Nf) + (g / Nf) * ld];
}

// The proto-hits of one tile, counted by the whole block; the total is valid
// in thread 0.  vec: the matrix is contiguous (ld == Nf) and 16-byte aligned.
__device__ int count_tile(const float* __restrict__ fdr, int64_t Nf, int64_t n, int64_t ld, double thr,
                          int64_t tile, int vec, int* sh) {
  const int64_t base = tile * HIT_TILE;
  int c = 0;
  if (vec && base + HIT_TILE <= n) {
    const float4* p = reinterpret_cast<const float4*>(fdr + base);
    float4 v[HIT_PER_THREAD / 4];
#pragma unroll
    for (int i = 0; i < HIT_PER_THREAD / 4; ++i) v[i] = p[threadIdx.x + i * HIT_THREADS];
#pragma unroll
    for (int i = 0; i < HIT_PER_THREAD / 4; ++i)
      c += (static_cast<double>(v[i].x) >= thr) + (static_cast<double>(v[i].y) >= thr) +
           (static_cast<double>(v[i].z) >= thr) + (static_cast<double>(v[i].w) >= thr);
  } else {
#pragma unroll 4
    for (int i = 0; i < HIT_PER_THREAD; ++i) {
      const int64_t g = base + static_cast<int64_t>(i) * HIT_THREADS + threadIdx.x;
      if (g < n && static_cast<double>(load_lin(fdr, Nf, ld, g)) >= thr) ++c;
    }
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
  constexpr int NW = HIT_THREADS / 64;
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
  __syncthreads();
  int t = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < NW; ++w) t += sh[w];
  __syncthreads();  // sh is reused by the caller's next tile
  return t;
}

__global__ __launch_bounds__(HIT_THREADS) void hits_count(const float* __restrict__ fdr, int64_t Nf, int64_t n,
                                                          int64_t ld, double thr_h, const double* thr_d,
                                                          int64_t* __restrict__ counts, int vec) {
  const double thr = thr_d ? *thr_d : thr_h;
  __shared__ int sh[HIT_THREADS / 64];
  const int t = count_tile(fdr, Nf, n, ld, thr, blockIdx.x, vec, sh);
  if (threadIdx.x == 0) counts[blockIdx.x] = t;
}

// Tiles per block of the sparse proto-hit kernels: a block looks at the
// bounds of HIT_GROUP tiles and works only on those that can hold a hit.
constexpr int HIT_GROUP = 16;

// hits_count with the FDR's run maxima (smax, written by k3, layout
// [column][N1/16][N2] with N1 = Nf/(2 N2)): a tile whose runs all lie under the
// threshold has no proto-hit and is not read.  Tile tc of a column holds the
// rows n2 of [tc R, tc R + R), R = 2048/N1, of every run column nb.  A block
// takes HIT_GROUP consecutive tiles of one column: 16 R rows of N1/16 run
// columns, 2048 maxima, 8 per thread (8 consecutive rows of one nb), folded
// per tile in LDS.
__global__ __launch_bounds__(HIT_THREADS) void hits_count_sparse(const float* __restrict__ fdr, int64_t Nf, int64_t n,
                                                                 int64_t ld, double thr_h, const double* thr_d,
                                                                 const float* __restrict__ smax, int64_t ntiles,
                                                                 int64_t* __restrict__ counts, int vec, int N2,
                                                                 int rpt) {
  static_assert(HIT_GROUP * (HIT_TILE / FDR_RUN) == HIT_THREADS * 8, "8 run maxima per thread");
  const double thr = thr_d ? *thr_d : thr_h;
  __shared__ int sh[HIT_THREADS / 64];
  __shared__ unsigned tm[HIT_GROUP];
  __shared__ unsigned need;
  if (threadIdx.x < HIT_GROUP) tm[threadIdx.x] = 0u;
  if (threadIdx.x == 0) need = 0u;
  __syncthreads();
  const int64_t tpc = Nf / HIT_TILE;
  const int64_t t0 = static_cast<int64_t>(blockIdx.x) * HIT_GROUP;  // tpc is a multiple of HIT_GROUP
  const int64_t jc = t0 / tpc, tc0 = t0 - jc * tpc;
  const int R = HIT_GROUP * rpt;  // rows of the group (a multiple of 16)
  const int v0 = threadIdx.x * 8;
  const int nb = v0 / R, r0 = v0 - nb * R;
  if (t0 < ntiles) {
    const float4* s = reinterpret_cast<const float4*>(smax + (jc * (Nf / (2 * static_cast<int64_t>(N2)) / 16) + nb) *
                                                                 static_cast<int64_t>(N2) + tc0 * rpt + r0);
    const float4 u = s[0], w = s[1];
    const float v[8] = {u.x, u.y, u.z, u.w, w.x, w.y, w.z, w.w};
    // fold runs of the same tile, one LDS max per tile met
    int tl = r0 / rpt;
    float m = v[0];
#pragma unroll
    for (int i = 1; i < 8; ++i) {
      const int ti = (r0 + i) / rpt;
      if (ti != tl) {
        atomicMax(&tm[tl], fdr_ordered(m));
        tl = ti;
        m = v[i];
      } else {
        m = fmaxf(m, v[i]);
      }
    }
    atomicMax(&tm[tl], fdr_ordered(m));
  }
  __syncthreads();
  if (threadIdx.x < HIT_GROUP) {
    const int64_t tile = t0 + threadIdx.x;
    if (tile < ntiles) {
      if (static_cast<double>(fdr_unordered(tm[threadIdx.x])) < thr) counts[tile] = 0;
      else atomicOr(&need, 1u << threadIdx.x);
    }
  }
  __syncthreads();
  for (unsigned mk = need; mk != 0u; mk &= mk - 1u) {  // uniform
    const int64_t tl = t0 + (__ffs(mk) - 1);
    const int c = count_tile(fdr, Nf, n, ld, thr, tl, vec, sh);
    if (threadIdx.x == 0) counts[tl] = c;
  }
}

// Exclusive scan of the tile counts in three steps: each block of
// SCAN_THREADS scans SCAN_CHUNK (2048) counts locally and writes its total
// (hits_scan_local), one block scans the totals (hits_scan), and the totals'
// offsets are added back (hits_scan_add).
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_PER_THREAD = 8;
constexpr int64_t SCAN_CHUNK = SCAN_THREADS * SCAN_PER_THREAD;

int64_t hits_scan_scratch(int64_t n) { return (n + SCAN_CHUNK - 1) / SCAN_CHUNK; }

__device__ __forceinline__ int64_t wave_incl_scan(int64_t v, int lane) {
  for (int off = 1; off < 64; off <<= 1) {
    const int64_t u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  return v;
}

__global__ __launch_bounds__(SCAN_THREADS) void hits_scan_local(int64_t* __restrict__ counts, int64_t n,
                                                                int64_t* __restrict__ btot) {
  constexpr int NW = SCAN_THREADS / 64;
  __shared__ int64_t sh[NW + 1];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t base = static_cast<int64_t>(blockIdx.x) * SCAN_CHUNK + threadIdx.x * SCAN_PER_THREAD;
  int64_t v[SCAN_PER_THREAD];
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_PER_THREAD; ++i) {
    v[i] = base + i < n ? counts[base + i] : 0;
    s += v[i];
  }
  const int64_t incl = wave_incl_scan(s, lane);
  if (lane == 63) sh[wv] = incl;
  __syncthreads();
  int64_t w = 0, wi = 0;
  if (wv == 0) {
    w = lane < NW ? sh[lane] : 0;
    wi = wave_incl_scan(w, lane);
  }
  __syncthreads();
  if (wv == 0) {
    if (lane < NW) sh[lane] = wi - w;  // exclusive offsets of the wavefronts
    if (lane == NW - 1) sh[NW] = wi;   // block total
  }
  __syncthreads();
  int64_t run = sh[wv] + incl - s;
#pragma unroll
  for (int i = 0; i < SCAN_PER_THREAD; ++i) {
    if (base + i < n) counts[base + i] = run;
    run += v[i];
  }
  if (threadIdx.x == 0) btot[blockIdx.x] = sh[NW];
}

__global__ __launch_bounds__(256) void hits_scan_add(int64_t* __restrict__ counts, int64_t n,
                                                     const int64_t* __restrict__ boff) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (g < n) counts[g] += boff[g / SCAN_CHUNK];
}

// exclusive scan of counts[0..n) in place, total in *total (one block; used
// on the per-chunk totals)
__global__ __launch_bounds__(1024) void hits_scan(int64_t* __restrict__ counts, int64_t n,
                                                  int64_t* __restrict__ total) {
  __shared__ int64_t sh[1024];
  int64_t carry = 0;
  for (int64_t base = 0; base < n; base += 1024) {
    const int64_t g = base + threadIdx.x;
    const int64_t v = g < n ? counts[g] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      int64_t add = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += add;
      __syncthreads();
    }
    const int64_t incl = sh[threadIdx.x];
    if (g < n) counts[g] = carry + incl - v;
    const int64_t blocksum = sh[1023];
    __syncthreads();
    carry += blocksum;
  }
  if (threadIdx.x == 0) *total = carry;
}

// The proto-hits of one tile written in order at hits[start...], by the whole
// block (which must not have returned early).
__device__ void write_tile(const float* __restrict__ fdr, int64_t Nf, int64_t n, int64_t ld, double thr, int64_t tile,
                           int64_t start, int64_t base_index, int64_t* __restrict__ hits, int64_t cap, int vec,
                           int* sh) {
  // thread handles HIT_PER_THREAD consecutive elements, so a block scan of
  // the per-thread counts gives each hit its ordered slot
  const int64_t base = tile * HIT_TILE + static_cast<int64_t>(threadIdx.x) * HIT_PER_THREAD;
  unsigned mask = 0;
  if (vec && tile * HIT_TILE + HIT_TILE <= n) {
    const float4* p = reinterpret_cast<const float4*>(fdr + base);
#pragma unroll
    for (int i = 0; i < HIT_PER_THREAD / 4; ++i) {
      const float4 v = p[i];
      mask |= (static_cast<double>(v.x) >= thr ? 1u : 0u) << (4 * i);
      mask |= (static_cast<double>(v.y) >= thr ? 1u : 0u) << (4 * i + 1);
      mask |= (static_cast<double>(v.z) >= thr ? 1u : 0u) << (4 * i + 2);
      mask |= (static_cast<double>(v.w) >= thr ? 1u : 0u) << (4 * i + 3);
    }
  } else {
#pragma unroll
    for (int i = 0; i < HIT_PER_THREAD; ++i) {
      const int64_t g = base + i;
      if (g < n && static_cast<double>(load_lin(fdr, Nf, ld, g)) >= thr) mask |= 1u << i;
    }
  }
  const int c = __popc(mask);
  sh[threadIdx.x] = c;
  __syncthreads();
  for (int off = 1; off < HIT_THREADS; off <<= 1) {
    int add = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += add;
    __syncthreads();
  }
  int64_t pos = start + sh[threadIdx.x] - c;
  __syncthreads();  // sh is reused by the next tile
  while (mask) {
    const int i = __ffs(mask) - 1;
    mask &= mask - 1;
    if (pos < cap) hits[pos] = base_index + base + i;
    ++pos;
  }
}

// HIT_GROUP tiles per block; a tile whose scanned count is zero (or whose
// hits lie past cap) is not read.
__global__ __launch_bounds__(HIT_THREADS) void hits_write(const float* __restrict__ fdr, int64_t Nf, int64_t n,
                                                          int64_t ld, double thr_h, const double* thr_d,
                                                          const int64_t* __restrict__ offsets, int64_t ntiles,
                                                          const int64_t* __restrict__ total, int64_t base_index,
                                                          const int64_t* __restrict__ pos_base,
                                                          int64_t* __restrict__ hits, int64_t cap, int vec,
                                                          int64_t nt_here) {
  const double thr = thr_d ? *thr_d : thr_h;
  __shared__ int sh[HIT_THREADS];
  __shared__ unsigned need;
  __shared__ int64_t starts[HIT_GROUP];
  const int64_t t0 = static_cast<int64_t>(blockIdx.x) * HIT_GROUP;
  if (threadIdx.x < 64) {
    const int64_t tile = t0 + threadIdx.x;
    bool work = false;
    if (threadIdx.x < HIT_GROUP && tile < nt_here) {
      const int64_t pb = pos_base ? *pos_base : 0;  // hits already written by earlier matrices
      const int64_t start = offsets[tile] + pb;
      const int64_t end = ((tile + 1 < ntiles) ? offsets[tile + 1] : *total) + pb;
      work = end > start && start < cap;
      starts[threadIdx.x] = start;
    }
    const unsigned long long b = __ballot(work);
    if (threadIdx.x == 0) need = static_cast<unsigned>(b);
  }
  __syncthreads();
  for (unsigned m = need; m != 0u; m &= m - 1u) {  // uniform
    const int q = __ffs(m) - 1;
    write_tile(fdr, Nf, n, ld, thr, t0 + q, starts[q], base_index, hits, cap, vec, sh);
  }
}

int launch_hits_count(const float* fdr, int64_t Nf, int64_t Nr, int64_t ld, double thr, const double* thr_dev,
                      int64_t* counts, hipStream_t st, const float* smax, int smax_n2) {
  const int64_t nt = hits_tiles(Nf, Nr);
  if (nt == 0) return ZDT_OK;
  const int vec = (ld == Nf && (reinterpret_cast<uintptr_t>(fdr) & 15) == 0) ? 1 : 0;
  if (smax != nullptr) {
    // the layout the fused output writes: N1 = Nf/(2 N2) a power of two in
    // [16, 2048] (whole rows per tile), tiles per column a multiple of the group
    const int64_t N1 = smax_n2 > 0 ? Nf / (2 * static_cast<int64_t>(smax_n2)) : 0;
    if (ld != Nf || Nf % (HIT_TILE * HIT_GROUP) != 0 || N1 < 16 || N1 > 2048 || (N1 & (N1 - 1)) != 0 ||
        2 * N1 * smax_n2 != Nf || (reinterpret_cast<uintptr_t>(smax) & 15) != 0)
      return fail(ZDT_EINVAL, "hits_count: run maxima of an unsupported shape (Nf=%lld, N2=%d)", (long long)Nf,
                  smax_n2);
    // bytes: the run maxima, and the counts
    ProfScope ps(st, "hits_count_sparse", 4.0 * (HIT_TILE / FDR_RUN) * nt + 8.0 * nt);
    hits_count_sparse<<<static_cast<unsigned>((nt + HIT_GROUP - 1) / HIT_GROUP), HIT_THREADS, 0, st>>>(
        fdr, Nf, Nf * Nr, ld, thr, thr_dev, smax, nt, counts, vec, smax_n2, static_cast<int>(2048 / N1));
    ZDT_CHECK_LAUNCH("hits_count_sparse");
    return ZDT_OK;
  }
  ProfScope ps(st, "hits_count", 4.0 * Nf * Nr + 8.0 * nt);
  hits_count<<<static_cast<unsigned>(nt), HIT_THREADS, 0, st>>>(fdr, Nf, Nf * Nr, ld, thr, thr_dev, counts, vec);
  ZDT_CHECK_LAUNCH("hits_count");
  return ZDT_OK;
}

int launch_hits_scan(int64_t* counts, int64_t n, int64_t* total, int64_t* scratch, hipStream_t st) {
  const int64_t nb = hits_scan_scratch(n);
  ProfScope ps(st, "hits_scan", 24.0 * n);
  if (nb > 0) {
    hits_scan_local<<<static_cast<unsigned>(nb), SCAN_THREADS, 0, st>>>(counts, n, scratch);
    ZDT_CHECK_LAUNCH("hits_scan_local");
  }
  hits_scan<<<1, 1024, 0, st>>>(scratch, nb, total);
  ZDT_CHECK_LAUNCH("hits_scan");
  if (nb > 1) {
    hits_scan_add<<<static_cast<unsigned>((n + 255) / 256), 256, 0, st>>>(counts, n, scratch);
    ZDT_CHECK_LAUNCH("hits_scan_add");
  }
  return ZDT_OK;
}

int launch_hits_write_impl(const float* fdr, int64_t Nf, int64_t Nr, int64_t ld, double thr,
                           const double* thr_dev, const int64_t* offsets, int64_t ntiles_all, int64_t tile0,
                           const int64_t* total_dev, int64_t base_index, int64_t* hits, int64_t cap,
                           hipStream_t st, const int64_t* pos_base) {
  const int64_t nt = hits_tiles(Nf, Nr);
  if (nt == 0) return ZDT_OK;
  // offsets of this matrix's tiles start at tile0 of the global scan; the end
  // of its last tile is the next matrix's first offset (or the total)
  const int vec = (ld == Nf && (reinterpret_cast<uintptr_t>(fdr) & 15) == 0) ? 1 : 0;
  ProfScope ps(st, "hits_write", 4.0 * Nf * Nr + 8.0 * nt);
  hits_write<<<static_cast<unsigned>((nt + HIT_GROUP - 1) / HIT_GROUP), HIT_THREADS, 0, st>>>(
      fdr, Nf, Nf * Nr, ld, thr, thr_dev, offsets + tile0, ntiles_all - tile0, total_dev, base_index, pos_base, hits,
      cap, vec, nt);
  ZDT_CHECK_LAUNCH("hits_write");
  return ZDT_OK;
}

__global__ void add_i64(int64_t* __restrict__ acc, const int64_t* __restrict__ v, int init) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *acc = (init ? 0 : *acc) + *v;
}

int launch_accumulate_i64(int64_t* acc, const int64_t* v, bool init, hipStream_t st) {
  add_i64<<<1, 64, 0, st>>>(acc, v, init ? 1 : 0);
  ZDT_CHECK_LAUNCH("add_i64");
  return ZDT_OK;
}

// fdrdenormalize on the device (src/fdrutils.jl:104-119 with :39-50):
// thr = snr * s + m, s = sqrt(min_j colvar[j]) (NaN if any is NaN),
// m = colmean[0]; one wavefront, the minimum taken lane-wise then across the
// lanes (min is exact, so the order does not matter)
__global__ __launch_bounds__(64) void hits_threshold(const double* __restrict__ colmean,
                                                     const double* __restrict__ colvar, int64_t Nr, double snr,
                                                     double* __restrict__ thr) {
  double vmin = INFINITY;
  int nan = 0;
  for (int64_t j = threadIdx.x; j < Nr; j += 64) {
    const double v = colvar[j];
    if (v != v) nan = 1;
    vmin = fmin(vmin, v);
  }
  for (int off = 32; off > 0; off >>= 1) {
    vmin = fmin(vmin, __shfl_xor(vmin, off, 64));
    nan |= __shfl_xor(nan, off, 64);
  }
  if (threadIdx.x == 0) {
    const double sd = nan ? NAN : sqrt(fmax(vmin, 0.0));
    *thr = snr * sd + colmean[0];
  }
}

__global__ void set_double(double* p, double v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

int launch_set_double(double* p, double v, hipStream_t st) {
  set_double<<<1, 64, 0, st>>>(p, v);
  ZDT_CHECK_LAUNCH("set_double");
  return ZDT_OK;
}

int launch_hits_threshold(const double* colmean, const double* colvar, int64_t Nr, double snr, double* thr,
                          hipStream_t st) {
  ProfScope ps(st, "hits_threshold", 16.0 * Nr);
  hits_threshold<<<1, 64, 0, st>>>(colmean, colvar, Nr, snr, thr);
  ZDT_CHECK_LAUNCH("hits_threshold");
  return ZDT_OK;
}

}  // namespace zdt

// ---------------------------------------------------------------- allocation-free findprotohits

namespace zdt {

// The tile counts, the scan's block totals and the total of
// zdt_findprotohits_noalloc live in this file-scope array instead of a
// per-call allocation: ZDT_HITS_NOALLOC_MAX_TILES counts, one total, and the
// scan's ceil(max / SCAN_CHUNK) block totals (16 MiB + 8 KiB per device, held
// until the process ends).  Host code never passes the array's name to a
// kernel or a copy: it uses the device address hipGetSymbolAddress returns,
// resolved once per device on the first call there (each device holds its
// own copy of the array).
constexpr int64_t NOALLOC_TILES = ZDT_HITS_NOALLOC_MAX_TILES;
constexpr int64_t NOALLOC_WORDS = NOALLOC_TILES + 1 + (NOALLOC_TILES + SCAN_CHUNK - 1) / SCAN_CHUNK;
static_assert(NOALLOC_TILES * HIT_TILE == ZDT_HITS_NOALLOC_MAX_CELLS, "header and tile size disagree");

static __device__ __align__(256) int64_t scratch[NOALLOC_WORDS];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
int64_t* g_scratch_dev[NOALLOC_MAX_DEVICES] = {};

int noalloc_scratch(int64_t** out) {
  int dev = 0;
  ZDT_TRY_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(ZDT_ENOTSUP, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_dev[dev] = static_cast<int64_t*>(p);
  }
  *out = g_scratch_dev[dev];
  return ZDT_OK;
}
}  // namespace

}  // namespace zdt

extern "C" int zdt_findprotohits_noalloc(const float* const* fdrs, int nb, int64_t Nf, int64_t Nr, int64_t ld,
                                         double threshold, int64_t* hits, int64_t cap, int64_t* nhits,
                                         void* stream) {
  using namespace zdt;
  if (fdrs == nullptr || nb < 1) return fail(ZDT_EINVAL, "need at least one FDR matrix");
  if (Nf < 1 || Nr < 1 || ld < Nf)
    return fail(ZDT_EINVAL, "bad FDR shape Nf=%lld Nr=%lld ld=%lld", (long long)Nf, (long long)Nr, (long long)ld);
  for (int b = 0; b < nb; ++b)
    if (fdrs[b] == nullptr) return fail(ZDT_EINVAL, "FDR matrix %d is NULL", b);
  if (cap < 0 || (cap > 0 && hits == nullptr)) return fail(ZDT_EINVAL, "findprotohits: bad output buffer");
  // the size limit, checked before anything is launched
  if (Nf > ZDT_HITS_NOALLOC_MAX_CELLS / Nr)
    return fail(ZDT_ENOTSUP, "findprotohits (allocation-free): %lld x %lld cells exceed its limit of %lld tiles "
                "of %lld cells", (long long)Nf, (long long)Nr, (long long)NOALLOC_TILES, (long long)HIT_TILE);
  const int64_t nt = hits_tiles(Nf, Nr);
  if (nt > NOALLOC_TILES / nb)
    return fail(ZDT_ENOTSUP, "findprotohits (allocation-free): %d x %lld tiles exceed its limit of %lld tiles", nb,
                (long long)nt, (long long)NOALLOC_TILES);
  int64_t* buf = nullptr;
  ZDT_TRY(noalloc_scratch(&buf));
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t ntall = nt * nb;
  int64_t* counts = buf;              // every tile's count is written by hits_count
  int64_t* total_dev = buf + ntall;   // written by hits_scan
  int64_t* scan_tmp = total_dev + 1;  // hits_scan_scratch(ntall) block totals, written by hits_scan_local
  for (int b = 0; b < nb; ++b) ZDT_TRY(launch_hits_count(fdrs[b], Nf, Nr, ld, threshold, nullptr, counts + b * nt, st));
  ZDT_TRY(launch_hits_scan(counts, ntall, total_dev, scan_tmp, st));
  int64_t total = 0;
  {
    hipError_t e = hipMemcpyAsync(&total, total_dev, sizeof(int64_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "findprotohits count copy");
  }
  if (cap > 0 && total > 0) {
    for (int b = 0; b < nb; ++b)
      ZDT_TRY(launch_hits_write_impl(fdrs[b], Nf, Nr, ld, threshold, nullptr, counts, ntall, b * nt, total_dev,
                                     static_cast<int64_t>(b) * Nf * Nr, hits, cap, st));
  }
  if (nhits) *nhits = total;
  if (total > cap)
    return fail(ZDT_ETRUNC, "findprotohits: %lld hits exceed capacity %lld", (long long)total, (long long)cap);
  return ZDT_OK;
}

