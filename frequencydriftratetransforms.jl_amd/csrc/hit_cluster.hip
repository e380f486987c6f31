// Proto-hit clustering: connected components of the proto-hit cells of an FDR.
//
// findprotohits (src/fdrutils.jl:131-169) returns every cell at or above the
// threshold; its docstring calls them proto-hits because "no clustering of
// neighboring proto-hits has been performed", and README.md:48-53 names that
// clustering as the next step of a search.  Here two proto-hits are neighbours
// when their channels differ by at most di and their (hcat'd) rate columns by
// at most dj, and a hit is one connected component of that neighbour graph.
//
// Input: the proto-hits as zdt_findprotohits writes them, 0-based linear
// indices h = i + Nf*J in strictly ascending order.  Sorted input turns every
// neighbour query into a contiguous range: the same column's neighbours are
// the next entries, and column J+d's are one binary search away.
//
// Union-find over proto-hit ordinals, after the ECL-CC scheme (Jaiganesh and
// Burtscher, "A High-Performance Connected Components Implementation for
// GPUs", HPDC 2018): a root is only ever hooked under a smaller root with one
// compare-and-swap, and finds shorten the path they walk.  So parent[x] <= x
// always, every walk descends, and each set's root is its smallest ordinal --
// its first proto-hit in column-major order.  Clusters are numbered by a scan
// over the roots, which orders them by first proto-hit, and their size,
// bounding box and peak are reduced with atomics that commute (add, min, max),
// so every output is independent of scheduling.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "zdt_internal.h"
#include "zdt_kernels.h"

namespace zdt {

namespace {

constexpr int CL_THREADS = 256;
// one proto-hit per thread in every kernel: the root walks and the numbering
// are chains of dependent loads, so parallelism, not work per thread, sets
// their time
constexpr int CL_PER_THREAD = 1;
constexpr int CL_TILE = CL_THREADS * CL_PER_THREAD;  // proto-hits per tile of the root count

__device__ __forceinline__ int ld_parent(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_parent(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x.  Each visited entry is pointed at its grandparent, an ancestor of
// it in the same set and smaller, so concurrent finds and hooks keep
// parent[y] <= y.  A root's entry is never written here (only hooks do that).
__device__ int uf_find(int* parent, int x) {
  int cur = ld_parent(parent + x);
  if (cur == x) return x;
  int prev = x;
  for (;;) {
    const int next = ld_parent(parent + cur);
    if (next == cur) return cur;
    st_parent(parent + prev, next);
    prev = cur;
    cur = next;
  }
}

// Merge the sets of a and b: hook the larger root under the smaller.  A failed
// compare-and-swap means another thread hooked that root first; both roots are
// found again and the attempt repeats.  Every failure is some thread's
// success, so the loop ends.
__device__ void uf_union(int* parent, int a, int b) {
  int ra = uf_find(parent, a), rb = uf_find(parent, b);
  while (ra != rb) {
    if (ra > rb) {
      const int t = ra;
      ra = rb;
      rb = t;
    }
    const int old = atomicCAS(parent + rb, rb, ra);
    if (old == rb) return;
    ra = uf_find(parent, ra);
    rb = uf_find(parent, old);
  }
}

// The nb matrices: one pointer by value (the usual single matrix), or a
// device table of nb pointers.
struct Mats {
  const float* one;
  const float* const* tab;
  __device__ const float* get(int64_t b) const { return tab ? tab[b] : one; }
};

struct InitArgs {
  const int64_t* hits;
  int64_t n;
  int64_t Nf;
  int64_t ncol;       // columns of the hcat'd FDR: nb * Nr
  int* parent;
  int64_t* colstart;  // ncol + 1 entries: first proto-hit of each column (null when dj = 0)
  int64_t* flags;     // [0] hit count (written by the scan), [1] input error
};

// The column of a proto-hit index, clamped to [0, ncol] so that malformed
// input cannot address outside colstart.
__device__ __forceinline__ int64_t col_of(int64_t h, int64_t Nf, int64_t ncol) {
  if (h < 0) return 0;
  const int64_t c = h / Nf;
  return c < ncol ? c : ncol;
}

// parent[k] = k; colstart[c] = first k in column >= c (every entry written
// exactly once for sorted input; unwritten ones are clamped where read).
__global__ __launch_bounds__(CL_THREADS) void cl_init(InitArgs a) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * CL_THREADS + threadIdx.x;
  if (k >= a.n) return;
  a.parent[k] = static_cast<int>(k);
  if (k == 0) a.flags[1] = 0;
  if (a.colstart == nullptr) return;
  const int64_t c = col_of(a.hits[k], a.Nf, a.ncol);
  const int64_t cp = k == 0 ? -1 : col_of(a.hits[k - 1], a.Nf, a.ncol);
  for (int64_t q = cp + 1; q <= c; ++q) a.colstart[q] = k;
  if (k == a.n - 1)
    for (int64_t q = c + 1; q <= a.ncol; ++q) a.colstart[q] = a.n;
}

struct LinkArgs {
  const int64_t* hits;
  int64_t n;
  int64_t Nf;
  int64_t ncol;
  int di, dj;
  int* parent;
  const int64_t* colstart;
  int64_t* flags;
};

// One thread per proto-hit k: union with every neighbour after it in the sorted
// order (the neighbours before it do the same towards k).  A proto-hit out of
// range or out of order sets the error flag and links nothing; every read
// stays in hits[0, n) either way.
__global__ __launch_bounds__(CL_THREADS) void cl_link(LinkArgs a) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * CL_THREADS + threadIdx.x;
  if (k >= a.n) return;
  const int64_t h = a.hits[k];
  if (h < 0 || h >= a.Nf * a.ncol || (k > 0 && a.hits[k - 1] >= h)) {
    atomicOr(reinterpret_cast<unsigned long long*>(a.flags + 1), 1ull);
    return;
  }
  const int64_t J = h / a.Nf, i = h - J * a.Nf;
  // the same column: channels i+1 .. i+di are the next entries (an entry not
  // above h is out of order, flagged by its own thread, and ends the walk, so
  // that unsorted input costs at most di steps here too)
  const int64_t colend = (J + 1) * a.Nf;
  for (int64_t m = k + 1; m < a.n; ++m) {
    const int64_t g = a.hits[m];
    if (g <= h || g > h + a.di || g >= colend) break;
    uf_union(a.parent, static_cast<int>(k), static_cast<int>(m));
  }
  // columns J+1 .. J+dj: channels max(0, i-di) .. min(Nf-1, i+di), one
  // contiguous run inside the column's own range of the sorted list
  const int64_t ilo = i > a.di ? i - a.di : 0;
  const int64_t ihi = i + a.di < a.Nf - 1 ? i + a.di : a.Nf - 1;
  for (int d = 1; d <= a.dj && J + d < a.ncol; ++d) {
    const int64_t base = (J + d) * a.Nf;
    const int64_t lo = base + ilo, hi = base + ihi;
    int64_t L = a.colstart[J + d], R = a.colstart[J + d + 1];
    L = L < k + 1 ? k + 1 : (L > a.n ? a.n : L);
    R = R < L ? L : (R > a.n ? a.n : R);
    int64_t e = R;
    while (L < R) {
      const int64_t mid = (L + R) >> 1;
      if (a.hits[mid] < lo) L = mid + 1;
      else R = mid;
    }
    for (int64_t m = L; m < e && a.hits[m] <= hi; ++m) uf_union(a.parent, static_cast<int>(k), static_cast<int>(m));
  }
}

// After linking: every entry is pointed straight at its root, and the roots
// (parent[k] == k, which flattening leaves unchanged) of each tile of CL_TILE
// proto-hits are counted.
__global__ __launch_bounds__(CL_THREADS) void cl_flatten_count(int* parent, int64_t n, int64_t* __restrict__ counts) {
  __shared__ int sh[CL_THREADS / 64];
  int s = 0;
#pragma unroll 4
  for (int q = 0; q < CL_PER_THREAD; ++q) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * CL_TILE + q * CL_THREADS + threadIdx.x;
    if (k >= n) break;
    int r = ld_parent(parent + k);
    for (int next = ld_parent(parent + r); next != r; next = ld_parent(parent + r)) r = next;
    st_parent(parent + k, r);
    s += r == static_cast<int>(k) ? 1 : 0;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < CL_THREADS / 64; ++w) t += sh[w];
    counts[blockIdx.x] = t;
  }
}

__device__ __forceinline__ int wave_incl_scan_i(int v, int lane) {
  for (int off = 1; off < 64; off <<= 1) {
    const int u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  return v;
}

struct NumberArgs {
  const int64_t* hits;
  int64_t n;
  int64_t Nf;
  const int* parent;
  const int64_t* offs;  // exclusive scan of the per-tile root counts
  int* cid;             // cluster ordinal, written at each root's entry
  int64_t cap;
  int64_t* peak;        // per cluster: the peak key while reducing, cleared here
  int64_t* size;
  int64_t* box;  // per cluster: channel min, channel max, column min, column max
};

// Number the roots in order and clear the accumulators of the clusters that
// fit in cap.  A root is its cluster's first proto-hit, so the column minimum
// is the root's column.
__global__ __launch_bounds__(CL_THREADS) void cl_number_roots(NumberArgs a) {
  __shared__ int sh[CL_THREADS / 64 + 1];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * CL_TILE + threadIdx.x * CL_PER_THREAD;
  int s = 0;
#pragma unroll
  for (int q = 0; q < CL_PER_THREAD; ++q) {
    const int64_t k = base + q;
    s += (k < a.n && a.parent[k] == static_cast<int>(k)) ? 1 : 0;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int incl = wave_incl_scan_i(s, lane);
  if (lane == 63) sh[wv] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < CL_THREADS / 64; ++w) {
      const int v = sh[w];
      sh[w] = t;
      t += v;
    }
  }
  __syncthreads();
  int64_t c = a.offs[blockIdx.x] + sh[wv] + incl - s;
  for (int q = 0; q < CL_PER_THREAD; ++q) {
    const int64_t k = base + q;
    if (k >= a.n || a.parent[k] != static_cast<int>(k)) continue;
    a.cid[k] = static_cast<int>(c);
    if (c < a.cap) {
      a.peak[c] = 0;
      a.size[c] = 0;
      a.box[4 * c + 0] = INT64_MAX;
      a.box[4 * c + 1] = 0;
      a.box[4 * c + 2] = a.hits[k] / a.Nf;
      a.box[4 * c + 3] = 0;
    }
    ++c;
  }
}

struct AccumArgs {
  const int64_t* hits;
  int64_t n;
  int64_t Nf, Nr, ld;
  int64_t ncol;
  Mats mats;
  const int* parent;  // flattened roots; null: cid is per proto-hit (the one-workgroup path)
  const int* cid;
  int64_t cap;
  int64_t* peak;
  int64_t* size;
  int64_t* box;
  int64_t* label;  // per proto-hit cluster ordinal, or null
  const int64_t* done;  // non-zero: cl_small has reduced the hits already (null on the general path)
};

// Order-preserving map of a float onto an unsigned integer.
__device__ __forceinline__ unsigned ordered_bits(float v) {
  const unsigned u = __float_as_uint(v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unordered_bits(unsigned o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

// One thread per proto-hit.  Proto-hits of one hit are often neighbours in the
// sorted order, and at dense thresholds one hit can hold most of them, so per
// proto-hit atomics would queue on one hit's five words.  Each wavefront first
// folds the lanes of up to three of its hits (taken in lane order) with
// cross-lane reductions, and one lane per such hit applies the result; the
// lanes left do their own atomics.  max/min/add commute, so the result is the
// same either way.  The peak is reduced as a key (value bits, then the
// reversed ordinal) in the peak array itself until cl_finish decodes it.
__global__ __launch_bounds__(CL_THREADS) void cl_accumulate(AccumArgs a) {
  if (a.done != nullptr && *a.done != 0) return;  // uniform
  const int64_t k = static_cast<int64_t>(blockIdx.x) * CL_THREADS + threadIdx.x;
  unsigned long long* keys = reinterpret_cast<unsigned long long*>(a.peak);
  bool live = false;
  int c = -1, i = 0, J = 0;
  unsigned long long key = 0ull;
  if (k < a.n) {
    const int64_t h = a.hits[k];
    if (h >= 0 && h < a.Nf * a.ncol) {  // flagged by cl_link otherwise
      c = a.parent ? a.cid[a.parent[k]] : a.cid[k];
      if (a.label) a.label[k] = c;
      if (c < a.cap) {
        const int64_t JJ = h / a.Nf, ii = h - JJ * a.Nf;
        const int64_t b = JJ / a.Nr, j = JJ - b * a.Nr;
        const float v = a.mats.get(b)[ii + a.ld * j];
        // peak: the largest value, ties to the smaller ordinal (the earlier proto-hit)
        key = (static_cast<unsigned long long>(ordered_bits(v)) << 32) |
              (0xffffffffull - static_cast<unsigned long long>(k));
        i = static_cast<int>(ii);
        J = static_cast<int>(JJ);
        live = true;
      }
    }
  }
  const int lane = threadIdx.x & 63;
  unsigned long long todo = __ballot(live);
  for (int it = 0; it < 3 && todo != 0ull; ++it) {
    const int leader = __ffsll(static_cast<long long>(todo)) - 1;
    const int cl = __shfl(c, leader, 64);
    const unsigned long long same = __ballot(live && c == cl);
    const int cnt = __popcll(same);
    if (cnt > 1) {
      const bool in = (same >> lane) & 1ull;
      unsigned long long kr = in ? key : 0ull;
      int imn = in ? i : INT32_MAX, imx = in ? i : 0, jmx = in ? J : 0;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long ko = __shfl_xor(kr, off, 64);
        kr = ko > kr ? ko : kr;
        imn = min(imn, __shfl_xor(imn, off, 64));
        imx = max(imx, __shfl_xor(imx, off, 64));
        jmx = max(jmx, __shfl_xor(jmx, off, 64));
      }
      if (lane == leader) {
        atomicMax(keys + cl, kr);
        atomicAdd(reinterpret_cast<unsigned long long*>(a.size + cl), static_cast<unsigned long long>(cnt));
        atomicMin(reinterpret_cast<unsigned long long*>(a.box + 4 * cl + 0), static_cast<unsigned long long>(imn));
        atomicMax(reinterpret_cast<unsigned long long*>(a.box + 4 * cl + 1), static_cast<unsigned long long>(imx));
        atomicMax(reinterpret_cast<unsigned long long*>(a.box + 4 * cl + 3), static_cast<unsigned long long>(jmx));
      }
      if (in) live = false;
    }
    todo &= ~same;
  }
  if (live) {
    atomicMax(keys + c, key);
    atomicAdd(reinterpret_cast<unsigned long long*>(a.size + c), 1ull);
    atomicMin(reinterpret_cast<unsigned long long*>(a.box + 4 * c + 0), static_cast<unsigned long long>(i));
    atomicMax(reinterpret_cast<unsigned long long*>(a.box + 4 * c + 1), static_cast<unsigned long long>(i));
    atomicMax(reinterpret_cast<unsigned long long*>(a.box + 4 * c + 3), static_cast<unsigned long long>(J));
  }
}

// ---- one workgroup for few proto-hits
//
// Up to CL_SMALL_MAX proto-hits (and Nf * ncol < 2^31) are clustered by one
// workgroup with the proto-hits and the parents in LDS: the same union-find,
// neighbour rule and numbering as cl_init .. cl_number_roots, in one launch,
// with every root walk and search on LDS instead of device-scope memory.
// The neighbours in column J+d are found by one binary search over the rest
// of the sorted list (no column table).  With at most CL_SMALL_ACC hits, the
// parents' LDS then takes the hits' accumulators: the same keyed peak, size
// and box reductions as cl_accumulate, on LDS atomics, and the workgroup
// writes the hits and labels itself and sets flags[2], on which cl_accumulate
// and cl_finish return at once.  With more hits it writes cid[k] = the hit
// ordinal of every proto-hit (cl_accumulate then reads it without the parent
// array) and clears the accumulators of the hits below cap, so that
// cl_accumulate and cl_finish complete it as they complete the general path.
// Input that cl_link would flag sets the error flag and gives every proto-hit
// the ordinal INT_MAX and a count of 0, so that the kernels after it touch
// nothing within cap.
constexpr int CL_SMALL_THREADS = 1024;
constexpr int CL_SMALL_MAX = 16384;
constexpr int CL_SMALL_ACC = CL_SMALL_MAX * 4 / 24;  // hits whose accumulators fit in the parents' LDS

__device__ __forceinline__ int lds_ld(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_st(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ int lds_find(int* parent, int x) {
  int cur = lds_ld(parent + x);
  if (cur == x) return x;
  int prev = x;
  for (;;) {
    const int next = lds_ld(parent + cur);
    if (next == cur) return cur;
    lds_st(parent + prev, next);
    prev = cur;
    cur = next;
  }
}
__device__ void lds_union(int* parent, int a, int b) {
  int ra = lds_find(parent, a), rb = lds_find(parent, b);
  while (ra != rb) {
    if (ra > rb) {
      const int t = ra;
      ra = rb;
      rb = t;
    }
    int expect = rb;
    if (__hip_atomic_compare_exchange_strong(parent + rb, &expect, ra, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP))
      return;
    ra = lds_find(parent, ra);
    rb = lds_find(parent, expect);
  }
}

struct SmallArgs {
  const int64_t* hits;
  int n;
  int64_t Nf;
  int64_t ncol;
  int di, dj;
  int* cid;        // out: hit ordinal of each proto-hit (when the reductions do not fold)
  int64_t* flags;  // [0] count, [1] input error, [2] 1 when the reductions folded
  int64_t cap;
  int64_t* peak;
  float* value;
  int64_t* size;
  int64_t* box;
  int64_t* label;
  int Nr;
  int64_t ld;
  Mats mats;
  int want;  // cap > 0 or labels asked for
};

__global__ __launch_bounds__(CL_SMALL_THREADS) void cl_small(SmallArgs a) {
  __shared__ int hv[CL_SMALL_MAX];   // the proto-hits, as int32 linear indices
  // the forest; after numbering, -(ordinal + 1) at the roots; at the end,
  // the hits' accumulators
  __shared__ __align__(8) int par[CL_SMALL_MAX];
  __shared__ int bad;
  __shared__ int wsum[CL_SMALL_THREADS / 64 + 1];
  const int n = a.n, tid = threadIdx.x;
  const int64_t lim = a.Nf * a.ncol;
  if (tid == 0) bad = 0;
  __syncthreads();
  {
    // every load of this thread in flight at once (at most 16 per thread)
    constexpr int PT = CL_SMALL_MAX / CL_SMALL_THREADS;
    int64_t hh[PT];
#pragma unroll
    for (int q = 0; q < PT; ++q) {
      const int k = tid + q * CL_SMALL_THREADS;
      hh[q] = k < n ? a.hits[k] : 0;
    }
#pragma unroll
    for (int q = 0; q < PT; ++q) {
      const int k = tid + q * CL_SMALL_THREADS;
      if (k < n) {
        const int64_t h = hh[q];
        if (h < 0 || h >= lim) lds_st(&bad, 1);
        hv[k] = static_cast<int>(h < 0 ? 0 : (h >= lim ? lim - 1 : h));
        par[k] = k;
      }
    }
  }
  __syncthreads();
  // order, on the LDS copy: exact once every entry is in range
  if (!bad)
    for (int k = tid + 1; k < n; k += CL_SMALL_THREADS)
      if (hv[k - 1] >= hv[k]) lds_st(&bad, 1);
  __syncthreads();
  if (bad) {  // uniform
    for (int k = tid; k < n; k += CL_SMALL_THREADS) a.cid[k] = INT_MAX;
    if (tid == 0) {
      a.flags[0] = 0;
      a.flags[1] = 1;
      a.flags[2] = 0;
    }
    return;
  }
  const int Nf = static_cast<int>(a.Nf), ncol = static_cast<int>(a.ncol);
  for (int k = tid; k < n; k += CL_SMALL_THREADS) {
    const int h = hv[k];
    const int J = h / Nf, i = h - J * Nf;
    // the same column: the next entries up to channel i + di
    const int colend = (J + 1) * Nf;  // <= Nf * ncol < 2^31
    for (int m = k + 1; m < n; ++m) {
      const int g = hv[m];
      if (g - h > a.di || g >= colend) break;
      lds_union(par, k, m);
    }
    // columns J+1 .. J+dj: channels max(0, i-di) .. min(Nf-1, i+di)
    const int ilo = i > a.di ? i - a.di : 0;
    const int ihi = i < Nf - 1 - a.di ? i + a.di : Nf - 1;
    for (int d = 1; d <= a.dj && J + d < ncol; ++d) {
      const int lo = (J + d) * Nf + ilo, hi = (J + d) * Nf + ihi;
      int L = k + 1, R = n;
      while (L < R) {
        const int mid = (L + R) >> 1;
        if (hv[mid] < lo) L = mid + 1;
        else R = mid;
      }
      for (int m = L; m < n && hv[m] <= hi; ++m) lds_union(par, k, m);
    }
  }
  __syncthreads();
  // roots for every entry
  for (int k = tid; k < n; k += CL_SMALL_THREADS) {
    int r = lds_ld(par + k);
    for (int next = lds_ld(par + r); next != r; next = lds_ld(par + r)) r = next;
    lds_st(par + k, r);
  }
  __syncthreads();
  // number the roots in order: thread t takes the entries [t*per, (t+1)*per)
  const int per = (n + CL_SMALL_THREADS - 1) / CL_SMALL_THREADS;
  const int k0 = tid * per, k1 = min(n, k0 + per);
  int s = 0;
  for (int k = k0; k < k1; ++k) s += par[k] == k ? 1 : 0;
  const int lane = tid & 63, wv = tid >> 6;
  const int incl = wave_incl_scan_i(s, lane);
  if (lane == 63) wsum[wv] = incl;
  __syncthreads();
  if (tid == 0) {
    int t = 0;
    for (int w = 0; w < CL_SMALL_THREADS / 64; ++w) {
      const int v = wsum[w];
      wsum[w] = t;
      t += v;
    }
    wsum[CL_SMALL_THREADS / 64] = t;
  }
  __syncthreads();
  const int C = wsum[CL_SMALL_THREADS / 64];
  // the hits' reductions fold into this workgroup when their accumulators
  // (24 bytes a hit) fit where the parents were; otherwise cl_accumulate and
  // cl_finish run them
  const bool fold = a.want && C <= CL_SMALL_ACC;
  int c = wsum[wv] + incl - s;
  for (int k = k0; k < k1; ++k) {
    if (par[k] != k) continue;
    if (c < a.cap) {
      a.box[4 * c + 2] = hv[k] / Nf;  // a root is its hit's first proto-hit: the column minimum
      if (!fold) {
        a.peak[c] = 0;
        a.size[c] = 0;
        a.box[4 * c + 0] = INT64_MAX;
        a.box[4 * c + 1] = 0;
        a.box[4 * c + 3] = 0;
      }
    }
    par[k] = -(c + 1);  // read only after the barrier below
    ++c;
  }
  if (tid == 0) {
    a.flags[0] = C;
    a.flags[1] = 0;
    a.flags[2] = fold ? 1 : 0;
  }
  __syncthreads();
  constexpr int PT = CL_SMALL_MAX / CL_SMALL_THREADS;
  int cc[PT];
#pragma unroll
  for (int q = 0; q < PT; ++q) {
    const int k = tid + q * CL_SMALL_THREADS;
    cc[q] = 0;
    if (k < n) {
      const int p = par[k];
      cc[q] = p < 0 ? -p - 1 : -par[p] - 1;
      if (!fold) a.cid[k] = cc[q];
    }
  }
  if (!fold) return;  // uniform
  __syncthreads();    // every read of the parents is done: their LDS takes the accumulators
  unsigned long long* akey = reinterpret_cast<unsigned long long*>(par);
  unsigned* asz = reinterpret_cast<unsigned*>(akey + C);
  int* aimn = reinterpret_cast<int*>(asz + C);
  int* aimx = aimn + C;
  int* ajmx = aimx + C;
  for (int q = tid; q < C; q += CL_SMALL_THREADS) {
    akey[q] = 0ull;
    asz[q] = 0u;
    aimn[q] = INT_MAX;
    aimx[q] = 0;
    ajmx[q] = 0;
  }
  // every FDR value of this thread in flight at once
  float vv[PT];
#pragma unroll
  for (int q = 0; q < PT; ++q) {
    const int k = tid + q * CL_SMALL_THREADS;
    vv[q] = 0.0f;
    if (k < n) {
      const int h = hv[k];
      const int J = h / Nf, i = h - J * Nf;
      const int b = J / a.Nr, j = J - b * a.Nr;
      vv[q] = a.mats.get(b)[i + a.ld * j];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PT; ++q) {
    const int k = tid + q * CL_SMALL_THREADS;
    if (k < n) {
      const int h = hv[k];
      const int J = h / Nf, i = h - J * Nf;
      const int cl = cc[q];
      // peak: the largest value, ties to the smaller ordinal (the earlier proto-hit)
      atomicMax(akey + cl, (static_cast<unsigned long long>(ordered_bits(vv[q])) << 32) |
                               (0xffffffffull - static_cast<unsigned long long>(k)));
      atomicAdd(asz + cl, 1u);
      atomicMin(aimn + cl, i);
      atomicMax(aimx + cl, i);
      atomicMax(ajmx + cl, J);
      if (a.label) a.label[k] = cl;
    }
  }
  __syncthreads();
  const int nout = a.cap < C ? static_cast<int>(a.cap) : C;
  for (int q = tid; q < nout; q += CL_SMALL_THREADS) {
    const unsigned long long key = akey[q];
    const int kk = static_cast<int>(0xffffffffull - (key & 0xffffffffull));  // every hit has a member
    a.peak[q] = hv[kk];
    a.value[q] = unordered_bits(static_cast<unsigned>(key >> 32));
    a.size[q] = asz[q];
    a.box[4 * q + 0] = aimn[q];
    a.box[4 * q + 1] = aimx[q];
    a.box[4 * q + 3] = ajmx[q];
  }
}

// Decode the peak keys of the first min(count, cap) hits in place.  A cluster
// whose only member is a proto-hit cl_link flagged (out of range) received no
// key: its key is still 0, which decodes to an ordinal past n; it gets peak -1
// and a NaN value (the call reports the input error).
__global__ __launch_bounds__(CL_THREADS) void cl_finish(const int64_t* __restrict__ hits, int64_t n,
                                                       const int64_t* __restrict__ flags, int64_t cap,
                                                       int64_t* __restrict__ peak, float* __restrict__ value,
                                                       const int64_t* __restrict__ done) {
  if (done != nullptr && *done != 0) return;  // uniform: cl_small wrote the hits
  const int64_t c = static_cast<int64_t>(blockIdx.x) * CL_THREADS + threadIdx.x;
  const int64_t ncl = flags[0] < cap ? flags[0] : cap;
  if (c >= ncl) return;
  const unsigned long long kk = static_cast<unsigned long long>(peak[c]);
  const int64_t k = static_cast<int64_t>(0xffffffffull - (kk & 0xffffffffull));
  if (k >= n) {
    peak[c] = -1;
    value[c] = __uint_as_float(0x7fc00000u);
    return;
  }
  peak[c] = hits[k];
  value[c] = unordered_bits(static_cast<unsigned>(kk >> 32));
}

unsigned blocks_for(int64_t n, int64_t per) { return static_cast<unsigned>((n + per - 1) / per); }

}  // namespace

int cluster_protohits_impl(const int64_t* hits, int64_t n, const float* const* fdrs, int nb, int64_t Nf,
                           int64_t Nr, int64_t ld, int di, int dj, int64_t* peak, float* value, int64_t* size,
                           int64_t* box, int64_t cap, int64_t* nclusters, int64_t* label, hipStream_t st) {
  if (n == 0) {
    *nclusters = 0;
    return ZDT_OK;
  }
  const int64_t ncol = static_cast<int64_t>(nb) * Nr;
  // few proto-hits: one workgroup in LDS (cl_small) instead of cl_init ..
  // cl_number_roots and the scan
  const bool small = n <= CL_SMALL_MAX && Nf * ncol < (int64_t(1) << 31);
  const int64_t nt = small ? 0 : (n + CL_TILE - 1) / CL_TILE;
  const int64_t nscan = small ? 0 : hits_scan_scratch(nt);
  // one scratch block, no other allocation and one synchronization (at the
  // end, for the count and the error flag):
  //   parent[n], cid[n] (int); counts[nt] + scan scratch, flags[3],
  //   colstart[ncol + 1] (dj > 0, general path), the matrix table (nb > 1)
  //   (int64 words)
  auto up = [](int64_t bytes) { return (bytes + 255) & ~int64_t(255); };
  const int64_t o_cid = up(4 * n), o_w = o_cid + up(4 * n);
  const int64_t o_flags = o_w + 8 * (nt + nscan), o_col = o_flags + 24;
  const int64_t o_tab = o_col + (dj > 0 && !small ? 8 * (ncol + 1) : 0);
  const int64_t bytes = o_tab + (nb > 1 ? 8 * static_cast<int64_t>(nb) : 0);
  char* buf = nullptr;
  ZDT_TRY_HIP(hipMallocAsync(reinterpret_cast<void**>(&buf), static_cast<size_t>(bytes), st));
  int* parent = reinterpret_cast<int*>(buf);
  int* cid = reinterpret_cast<int*>(buf + o_cid);
  int64_t* counts = reinterpret_cast<int64_t*>(buf + o_w);
  int64_t* flags_dev = reinterpret_cast<int64_t*>(buf + o_flags);
  int64_t* colstart = dj > 0 && !small ? reinterpret_cast<int64_t*>(buf + o_col) : nullptr;
  Mats mats{fdrs[0], nb > 1 ? reinterpret_cast<const float* const*>(buf + o_tab) : nullptr};
  int64_t flags[2] = {0, 0};
  auto run = [&]() -> int {
    if (nb > 1)
      ZDT_TRY_HIP(hipMemcpyAsync(buf + o_tab, fdrs, sizeof(float*) * nb, hipMemcpyHostToDevice, st));
    if (small) {
      ProfScope ps(st, "cluster_small", 16.0 * n);
      SmallArgs sa{hits, static_cast<int>(n), Nf, ncol, di, dj, cid, flags_dev, cap, peak, value, size, box, label,
                   static_cast<int>(Nr), ld, mats, cap > 0 || label != nullptr ? 1 : 0};
      cl_small<<<1, CL_SMALL_THREADS, 0, st>>>(sa);
      ZDT_CHECK_LAUNCH("cl_small");
    } else {
      {
        ProfScope ps(st, "cluster_link", 20.0 * n);
        InitArgs ia{hits, n, Nf, ncol, parent, colstart, flags_dev};
        cl_init<<<blocks_for(n, CL_THREADS), CL_THREADS, 0, st>>>(ia);
        ZDT_CHECK_LAUNCH("cl_init");
        LinkArgs a{hits, n, Nf, ncol, di, dj, parent, colstart, flags_dev};
        cl_link<<<blocks_for(n, CL_THREADS), CL_THREADS, 0, st>>>(a);
        ZDT_CHECK_LAUNCH("cl_link");
      }
      {
        ProfScope ps(st, "cluster_roots", 8.0 * n);
        cl_flatten_count<<<static_cast<unsigned>(nt), CL_THREADS, 0, st>>>(parent, n, counts);
        ZDT_CHECK_LAUNCH("cl_flatten_count");
      }
      ZDT_TRY(launch_hits_scan(counts, nt, flags_dev, counts + nt, st));
      {
        ProfScope ps(st, "cluster_number", 12.0 * n);
        NumberArgs a{hits, n, Nf, parent, counts, cid, cap, peak, size, box};
        cl_number_roots<<<static_cast<unsigned>(nt), CL_THREADS, 0, st>>>(a);
        ZDT_CHECK_LAUNCH("cl_number_roots");
      }
    }
    if (cap > 0 || label != nullptr) {
      ProfScope ps(st, "cluster_reduce", 28.0 * n);
      AccumArgs a{hits, n, Nf, Nr, ld, ncol, mats, small ? nullptr : parent, cid, cap, peak, size, box,
                  label, small ? flags_dev + 2 : nullptr};
      cl_accumulate<<<blocks_for(n, CL_THREADS), CL_THREADS, 0, st>>>(a);
      ZDT_CHECK_LAUNCH("cl_accumulate");
    }
    if (cap > 0) {
      cl_finish<<<blocks_for(cap < n ? cap : n, CL_THREADS), CL_THREADS, 0, st>>>(
          hits, n, flags_dev, cap, peak, value, small ? flags_dev + 2 : nullptr);
      ZDT_CHECK_LAUNCH("cl_finish");
    }
    ZDT_TRY_HIP(hipMemcpyAsync(flags, flags_dev, 16, hipMemcpyDeviceToHost, st));
    ZDT_TRY_HIP(hipStreamSynchronize(st));
    return ZDT_OK;
  };
  const int rc = run();
  (void)hipFreeAsync(buf, st);
  if (rc != ZDT_OK) return rc;
  if (flags[1] != 0)
    return fail(ZDT_EINVAL, "clusterprotohits: proto-hits must be strictly ascending linear indices below Nf*nb*Nr");
  *nclusters = flags[0];
  if (flags[0] > cap)
    return fail(ZDT_ETRUNC, "clusterprotohits: %lld hits exceed capacity %lld", (long long)flags[0], (long long)cap);
  return ZDT_OK;
}

}  // namespace zdt
