// C-ABI of libzdt_hip.so: workspace management and the entry points declared
// in include/zdt.h.  Each entry point names the reference method it replaces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fft_generic.h"
#include "zdt_fast.h"
#include "zdt_ws.h"
#include "zdt_internal.h"
#include "zdt_kernels.h"

namespace zdt {

namespace {
thread_local std::string g_err;
}

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// ---------------------------------------------------------------- profiler

struct Profiler {
  struct Rec {
    const char* name;
    double bytes;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  struct Agg {
    std::string name;
    double ms = 0, bytes = 0;
    int64_t n = 0;
  };
  std::vector<Agg> agg;

  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    // timing-only events: no system-scope cache writeback/invalidate at the
    // record, which would otherwise stall the queue between the kernels
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    return e;
  }
  void reset() {
    for (auto& r : recs) {
      if (r.a) pool.push_back(r.a);
      if (r.b) pool.push_back(r.b);
    }
    recs.clear();
    agg.clear();
  }
  ~Profiler() {
    reset();
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

thread_local Profiler* g_prof = nullptr;

void prof_begin(hipStream_t st, const char* name, double bytes) {
  Profiler* p = g_prof;
  Profiler::Rec r{name, bytes, p->get(), nullptr};
  if (r.a) (void)hipEventRecord(r.a, st);
  p->recs.push_back(r);
}

void prof_end(hipStream_t st) {
  Profiler* p = g_prof;
  if (p->recs.empty()) return;
  Profiler::Rec& r = p->recs.back();
  r.b = p->get();
  if (r.b) (void)hipEventRecord(r.b, st);
}

int hip_fail(hipError_t e, const char* what) {
  if (e == hipErrorOutOfMemory) return fail(ZDT_ENOMEM, "%s: %s", what, hipGetErrorString(e));
  return fail(ZDT_EHIP, "%s: %s", what, hipGetErrorString(e));
}

// ---------------------------------------------------------------- sizing

static bool smooth(int64_t n, const std::vector<int>& f) {
  for (int p : f) {
    while (n > 1 && n % p == 0) n /= p;
  }
  return n == 1;
}

static int parse_factors(const int* factors, int nfactors, std::vector<int>* out) {
  out->clear();
  if (factors == nullptr || nfactors <= 0) {
    *out = {2, 3, 5};
    return ZDT_OK;
  }
  for (int i = 0; i < nfactors; ++i) {
    if (factors[i] < 2) return fail(ZDT_EINVAL, "factors must be integers >= 2 (got %d)", factors[i]);
    out->push_back(factors[i]);
  }
  return ZDT_OK;
}

// nextprod(factors, n) (Julia Base), used by calcNl (src/zdtutils.jl:7-9)
static int64_t nextprod(const std::vector<int>& f, int64_t n) {
  if (n < 1) n = 1;
  for (int64_t m = n;; ++m) {
    if (smooth(m, f)) return m;
  }
}

}  // namespace zdt

using namespace zdt;

// ---------------------------------------------------------------- workspace

struct zdt_ws {
  int device = 0;
  int64_t Nf = 0, Nt = 0, Nr = 0, Nl = 0, Nk = 0;
  float r0 = 0, dr = 0;
  std::vector<int> factors;
  int flags = 0;
  hipStream_t stream = nullptr;
  double hd = 0;  // dr/(2 Nf)

  // generic path
  c32* F = nullptr;   // Nk x Nt
  c32* Ys = nullptr;  // Nk x Nr
  void* s[3] = {nullptr, nullptr, nullptr};
  size_t s_bytes = 0;

  // step functions (materialized Y and V)
  c32* Y = nullptr;  // Nk x Nl
  c32* V = nullptr;  // Nk x Nl
  bool have_V = false;

  // fused fast path (power-of-two Nf)
  FastPlan* fast = nullptr;
  // fused output for the generic chirp-Z (power-of-two Nf, 16 < Nt <= FOUT_MAX_NT)
  FastPlan* fout = nullptr;
  c32* Ft = nullptr;  // its input spectrum, Nk x Nt entries in the transposed layout
  bool ys_fout = false;  // last run left its spectra in fout

  bool have_input = false;
  const c32* ys_src = nullptr;  // post-processed spectra for zdt_output
  int64_t ys_ld = 0;
  bool ys_fast = false;  // last run left its result in the fast path's buffers

  int64_t bytes = 0;
  Profiler* prof = nullptr;
  int maxr = 16;  // radix cap of the generic FFTs (autotuned by plans)
  int precision = ZDT_PRECISION_FP32;
};

Profiler* zdt::ws_profiler(const zdt_ws* ws) { return ws ? ws->prof : nullptr; }

namespace {

// installs the workspace's profiler (if any) for the duration of an entry point
struct ProfInstall {
  Profiler* prev;
  explicit ProfInstall(const zdt_ws* ws) : prev(g_prof) { g_prof = ws ? ws->prof : nullptr; }
  ~ProfInstall() { g_prof = prev; }
};

int dev_alloc(zdt_ws* ws, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) return ZDT_OK;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(ZDT_ENOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
  }
  ws->bytes += static_cast<int64_t>(bytes);
  return ZDT_OK;
}

void dev_free(void* p) {
  if (p) (void)hipFree(p);
}

int ensure_scratch(zdt_ws* ws, size_t bytes) {
  if (ws->s_bytes >= bytes) return ZDT_OK;
  for (int i = 0; i < 3; ++i) {
    dev_free(ws->s[i]);
    ws->s[i] = nullptr;
  }
  ws->bytes -= static_cast<int64_t>(3 * ws->s_bytes);
  ws->s_bytes = 0;
  for (int i = 0; i < 3; ++i) ZDT_TRY(dev_alloc(ws, &ws->s[i], bytes));
  ws->s_bytes = bytes;
  return ZDT_OK;
}

size_t real_scratch_bytes(const zdt_ws* ws) {
  const int64_t b = std::max(ws->Nt, ws->Nr);
  return fft_real_scratch_elems(ws->Nf, b) * sizeof(c32);
}

int check_ws(const zdt_ws* ws) {
  if (ws == nullptr) return fail(ZDT_EINVAL, "workspace is NULL");
  return ZDT_OK;
}

int set_device(const zdt_ws* ws) {
  ZDT_TRY_HIP(hipSetDevice(ws->device));
  return ZDT_OK;
}

int check_window(int wkind, const float* window) {
  if (wkind != ZDT_WINDOW_RECT && wkind != ZDT_WINDOW_HAMMING && wkind != ZDT_WINDOW_CUSTOM)
    return fail(ZDT_EINVAL, "unsupported window type (%d)", wkind);
  if (wkind == ZDT_WINDOW_CUSTOM && window == nullptr)
    return fail(ZDT_EINVAL, "custom window requires a device vector of Nf/2+1 values");
  return ZDT_OK;
}

double out_scale(const zdt_ws* ws) {
  return (ws->flags & ZDT_FLAG_UNNORMALIZED) ? static_cast<double>(ws->Nf) * static_cast<double>(ws->Nl) : 1.0;
}

bool use_fast(const zdt_ws* ws) { return ws->fast != nullptr; }

int ensure_stepwise(zdt_ws* ws) {
  if (ws->Y == nullptr) {
    ZDT_TRY(dev_alloc(ws, reinterpret_cast<void**>(&ws->Y), static_cast<size_t>(ws->Nk * ws->Nl) * sizeof(c32)));
  }
  ZDT_TRY(ensure_scratch(ws, std::max(real_scratch_bytes(ws), static_cast<size_t>(ws->Nk * ws->Nl) * sizeof(c32))));
  return ZDT_OK;
}

// V = fft_l(chirp kernel), computed in FP64 (src/zdtfdr.jl:173-187)
int ensure_V(zdt_ws* ws) {
  if (ws->have_V) return ZDT_OK;
  ZDT_TRY(ensure_stepwise(ws));
  if (ws->V == nullptr) {
    ZDT_TRY(dev_alloc(ws, reinterpret_cast<void**>(&ws->V), static_cast<size_t>(ws->Nk * ws->Nl) * sizeof(c32)));
  }
  const FftPlan* p = nullptr;
  ZDT_TRY(fft_plan_get(ws->Nl, true, &p));
  const size_t n = static_cast<size_t>(ws->Nk * ws->Nl);
  void *v64 = nullptr, *a = nullptr, *b = nullptr;
  int rc = ZDT_OK;
  hipStream_t st = ws->stream;
  if (hipMallocAsync(&v64, n * sizeof(c64), st) != hipSuccess || hipMallocAsync(&a, n * sizeof(c64), st) != hipSuccess ||
      hipMallocAsync(&b, n * sizeof(c64), st) != hipSuccess) {
    (void)hipGetLastError();
    rc = fail(ZDT_ENOMEM, "allocation of the FP64 chirp-kernel buffers (%zu bytes) failed", 3 * n * sizeof(c64));
  }
  if (rc == ZDT_OK) rc = launch_step_vgen(static_cast<c64*>(v64), ws->Nk, (int)ws->Nt, (int)ws->Nr, (int)ws->Nl, ws->hd, st);
  if (rc == ZDT_OK) rc = fft_c2c(p, v64, ws->Nk, 1, v64, ws->Nk, 1, ws->Nk, false, 1.0, a, b, st);
  if (rc == ZDT_OK) rc = launch_c64_to_c32(static_cast<const c64*>(v64), ws->V, static_cast<int64_t>(n), st);
  if (v64) (void)hipFreeAsync(v64, st);
  if (a) (void)hipFreeAsync(a, st);
  if (b) (void)hipFreeAsync(b, st);
  if (rc == ZDT_OK) ws->have_V = true;
  return rc;
}

int output_impl(zdt_ws* ws, float* dest) {
  if (ws->ys_fast) return fast_output(ws->fast, dest, out_scale(ws), ws->stream);
  if (ws->ys_fout) return fast_output(ws->fout, dest, out_scale(ws), ws->stream);
  if (ws->ys_src == nullptr) return fail(ZDT_ESTATE, "output!: no post-processed spectra (run zdtfdr! or the steps first)");
  ZDT_TRY(ensure_scratch(ws, real_scratch_bytes(ws)));
  return fft_c2r(ws->Nf, ws->ys_src, ws->ys_ld, dest, ws->Nf, ws->Nr, out_scale(ws), ws->s[0], ws->s[1], ws->s[2],
                 ws->stream, ws->maxr);
}

int run_impl(zdt_ws* ws, float r0, int wkind, const float* window, float* dest) {
  if (!ws->have_input) return fail(ZDT_ESTATE, "zdtfdr!: no input (call input! or pass a spectrogram)");
  ZDT_TRY(check_window(wkind, window));
  if (use_fast(ws)) {
    ZDT_TRY(fast_run(ws->fast, r0, wkind, window, dest, out_scale(ws), ws->stream));
    ws->ys_fast = (dest == nullptr);
    ws->ys_src = nullptr;
    return ZDT_OK;
  }
  if (ws->fout != nullptr) {
    ZDT_TRY(fast_out_run(ws->fout, ws->Ft, ws->Nk, true, static_cast<int>(ws->Nt), r0, ws->dr, wkind, window, dest,
                         out_scale(ws), ws->stream));
    ws->ys_fast = false;
    ws->ys_fout = (dest == nullptr);
    ws->ys_src = nullptr;
    return ZDT_OK;
  }
  CztArgs a;
  a.F = ws->F;
  a.ldF = ws->Nk;
  a.Ys = ws->Ys;
  a.ldY = ws->Nk;
  a.Nk = ws->Nk;
  a.Nf = ws->Nf;
  a.Nt = static_cast<int>(ws->Nt);
  a.Nr = static_cast<int>(ws->Nr);
  a.hd = ws->hd;
  a.h0 = static_cast<double>(r0) / static_cast<double>(ws->Nf);
  {
    // rates in mirrored pairs around the centre rate (a single rate is its own partner)
    const double rc = static_cast<double>(r0) + 0.5 * static_cast<double>(ws->Nr - 1) * static_cast<double>(ws->dr);
    a.hc = rc / static_cast<double>(ws->Nf);
    a.hz = (static_cast<double>(r0) - rc) / static_cast<double>(ws->Nf);
  }
  a.wkind = wkind;
  a.wcustom = window;
  a.scale = 1.0f;
  ZDT_TRY(launch_czt(a, ws->stream));
  ws->ys_src = ws->Ys;
  ws->ys_ld = ws->Nk;
  ws->ys_fast = false;
  ws->ys_fout = false;
  if (dest != nullptr) ZDT_TRY(output_impl(ws, dest));
  return ZDT_OK;
}

}  // namespace

extern "C" {

const char* zdt_last_error(void) { return g_err.c_str(); }

int zdt_version(void) { return 2; }

int64_t zdt_calc_nl(int64_t Nt, int64_t Nr, const int* factors, int nfactors) {
  std::vector<int> f;
  if (parse_factors(factors, nfactors, &f) != ZDT_OK || Nt < 1 || Nr < 1) return -1;
  return nextprod(f, Nt + Nr - 1);
}

int64_t zdt_grow_nr(int64_t Nt, int64_t Nr, const int* factors, int nfactors) {
  const int64_t Nl = zdt_calc_nl(Nt, Nr, factors, nfactors);
  return Nl < 0 ? -1 : Nl - Nt + 1;
}

int64_t zdt_estimate_memory(int64_t Nf, int64_t Nt, int64_t Nr, int64_t Ni, int64_t No, const int* factors,
                            int nfactors) {
  const int64_t Nl = zdt_calc_nl(Nt, Nr, factors, nfactors);
  if (Nl < 0) return -1;
  Ni += 3;
  const int64_t Nz = 3;
  return 4 * ((Nf * Nt) * Ni + (Nf * Nl) * Nz + (Nf * Nr) * No);
}

int64_t zdt_workspace_bytes(int64_t Nf, int64_t Nt, int64_t Nr, const int* factors, int nfactors, int flags) {
  const int64_t Nl = zdt_calc_nl(Nt, Nr, factors, nfactors);
  if (Nl < 0 || Nf < 1) return -1;
  const int64_t Nk = Nf / 2 + 1;
  if (!(flags & ZDT_FLAG_GENERIC) && fast_supported(Nf, Nt, Nr)) return fast_bytes(Nf, Nt, Nr);
  if (!(flags & ZDT_FLAG_GENERIC) && fast_out_supported(Nf, Nt, Nr)) return Nk * Nt * 8 + fast_out_bytes(Nf, Nr);
  const int64_t b = std::max(Nt, Nr);
  const int64_t sc = static_cast<int64_t>(fft_real_scratch_elems(Nf, b) * sizeof(c32));
  return Nk * Nt * 8 + Nk * Nr * 8 + 3 * sc;
}

}  // extern "C"

namespace zdt {

// zdt_create with the tunables the plan autotuner chooses: n2 (the fused
// path's strided FFT length, 0 = default) and maxr (generic FFT radix cap).
int create_ws_tuned(const float* spec, int64_t Nf, int64_t Nt, float r0, float dr, int64_t Nr, const int* factors,
                    int nfactors, int flags, int precision, void* stream, int n2, int maxr, zdt_ws** out) {
  if (out == nullptr) return fail(ZDT_EINVAL, "out is NULL");
  *out = nullptr;
  if (precision != ZDT_PRECISION_FP32 && precision != ZDT_PRECISION_SPLIT_BF16 &&
      precision != ZDT_PRECISION_SPLIT_FP16)
    return fail(ZDT_EINVAL, "unknown precision %d", precision);
  if (Nf < 1 || Nt < 1 || Nr < 1)
    return fail(ZDT_EINVAL, "invalid shape Nf=%lld Nt=%lld Nr=%lld", (long long)Nf, (long long)Nt, (long long)Nr);
  if (Nt > (1 << 20) || Nr > (1 << 24)) return fail(ZDT_ENOTSUP, "Nt or Nr too large");
  if (!std::isfinite(r0) || !std::isfinite(dr)) return fail(ZDT_EINVAL, "r0 and dr must be finite");
  std::vector<int> f;
  ZDT_TRY(parse_factors(factors, nfactors, &f));
  std::vector<int> rad;
  if (!fft_factorize(Nf, &rad))
    return fail(ZDT_ENOTSUP, "Nf=%lld has a prime factor outside {2,3,5,7,11,13}", (long long)Nf);
  zdt_ws* ws = new zdt_ws();
  ws->Nf = Nf;
  ws->Nt = Nt;
  ws->Nr = Nr;
  ws->Nk = Nf / 2 + 1;
  ws->Nl = nextprod(f, Nt + Nr - 1);
  ws->r0 = r0;
  ws->dr = dr;
  ws->factors = f;
  ws->flags = flags & ~ZDT_FLAG_NO_FUSED_INTERNAL;
  ws->stream = static_cast<hipStream_t>(stream);
  ws->hd = static_cast<double>(dr) / (2.0 * static_cast<double>(Nf));
  int rc = ZDT_OK;
  if (hipGetDevice(&ws->device) != hipSuccess) rc = fail(ZDT_EHIP, "hipGetDevice failed");
  if (flags & ZDT_FLAG_PROFILE) ws->prof = new Profiler();
  ws->maxr = maxr;
  ws->precision = precision;
  const bool fused = !(flags & (ZDT_FLAG_GENERIC | ZDT_FLAG_NO_FUSED_INTERNAL)) && fast_supported(Nf, Nt, Nr);
  const bool fout_ok = !(flags & ZDT_FLAG_GENERIC) && !fused && fast_out_supported(Nf, Nt, Nr);
  if (rc == ZDT_OK && precision != ZDT_PRECISION_FP32 &&
      !((fused && fast_precision_ok(Nf, n2, precision)) || (fout_ok && fast_precision_ok(Nf, 0, precision))))
    rc = fail(ZDT_ENOTSUP, "precision %d needs the fused path or the fused output with N2 = 1024 "
              "(power-of-two Nf in [2^17, 2^22]); Nf=%lld Nt=%lld", precision, (long long)Nf, (long long)Nt);
  if (rc == ZDT_OK && fused) {
    rc = fast_create(Nf, Nt, Nr, dr, n2, precision, &ws->fast);
    if (rc == ZDT_OK) ws->bytes += fast_bytes(Nf, Nt, Nr);
  } else if (rc == ZDT_OK) {
    const bool fout = !(flags & ZDT_FLAG_GENERIC) && fast_out_supported(Nf, Nt, Nr);
    if (fout) {
      // the natural-layout F is unpacked from Ft only for the step functions
      rc = dev_alloc(ws, reinterpret_cast<void**>(&ws->Ft), static_cast<size_t>(ws->Nk * Nt) * sizeof(c32));
      if (rc == ZDT_OK) rc = fast_out_create(Nf, Nr, precision, &ws->fout);
      if (rc == ZDT_OK) ws->bytes += fast_out_bytes(Nf, Nr);
    } else {
      rc = dev_alloc(ws, reinterpret_cast<void**>(&ws->F), static_cast<size_t>(ws->Nk * Nt) * sizeof(c32));
      if (rc == ZDT_OK)
        rc = dev_alloc(ws, reinterpret_cast<void**>(&ws->Ys), static_cast<size_t>(ws->Nk * Nr) * sizeof(c32));
      if (rc == ZDT_OK) rc = ensure_scratch(ws, real_scratch_bytes(ws));
    }
    if (rc == ZDT_OK) {
      const FftPlan* p = nullptr;
      rc = fft_plan_get(Nf, false, &p, maxr);
      if (rc == ZDT_OK && Nf % 2 == 0 && Nf > 2) rc = fft_plan_get(Nf / 2, false, &p, maxr);
    }
  }
  if (rc == ZDT_OK && spec != nullptr) rc = zdt_input(ws, spec);
  if (rc != ZDT_OK) {
    std::string msg = g_err;
    zdt_destroy(ws);
    g_err = msg;
    return rc;
  }
  *out = ws;
  return ZDT_OK;
}

bool ws_is_fused(const zdt_ws* ws) { return ws != nullptr && ws->fast != nullptr; }

int64_t devstats_scratch_doubles(int64_t Nf, int64_t Nr) { return 2 * static_cast<int64_t>(colstats_nchunk(Nf)) * Nr; }

int run_devstats(zdt_ws* ws, float r0, float* dest, double* colmean, double* colvar, double* part,
                 float* smax) {
  ZDT_TRY(check_ws(ws));
  ProfInstall prof_(ws);
  if (dest == nullptr || colmean == nullptr || colvar == nullptr)
    return fail(ZDT_EINVAL, "zdtfdr! with statistics: dest and statistics buffers are required");
  ZDT_TRY(set_device(ws));
  if (!ws->have_input) return fail(ZDT_ESTATE, "zdtfdr!: no input (call input! or pass a spectrogram)");
  // the run maxima ride on this one run's k3
  struct RunMax {
    FastPlan* p;
    RunMax(FastPlan* q, float* s) : p(q) { fast_set_smax(p, s); }
    ~RunMax() { fast_set_smax(p, nullptr); }
  };
  if (use_fast(ws)) {
    RunMax rm_(ws->fast, smax);
    ZDT_TRY(fast_run(ws->fast, r0, ZDT_WINDOW_RECT, nullptr, dest, out_scale(ws), ws->stream, colmean, colvar));
    ws->ys_fast = false;
    ws->ys_fout = false;
    ws->ys_src = nullptr;
    return ZDT_OK;
  }
  if (ws->fout != nullptr) {
    RunMax rm_(ws->fout, smax);
    ZDT_TRY(fast_out_run(ws->fout, ws->Ft, ws->Nk, true, static_cast<int>(ws->Nt), r0, ws->dr, ZDT_WINDOW_RECT,
                         nullptr, dest, out_scale(ws), ws->stream, colmean, colvar));
    ws->ys_fast = false;
    ws->ys_fout = false;
    ws->ys_src = nullptr;
    return ZDT_OK;
  }
  if (smax != nullptr) return fail(ZDT_EINVAL, "zdtfdr! with run maxima: the generic path has none");
  if (part == nullptr) return fail(ZDT_EINVAL, "zdtfdr! with statistics: the generic path needs scratch");
  ZDT_TRY(run_impl(ws, r0, ZDT_WINDOW_RECT, nullptr, dest));
  return launch_colstats2(dest, ws->Nf, ws->Nr, ws->Nf, part, colmean, colvar, ws->stream);
}

int ws_path(const zdt_ws* ws) {
  if (ws == nullptr) return 0;
  return ws->fast ? 1 : (ws->fout ? 2 : 0);
}

int ws_fused_n2(const zdt_ws* ws) {
  if (ws == nullptr) return 0;
  return ws->fast ? fast_plan_n2(ws->fast) : (ws->fout ? fast_plan_n2(ws->fout) : 0);
}

}  // namespace zdt

extern "C" {

int zdt_create(const float* spec, int64_t Nf, int64_t Nt, float r0, float dr, int64_t Nr, const int* factors,
               int nfactors, int flags, int precision, void* stream, zdt_ws** out) {
  int n2 = 0;
  if (flags & ZDT_FLAG_N2_512) {
    if (!zdt::fast_n2_ok(Nf, 512) || (Nf & (Nf - 1)) != 0)
      return zdt::fail(ZDT_ENOTSUP, "ZDT_FLAG_N2_512: no fused N2 = 512 variant for Nf=%lld", (long long)Nf);
    n2 = 512;
  }
  const int maxr = (flags & ZDT_FLAG_RADIX8) ? 8 : 16;
  return zdt::create_ws_tuned(spec, Nf, Nt, r0, dr, Nr, factors, nfactors, flags & ~(ZDT_FLAG_N2_512 | ZDT_FLAG_RADIX8),
                              precision, stream, n2, maxr, out);
}

int zdt_destroy(zdt_ws* ws) {
  if (ws == nullptr) return ZDT_OK;
  (void)hipSetDevice(ws->device);
  if (ws->stream) (void)hipStreamSynchronize(ws->stream);
  else (void)hipDeviceSynchronize();
  dev_free(ws->F);
  dev_free(ws->Ft);
  dev_free(ws->Ys);
  dev_free(ws->Y);
  dev_free(ws->V);
  for (int i = 0; i < 3; ++i) dev_free(ws->s[i]);
  if (ws->fast) fast_destroy(ws->fast);
  if (ws->fout) fast_destroy(ws->fout);
  delete ws->prof;
  delete ws;
  return ZDT_OK;
}

int zdt_info(const zdt_ws* ws, int64_t* Nf, int64_t* Nt, float* r0, float* dr, int64_t* Nr, int64_t* Nl) {
  ZDT_TRY(check_ws(ws));
  if (Nf) *Nf = ws->Nf;
  if (Nt) *Nt = ws->Nt;
  if (r0) *r0 = ws->r0;
  if (dr) *dr = ws->dr;
  if (Nr) *Nr = ws->Nr;
  if (Nl) *Nl = ws->Nl;
  return ZDT_OK;
}

int zdt_path(const zdt_ws* ws, int* path, int* precision) {
  ZDT_TRY(check_ws(ws));
  if (path) *path = ws_path(ws);
  if (precision) *precision = ws->precision;
  return ZDT_OK;
}

int zdt_sizeof(const zdt_ws* ws, int64_t* bytes) {
  ZDT_TRY(check_ws(ws));
  if (bytes) *bytes = ws->bytes;
  return ZDT_OK;
}

int zdt_set_stream(zdt_ws* ws, void* stream) {
  ZDT_TRY(check_ws(ws));
  ws->stream = static_cast<hipStream_t>(stream);
  return ZDT_OK;
}

int zdt_sync(zdt_ws* ws) {
  ZDT_TRY(check_ws(ws));
  ZDT_TRY(set_device(ws));
  ZDT_TRY_HIP(hipStreamSynchronize(ws->stream));
  return ZDT_OK;
}

int zdt_input(zdt_ws* ws, const float* spec) {
  ZDT_TRY(check_ws(ws));
  ProfInstall prof_(ws);
  if (spec == nullptr) return fail(ZDT_EINVAL, "input!: spectrogram is NULL");
  ZDT_TRY(set_device(ws));
  if (use_fast(ws)) {
    ZDT_TRY(fast_input(ws->fast, spec, ws->stream));
    if (ws->F != nullptr) ZDT_TRY(fast_unpack_F(ws->fast, ws->F, ws->stream));
  } else if (ws->fout != nullptr) {
    ZDT_TRY(fast_out_input(ws->fout, spec, ws->Nt, ws->Ft, ws->stream));
    if (ws->F != nullptr) ZDT_TRY(fast_out_unpack_F(ws->fout, ws->Ft, ws->Nt, ws->F, ws->stream));
  } else {
    ZDT_TRY(fft_r2c(ws->Nf, spec, ws->Nf, ws->F, ws->Nk, ws->Nt, ws->s[0], ws->s[1], ws->s[2], ws->stream, ws->maxr));
  }
  ws->have_input = true;
  return ZDT_OK;
}

int zdt_run(zdt_ws* ws, float r0, int window_kind, const float* window, float* dest) {
  ZDT_TRY(check_ws(ws));
  ProfInstall prof_(ws);
  ZDT_TRY(set_device(ws));
  return run_impl(ws, r0, window_kind, window, dest);
}

int zdt_run_batches(zdt_ws* ws, const float* r0s, int nb, int window_kind, const float* window, float* const* dests) {
  ZDT_TRY(check_ws(ws));
  ProfInstall prof_(ws);
  if (nb < 0 || (nb > 0 && (r0s == nullptr || dests == nullptr)))
    return fail(ZDT_EINVAL, "zdtfdr!: bad batch arrays");
  ZDT_TRY(set_device(ws));
  for (int b = 0; b < nb; ++b) {
    if (dests[b] == nullptr) return fail(ZDT_EINVAL, "zdtfdr!: dests[%d] is NULL", b);
    ZDT_TRY(run_impl(ws, r0s[b], window_kind, window, dests[b]));
  }
  return ZDT_OK;
}

int zdt_output(zdt_ws* ws, float* dest) {
  ZDT_TRY(check_ws(ws));
  ProfInstall prof_(ws);
  if (dest == nullptr) return fail(ZDT_EINVAL, "output!: dest is NULL");
  ZDT_TRY(set_device(ws));
  return output_impl(ws, dest);
}

// F in natural layout for the step functions (the fused paths keep their own
// layouts and unpack on demand; input! refreshes an unpacked F)
static int ensure_F_natural(zdt_ws* ws) {
  if (ws->F != nullptr) return ZDT_OK;
  ZDT_TRY(dev_alloc(ws, reinterpret_cast<void**>(&ws->F), static_cast<size_t>(ws->Nk * ws->Nt) * sizeof(c32)));
  if (ws->have_input) {
    if (use_fast(ws)) ZDT_TRY(fast_unpack_F(ws->fast, ws->F, ws->stream));
    else if (ws->fout != nullptr) ZDT_TRY(fast_out_unpack_F(ws->fout, ws->Ft, ws->Nt, ws->F, ws->stream));
  }
  return ZDT_OK;
}

int zdt_preprocess(zdt_ws* ws, float r0) {
  ZDT_TRY(check_ws(ws));
  ProfInstall prof_(ws);
  ZDT_TRY(set_device(ws));
  if (!ws->have_input) return fail(ZDT_ESTATE, "preprocess!: no input");
  ZDT_TRY(ensure_F_natural(ws));
  ZDT_TRY(ensure_stepwise(ws));
  const double h0 = static_cast<double>(r0) / static_cast<double>(ws->Nf);
  return launch_step_prephase(ws->F, ws->Nk, ws->Y, ws->Nk, ws->Nk, (int)ws->Nt, (int)ws->Nl, ws->hd, h0, ws->stream);
}

int zdt_convolve(zdt_ws* ws) {
  ZDT_TRY(check_ws(ws));
  ProfInstall prof_(ws);
  ZDT_TRY(set_device(ws));
  ZDT_TRY(ensure_V(ws));
  const FftPlan* p = nullptr;
  ZDT_TRY(fft_plan_get(ws->Nl, false, &p));
  ZDT_TRY(fft_c2c(p, ws->Y, ws->Nk, 1, ws->Y, ws->Nk, 1, ws->Nk, false, 1.0, ws->s[0], ws->s[1], ws->stream));
  ZDT_TRY(launch_step_vmul(ws->Y, ws->V, ws->Nk * ws->Nl, ws->stream));
  ZDT_TRY(fft_c2c(p, ws->Y, ws->Nk, 1, ws->Y, ws->Nk, 1, ws->Nk, true, 1.0 / static_cast<double>(ws->Nl), ws->s[0],
                  ws->s[1], ws->stream));
  return ZDT_OK;
}

int zdt_postprocess(zdt_ws* ws, int window_kind, const float* window) {
  ZDT_TRY(check_ws(ws));
  ProfInstall prof_(ws);
  ZDT_TRY(check_window(window_kind, window));
  ZDT_TRY(set_device(ws));
  ZDT_TRY(ensure_stepwise(ws));
  ZDT_TRY(launch_step_postphase(ws->Y, ws->Nk, ws->Nk, ws->Nf, (int)ws->Nr, ws->hd, window_kind, window, 1.0f,
                                ws->stream));
  ws->ys_src = ws->Y;
  ws->ys_ld = ws->Nk;
  ws->ys_fast = false;
  ws->ys_fout = false;
  return ZDT_OK;
}

// ---------------------------------------------------------------- fftfdr / fdshift

namespace {
struct AsyncBufs {
  hipStream_t st;
  std::vector<void*> p;
  explicit AsyncBufs(hipStream_t s) : st(s) {}
  int get(void** out, size_t bytes) {
    *out = nullptr;
    if (hipMallocAsync(out, bytes, st) != hipSuccess) {
      (void)hipGetLastError();
      *out = nullptr;
      return fail(ZDT_ENOMEM, "temporary allocation of %zu bytes failed", bytes);
    }
    p.push_back(*out);
    return ZDT_OK;
  }
  ~AsyncBufs() {
    for (void* q : p) (void)hipFreeAsync(q, st);
  }
};
}  // namespace

int zdt_fftfdr(zdt_ws* ws, const float* rates, int64_t nrates, float* dest) {
  ZDT_TRY(check_ws(ws));
  ProfInstall prof_(ws);
  if (rates == nullptr || dest == nullptr || nrates < 1 || nrates > (1 << 24))
    return fail(ZDT_EINVAL, "fftfdr!: need a host rate list (1..2^24 rates) and a dest");
  ZDT_TRY(set_device(ws));
  if (!ws->have_input) return fail(ZDT_ESTATE, "fftfdr!: no input");
  ZDT_TRY(ensure_F_natural(ws));
  // phasor(k, t, Float32(rate), Nf) = cispi(2 k t r/Nf) (src/fftfdr.jl:10-12)
  std::vector<double> rh(static_cast<size_t>(nrates));
  for (int64_t j = 0; j < nrates; ++j) {
    if (!std::isfinite(rates[j])) return fail(ZDT_EINVAL, "fftfdr!: rate %lld is not finite", (long long)j);
    rh[static_cast<size_t>(j)] = static_cast<double>(rates[j]) / static_cast<double>(ws->Nf);
  }
  AsyncBufs tmp(ws->stream);
  void *rd = nullptr, *Y = nullptr, *s0 = nullptr, *s1 = nullptr, *s2 = nullptr;
  const size_t sc = fft_real_scratch_elems(ws->Nf, nrates) * sizeof(c32);
  ZDT_TRY(tmp.get(&rd, rh.size() * sizeof(double)));
  ZDT_TRY(tmp.get(&Y, static_cast<size_t>(ws->Nk * nrates) * sizeof(c32)));
  ZDT_TRY(tmp.get(&s0, sc));
  ZDT_TRY(tmp.get(&s1, sc));
  ZDT_TRY(tmp.get(&s2, sc));
  ZDT_TRY_HIP(hipMemcpyAsync(rd, rh.data(), rh.size() * sizeof(double), hipMemcpyHostToDevice, ws->stream));
  ZDT_TRY(launch_czt_rates(ws->F, ws->Nk, static_cast<c32*>(Y), ws->Nk, ws->Nk, ws->Nf, static_cast<int>(ws->Nt),
                           static_cast<const double*>(rd), static_cast<int>(nrates), ws->stream));
  ZDT_TRY(fft_c2r(ws->Nf, static_cast<const c32*>(Y), ws->Nk, dest, ws->Nf, nrates, 1.0, s0, s1, s2, ws->stream,
                  ws->maxr));
  return ZDT_OK;
}

int zdt_fdshift(zdt_ws* ws, float rate, float* dest) {
  ZDT_TRY(check_ws(ws));
  ProfInstall prof_(ws);
  if (dest == nullptr || !std::isfinite(rate)) return fail(ZDT_EINVAL, "fdshift!: need a finite rate and a dest");
  ZDT_TRY(set_device(ws));
  if (!ws->have_input) return fail(ZDT_ESTATE, "fdshift!: no input");
  ZDT_TRY(ensure_F_natural(ws));
  AsyncBufs tmp(ws->stream);
  void *Y = nullptr, *s0 = nullptr, *s1 = nullptr, *s2 = nullptr;
  const size_t sc = fft_real_scratch_elems(ws->Nf, ws->Nt) * sizeof(c32);
  ZDT_TRY(tmp.get(&Y, static_cast<size_t>(ws->Nk * ws->Nt) * sizeof(c32)));
  ZDT_TRY(tmp.get(&s0, sc));
  ZDT_TRY(tmp.get(&s1, sc));
  ZDT_TRY(tmp.get(&s2, sc));
  ZDT_TRY(launch_shift_phase(ws->F, ws->Nk, static_cast<c32*>(Y), ws->Nk, ws->Nk, static_cast<int>(ws->Nt),
                             static_cast<double>(rate) / static_cast<double>(ws->Nf), ws->stream));
  ZDT_TRY(fft_c2r(ws->Nf, static_cast<const c32*>(Y), ws->Nk, dest, ws->Nf, ws->Nt, 1.0, s0, s1, s2, ws->stream,
                  ws->maxr));
  return ZDT_OK;
}

int zdt_buffer(zdt_ws* ws, int which, void** ptr, int64_t* rows, int64_t* cols) {
  ZDT_TRY(check_ws(ws));
  ZDT_TRY(set_device(ws));
  void* p = nullptr;
  int64_t c = 0;
  switch (which) {
    case ZDT_BUF_F:
      if (!ws->have_input) return fail(ZDT_ESTATE, "F is not computed before input!");
      ZDT_TRY(ensure_F_natural(ws));
      p = ws->F;
      c = ws->Nt;
      break;
    case ZDT_BUF_Y:
      ZDT_TRY(ensure_stepwise(ws));
      p = ws->Y;
      c = ws->Nl;
      break;
    case ZDT_BUF_V:
      ZDT_TRY(ensure_V(ws));
      p = ws->V;
      c = ws->Nl;
      break;
    default:
      return fail(ZDT_EINVAL, "unknown buffer %d", which);
  }
  if (ptr) *ptr = p;
  if (rows) *rows = ws->Nk;
  if (cols) *cols = c;
  return ZDT_OK;
}

// ---------------------------------------------------------------- profiling

int zdt_profile_reset(zdt_ws* ws) {
  ZDT_TRY(check_ws(ws));
  if (ws->prof == nullptr) return fail(ZDT_ESTATE, "workspace was created without ZDT_FLAG_PROFILE");
  ZDT_TRY(set_device(ws));
  ZDT_TRY_HIP(hipStreamSynchronize(ws->stream));
  ws->prof->reset();
  return ZDT_OK;
}

int zdt_profile_read(zdt_ws* ws, int idx, int* count, char* name, int name_len, double* total_ms,
                     int64_t* launches, double* bytes_per_launch) {
  ZDT_TRY(check_ws(ws));
  Profiler* p = ws->prof;
  if (p == nullptr) return fail(ZDT_ESTATE, "workspace was created without ZDT_FLAG_PROFILE");
  ZDT_TRY(set_device(ws));
  if (p->agg.empty() && !p->recs.empty()) {
    ZDT_TRY_HIP(hipStreamSynchronize(ws->stream));
    for (auto& r : p->recs) {
      if (!r.a || !r.b) continue;
      float ms = 0;
      ZDT_TRY_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      Profiler::Agg* a = nullptr;
      for (auto& x : p->agg)
        if (x.name == r.name) a = &x;
      if (a == nullptr) {
        p->agg.push_back(Profiler::Agg());
        a = &p->agg.back();
        a->name = r.name;
      }
      a->ms += ms;
      a->bytes += r.bytes;
      a->n += 1;
    }
  }
  if (count) *count = static_cast<int>(p->agg.size());
  if (idx < 0) return ZDT_OK;
  if (idx >= static_cast<int>(p->agg.size())) return fail(ZDT_EINVAL, "profile index %d out of range", idx);
  const auto& a = p->agg[idx];
  if (name && name_len > 0) {
    std::strncpy(name, a.name.c_str(), static_cast<size_t>(name_len) - 1);
    name[name_len - 1] = 0;
  }
  if (total_ms) *total_ms = a.ms;
  if (launches) *launches = a.n;
  if (bytes_per_launch) *bytes_per_launch = a.n ? a.bytes / static_cast<double>(a.n) : 0.0;
  return ZDT_OK;
}

// ---------------------------------------------------------------- statistics

static int check_fdrs(const void* const* fdrs, int nb, int64_t Nf, int64_t Nr, int64_t ld) {
  if (fdrs == nullptr || nb < 1) return fail(ZDT_EINVAL, "need at least one FDR matrix");
  if (Nf < 1 || Nr < 1 || ld < Nf) return fail(ZDT_EINVAL, "bad FDR shape Nf=%lld Nr=%lld ld=%lld",
                                               (long long)Nf, (long long)Nr, (long long)ld);
  for (int b = 0; b < nb; ++b)
    if (fdrs[b] == nullptr) return fail(ZDT_EINVAL, "FDR matrix %d is NULL", b);
  return ZDT_OK;
}

int zdt_fdrstats(const float* const* fdrs, int nb, int64_t Nf, int64_t Nr, int64_t ld, void* stream, double* m,
                 double* s) {
  ZDT_TRY(check_fdrs(reinterpret_cast<const void* const*>(fdrs), nb, Nf, Nr, ld));
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int nchunk = colstats_nchunk(Nf);
  const size_t part_n = static_cast<size_t>(2) * nchunk * Nr;
  const size_t per = part_n + 2 * static_cast<size_t>(Nr);
  double* buf = nullptr;
  ZDT_TRY_HIP(hipMallocAsync(reinterpret_cast<void**>(&buf), per * nb * sizeof(double), st));
  int rc = ZDT_OK;
  for (int b = 0; b < nb && rc == ZDT_OK; ++b) {
    double* base = buf + per * b;
    rc = launch_colstats2(fdrs[b], Nf, Nr, ld, base, base + part_n, base + part_n + Nr, st);
  }
  std::vector<double> host(per * nb);
  if (rc == ZDT_OK) {
    hipError_t e = hipMemcpyAsync(host.data(), buf, per * nb * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = hip_fail(e, "fdrstats copy");
  }
  (void)hipFreeAsync(buf, st);
  if (rc != ZDT_OK) return rc;
  const double mean0 = host[part_n];
  double vmin = INFINITY;
  bool nan = false;
  for (int b = 0; b < nb; ++b) {
    const double* var = host.data() + per * b + part_n + Nr;
    for (int64_t j = 0; j < Nr; ++j) {
      if (std::isnan(var[j])) nan = true;
      vmin = std::min(vmin, var[j]);
    }
  }
  if (m) *m = mean0;
  if (s) *s = nan ? NAN : std::sqrt(std::max(vmin, 0.0));
  return ZDT_OK;
}

int zdt_fdrnormalize(float* const* fdrs, int nb, int64_t Nf, int64_t Nr, int64_t ld, double m, double s,
                     void* stream) {
  ZDT_TRY(check_fdrs(reinterpret_cast<const void* const*>(fdrs), nb, Nf, Nr, ld));
  for (int b = 0; b < nb; ++b)
    ZDT_TRY(launch_normalize(fdrs[b], Nf, Nr, ld, m, s, static_cast<hipStream_t>(stream)));
  return ZDT_OK;
}

int zdt_findprotohits(const float* const* fdrs, int nb, int64_t Nf, int64_t Nr, int64_t ld, double threshold,
                      int64_t* hits, int64_t cap, int64_t* nhits, void* stream) {
  ZDT_TRY(check_fdrs(reinterpret_cast<const void* const*>(fdrs), nb, Nf, Nr, ld));
  if (cap < 0 || (cap > 0 && hits == nullptr)) return fail(ZDT_EINVAL, "findprotohits: bad output buffer");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t nt = hits_tiles(Nf, Nr);
  const int64_t ntall = nt * nb;
  int64_t* buf = nullptr;
  ZDT_TRY_HIP(hipMallocAsync(reinterpret_cast<void**>(&buf), (ntall + 1 + hits_scan_scratch(ntall)) * sizeof(int64_t),
                              st));
  int64_t* counts = buf;
  int64_t* total_dev = buf + ntall;
  int rc = ZDT_OK;
  for (int b = 0; b < nb && rc == ZDT_OK; ++b)
    rc = launch_hits_count(fdrs[b], Nf, Nr, ld, threshold, nullptr, counts + b * nt, st);
  if (rc == ZDT_OK) rc = launch_hits_scan(counts, ntall, total_dev, total_dev + 1, st);
  int64_t total = 0;
  if (rc == ZDT_OK) {
    hipError_t e = hipMemcpyAsync(&total, total_dev, sizeof(int64_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = hip_fail(e, "findprotohits count copy");
  }
  if (rc == ZDT_OK && cap > 0 && total > 0) {
    for (int b = 0; b < nb && rc == ZDT_OK; ++b)
      rc = launch_hits_write_impl(fdrs[b], Nf, Nr, ld, threshold, nullptr, counts, ntall, b * nt, total_dev,
                                  static_cast<int64_t>(b) * Nf * Nr, hits, cap, st);
  }
  (void)hipFreeAsync(buf, st);
  if (rc != ZDT_OK) return rc;
  if (nhits) *nhits = total;
  if (total > cap) return fail(ZDT_ETRUNC, "findprotohits: %lld hits exceed capacity %lld", (long long)total,
                               (long long)cap);
  return ZDT_OK;
}

int zdt_clusterprotohits(const int64_t* hits, int64_t n, const float* const* fdrs, int nb, int64_t Nf,
                         int64_t Nr, int64_t ld, int di, int dj, int64_t* peak, float* value, int64_t* size,
                         int64_t* box, int64_t cap, int64_t* nclusters, int64_t* label, void* stream) {
  ZDT_TRY(check_fdrs(reinterpret_cast<const void* const*>(fdrs), nb, Nf, Nr, ld));
  if (n < 0 || (n > 0 && hits == nullptr)) return fail(ZDT_EINVAL, "clusterprotohits: bad proto-hit list");
  if (n > INT32_MAX) return fail(ZDT_ENOTSUP, "clusterprotohits: %lld proto-hits exceed 2^31-1", (long long)n);
  if (Nf > INT32_MAX || static_cast<int64_t>(nb) * Nr > INT32_MAX)
    return fail(ZDT_ENOTSUP, "clusterprotohits: channels or columns exceed 2^31-1");
  if (di < 0 || dj < 0) return fail(ZDT_EINVAL, "clusterprotohits: negative neighbourhood (%d, %d)", di, dj);
  if (nclusters == nullptr) return fail(ZDT_EINVAL, "clusterprotohits: nclusters is NULL");
  if (cap < 0 || (cap > 0 && (peak == nullptr || value == nullptr || size == nullptr || box == nullptr)))
    return fail(ZDT_EINVAL, "clusterprotohits: bad output buffers");
  return cluster_protohits_impl(hits, n, fdrs, nb, Nf, Nr, ld, di, dj, peak, value, size, box, cap, nclusters,
                                label, static_cast<hipStream_t>(stream));
}

}  // extern "C"
