// Ragged multi-channel plans: many spectrograms of uneven Nf and Nt through
// input!, one zdtfdr! batch, fdrdenormalize and findprotohits, captured once
// into a hipGraph and replayed with a single launch.
//
// The reference leaves this loop to its user (one ZDTWorkspace per shape,
// zdtfdr! per channel, src/zdtfdr.jl:349-421, then fdrstats/findprotohits,
// src/fdrutils.jl:39-169).  Here channels of one shape share a workspace and
// run in order on that shape's stream; the shapes' streams fork from and
// join into the plan's origin stream, so the captured graph has one branch
// per shape.  The detection stage stays on the device: column statistics,
// the threshold snr*s + m, a tile count, a scan and an ordered write.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "zdt_fast.h"
#include "zdt_internal.h"
#include "zdt_kernels.h"
#include "zdt_ws.h"

namespace zdt {
namespace {

struct Shape {
  int64_t Nf = 0, Nt = 0, Nr = 0;
  float r0 = 0, dr = 0;
  std::vector<int> channels;
  zdt_ws* ws = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t done = nullptr;
  int fused = 0, n2 = 0, maxr = 16;
  int tune_src = 0;      // 0: defaults, 1: timed at plan creation, 2: from the tuning table
  float tune_ms = 0.0f;  // the chosen candidate's time (input! + one batch), when known
  // detection scratch (reused by the shape's channels, which run in order)
  double* stats = nullptr;  // partial sums, colmean, colvar, threshold
  int64_t* counts = nullptr;  // tile counts / offsets
  float* fdr = nullptr;       // FDR for channels bound without a dest
  float* smax = nullptr;      // the FDR's run maxima (fused paths): the hit count reads only tiles that can hold one
};

}  // namespace
}  // namespace zdt

struct zdt_plan {
  int device = 0;
  int flags = 0;
  hipStream_t user = nullptr;    // caller's stream
  hipStream_t origin = nullptr;  // capture / launch stream
  hipEvent_t fork = nullptr, join = nullptr;
  std::vector<zdt::Shape> shapes;
  std::vector<int> ch_shape;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace zdt {
namespace {

int hip_alloc(void** p, size_t bytes) {
  *p = nullptr;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(ZDT_ENOMEM, "plan: hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
  }
  return ZDT_OK;
}

size_t stats_doubles(const Shape& s) {
  return static_cast<size_t>(devstats_scratch_doubles(s.Nf, s.Nr)) + 2 * static_cast<size_t>(s.Nr) + 1;
}

// Times input! + one zdtfdr! batch of a candidate workspace (ms, best of 3).
int time_candidate(const Shape& s, int flags, int n2, int maxr, const float* spec, float* dest, float* ms_out) {
  zdt_ws* ws = nullptr;
  ZDT_TRY(create_ws_tuned(nullptr, s.Nf, s.Nt, s.r0, s.dr, s.Nr, nullptr, 0, flags, ZDT_PRECISION_FP32, s.st, n2, maxr,
                          &ws));
  if (!(flags & (ZDT_FLAG_GENERIC | ZDT_FLAG_NO_FUSED_INTERNAL)) && n2 != 0 && ws_path(ws) != 1) {
    zdt_destroy(ws);
    return fail(ZDT_EHIP, "plan: fused candidate did not get the fused path");
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = ZDT_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = fail(ZDT_EHIP, "plan: event");
  float best = INFINITY;
  for (int it = 0; it < 4 && rc == ZDT_OK; ++it) {
    if (hipEventRecord(e0, s.st) != hipSuccess) rc = fail(ZDT_EHIP, "plan: event record");
    if (rc == ZDT_OK) rc = zdt_input(ws, spec);
    if (rc == ZDT_OK) rc = zdt_run(ws, s.r0, ZDT_WINDOW_RECT, nullptr, dest);
    if (rc == ZDT_OK && hipEventRecord(e1, s.st) != hipSuccess) rc = fail(ZDT_EHIP, "plan: event record");
    if (rc == ZDT_OK && hipEventSynchronize(e1) != hipSuccess) rc = fail(ZDT_EHIP, "plan: event sync");
    float ms = 0;
    if (rc == ZDT_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = fail(ZDT_EHIP, "plan: elapsed");
    if (rc == ZDT_OK && it > 0) best = std::min(best, ms);  // iteration 0 warms up
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  zdt_destroy(ws);
  *ms_out = best;
  return rc;
}

// ---------------------------------------------------------------- tuning table
//
// zdt_tune_table names a text file that keeps the autotuner's choices across
// processes: one line per (device, shape),
//   zdt-tune-v1 <gcnArchName> <CUs> <Nf> <Nt> <Nr> <forced generic 0/1> <path> <n2> <maxr> <ms>
// Plans look their shapes up before timing anything and append what they
// time.  Lines that do not parse are ignored; the last entry for a key wins.

std::mutex g_tune_mu;
std::string g_tune_path;  // empty: no table

struct TuneKey {
  std::string arch;
  int cus = 0;
  int64_t Nf = 0, Nt = 0, Nr = 0;
  int generic = 0;
  bool operator<(const TuneKey& o) const {
    return std::tie(arch, cus, Nf, Nt, Nr, generic) < std::tie(o.arch, o.cus, o.Nf, o.Nt, o.Nr, o.generic);
  }
};
struct TuneVal {
  int path = 0, n2 = 0, maxr = 16;
  float ms = 0.0f;
};

int device_key(TuneKey& k) {
  int dev = 0;
  ZDT_TRY_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  ZDT_TRY_HIP(hipGetDeviceProperties(&prop, dev));
  k.arch = prop.gcnArchName;
  for (char& ch : k.arch)
    if (ch == ' ') ch = '_';
  k.cus = prop.multiProcessorCount;
  return ZDT_OK;
}

bool tune_lookup(const TuneKey& key, TuneVal& out) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  if (g_tune_path.empty()) return false;
  FILE* f = std::fopen(g_tune_path.c_str(), "r");
  if (f == nullptr) return false;
  char line[512];
  bool found = false;
  while (std::fgets(line, sizeof line, f)) {
    char tag[32], arch[128];
    int cus, gen, path, n2, maxr;
    long long Nf, Nt, Nr;
    float ms;
    if (std::sscanf(line, "%31s %127s %d %lld %lld %lld %d %d %d %d %f", tag, arch, &cus, &Nf, &Nt, &Nr, &gen, &path,
                    &n2, &maxr, &ms) != 11)
      continue;
    if (std::strcmp(tag, "zdt-tune-v1") != 0 || key.arch != arch || key.cus != cus || key.Nf != Nf ||
        key.Nt != Nt || key.Nr != Nr || key.generic != gen)
      continue;
    if (path < 0 || path > 2 || (path == 1 && n2 != 512 && n2 != 1024) || (path == 0 && maxr != 8 && maxr != 16))
      continue;
    out = {path, n2, maxr, ms};
    found = true;
  }
  std::fclose(f);
  return found;
}

void tune_store(const TuneKey& key, const TuneVal& v) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  if (g_tune_path.empty()) return;
  FILE* f = std::fopen(g_tune_path.c_str(), "a");
  if (f == nullptr) return;  // a table that cannot be written only costs the next plan a retune
  std::fprintf(f, "zdt-tune-v1 %s %d %lld %lld %lld %d %d %d %d %.6g\n", key.arch.c_str(), key.cus, (long long)key.Nf,
               (long long)key.Nt, (long long)key.Nr, key.generic, v.path, v.n2, v.maxr, v.ms);
  std::fclose(f);
}

// workspace flags that select a candidate's path
int path_flags(int wflags, int path) {
  if (path == 0) return wflags | ZDT_FLAG_GENERIC;
  if (path == 2) return wflags | ZDT_FLAG_NO_FUSED_INTERNAL;
  return wflags;
}

// Chooses the kernel path and its shape parameters for a shape by timing the
// candidates (or from the tuning table): the fused kernels with strided FFTs
// of 1024 or 512 points (k1/k3 tiles of 32 or 16 rows, 138 or 70 KB of LDS),
// the fused output, and the generic kernels with radices up to 16 or 8.
int autotune(Shape& s, int flags) {
  struct Cand {
    int path, n2, maxr;
  };
  const bool generic_only = (flags & ZDT_FLAG_GENERIC) != 0;
  std::vector<Cand> cands;
  if (!generic_only && fast_supported(s.Nf, s.Nt, s.Nr))
    for (int n2 : {1024, 512})
      if (fast_n2_ok(s.Nf, n2)) cands.push_back({1, n2, 16});
  if (!generic_only && fast_out_supported(s.Nf, s.Nt, s.Nr)) cands.push_back({2, 0, 16});
  for (int maxr : {16, 8}) cands.push_back({0, 0, maxr});
  TuneKey key;
  ZDT_TRY(device_key(key));
  key.Nf = s.Nf;
  key.Nt = s.Nt;
  key.Nr = s.Nr;
  key.generic = generic_only ? 1 : 0;
  TuneVal hit;
  if (tune_lookup(key, hit)) {
    s.fused = hit.path;
    s.n2 = hit.n2;
    s.maxr = hit.maxr;
    s.tune_src = 2;
    s.tune_ms = hit.ms;
    return ZDT_OK;
  }
  void *spec = nullptr, *dest = nullptr;
  int rc = hip_alloc(&spec, static_cast<size_t>(s.Nf * s.Nt) * sizeof(float));
  if (rc == ZDT_OK) rc = hip_alloc(&dest, static_cast<size_t>(s.Nf * s.Nr) * sizeof(float));
  // the timing does not depend on the values; fill with a constant
  if (rc == ZDT_OK && hipMemsetAsync(spec, 0, static_cast<size_t>(s.Nf * s.Nt) * sizeof(float), s.st) != hipSuccess)
    rc = fail(ZDT_EHIP, "plan: memset");
  float best = INFINITY;
  for (const Cand& c : cands) {
    if (rc != ZDT_OK) break;
    float ms = INFINITY;
    rc = time_candidate(s, path_flags(flags, c.path), c.n2, c.maxr, static_cast<float*>(spec),
                        static_cast<float*>(dest), &ms);
    if (rc == ZDT_OK && ms < best) {
      best = ms;
      s.fused = c.path;
      s.n2 = c.n2;
      s.maxr = c.maxr;
    }
  }
  (void)hipStreamSynchronize(s.st);
  if (spec) (void)hipFree(spec);
  if (dest) (void)hipFree(dest);
  if (rc == ZDT_OK) {
    s.tune_src = 1;
    s.tune_ms = best;
    tune_store(key, {s.fused, s.n2, s.maxr, best});
  }
  return rc;
}

void destroy_graph(zdt_plan* p) {
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->exec = nullptr;
  p->graph = nullptr;
}

// Enqueues channel c (input!, zdtfdr!, stats, threshold, hits) on its shape's stream.
int enqueue_channel(Shape& s, int c, const float* spec, float* dest, double thr, int snr, int64_t* hits, int64_t cap,
                    int64_t* nhits, double* thr_out) {
  ZDT_TRY(zdt_input(s.ws, spec));
  const int64_t ld = s.Nf;
  const double* thr_dev = nullptr;
  double thr_h = thr;
  if (snr) {
    // the transform and the FDR statistics (fused into the output kernel
    // where the workspace runs the fused kernels)
    const size_t part_n = static_cast<size_t>(devstats_scratch_doubles(s.Nf, s.Nr));
    double* colmean = s.stats + part_n;
    double* colvar = colmean + s.Nr;
    double* t = thr_out ? thr_out + c : colvar + s.Nr;
    ZDT_TRY(run_devstats(s.ws, s.r0, dest, colmean, colvar, s.stats, s.smax));
    ZDT_TRY(launch_hits_threshold(colmean, colvar, s.Nr, thr, t, s.st));
    thr_dev = t;
  } else {
    ZDT_TRY(zdt_run(s.ws, s.r0, ZDT_WINDOW_RECT, nullptr, dest));
    if (thr_out) ZDT_TRY(launch_set_double(thr_out + c, thr_h, s.st));
  }
  const int64_t nt = hits_tiles(s.Nf, s.Nr);
  // the run maxima exist when the statistics pass wrote them
  const float* smax = snr ? s.smax : nullptr;
  ZDT_TRY(launch_hits_count(dest, s.Nf, s.Nr, ld, thr_h, thr_dev, s.counts, s.st, smax, ws_fused_n2(s.ws)));
  ZDT_TRY(launch_hits_scan(s.counts, nt, nhits + c, s.counts + nt, s.st));
  if (cap > 0)
    ZDT_TRY(launch_hits_write_impl(dest, s.Nf, s.Nr, ld, thr_h, thr_dev, s.counts, nt, 0, nhits + c, 0,
                                   hits + static_cast<int64_t>(c) * cap, cap, s.st));
  return ZDT_OK;
}

}  // namespace
}  // namespace zdt

using namespace zdt;

extern "C" {

int zdt_plan_destroy(zdt_plan* p) {
  if (p == nullptr) return ZDT_OK;
  (void)hipSetDevice(p->device);
  if (p->origin) (void)hipStreamSynchronize(p->origin);
  destroy_graph(p);
  for (Shape& s : p->shapes) {
    if (s.st) (void)hipStreamSynchronize(s.st);
    if (s.ws) zdt_destroy(s.ws);
    if (s.stats) (void)hipFree(s.stats);
    if (s.counts) (void)hipFree(s.counts);
    if (s.smax) (void)hipFree(s.smax);
    if (s.fdr) (void)hipFree(s.fdr);
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.st) (void)hipStreamDestroy(s.st);
  }
  if (p->fork) (void)hipEventDestroy(p->fork);
  if (p->join) (void)hipEventDestroy(p->join);
  if (p->origin) (void)hipStreamDestroy(p->origin);
  delete p;
  return ZDT_OK;
}

int zdt_plan_ragged(int nch, const int64_t* Nf, const int64_t* Nt, const float* r0, const float* dr,
                    const int64_t* Nr, int flags, void* stream, zdt_plan** out) {
  if (out == nullptr) return fail(ZDT_EINVAL, "plan: out is NULL");
  *out = nullptr;
  if (nch < 1 || !Nf || !Nt || !r0 || !dr || !Nr) return fail(ZDT_EINVAL, "plan: need nch >= 1 and shape arrays");
  if (flags & ~(ZDT_FLAG_UNNORMALIZED | ZDT_FLAG_GENERIC | ZDT_PLAN_AUTOTUNE))
    return fail(ZDT_EINVAL, "plan: unsupported flags 0x%x", flags);
  zdt_plan* p = new zdt_plan();
  p->flags = flags;
  p->user = static_cast<hipStream_t>(stream);
  int rc = ZDT_OK;
  if (hipGetDevice(&p->device) != hipSuccess) rc = fail(ZDT_EHIP, "plan: hipGetDevice");
  if (rc == ZDT_OK && (hipStreamCreateWithFlags(&p->origin, hipStreamNonBlocking) != hipSuccess ||
                       hipEventCreateWithFlags(&p->fork, hipEventDisableTiming) != hipSuccess ||
                       hipEventCreateWithFlags(&p->join, hipEventDisableTiming) != hipSuccess))
    rc = fail(ZDT_EHIP, "plan: stream/event creation failed");
  std::map<std::tuple<int64_t, int64_t, int64_t, float, float>, int> index;
  for (int c = 0; c < nch && rc == ZDT_OK; ++c) {
    if (Nf[c] < 1 || Nt[c] < 1 || Nr[c] < 1) {
      rc = fail(ZDT_EINVAL, "plan: channel %d has shape Nf=%lld Nt=%lld Nr=%lld", c, (long long)Nf[c],
                (long long)Nt[c], (long long)Nr[c]);
      break;
    }
    auto key = std::make_tuple(Nf[c], Nt[c], Nr[c], r0[c], dr[c]);
    auto it = index.find(key);
    int si;
    if (it == index.end()) {
      si = static_cast<int>(p->shapes.size());
      index[key] = si;
      Shape s;
      s.Nf = Nf[c];
      s.Nt = Nt[c];
      s.Nr = Nr[c];
      s.r0 = r0[c];
      s.dr = dr[c];
      p->shapes.push_back(s);
    } else {
      si = it->second;
    }
    p->shapes[si].channels.push_back(c);
    p->ch_shape.push_back(si);
  }
  const int wflags = flags & (ZDT_FLAG_UNNORMALIZED | ZDT_FLAG_GENERIC);
  for (Shape& s : p->shapes) {
    if (rc != ZDT_OK) break;
    if (hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) {
      rc = fail(ZDT_EHIP, "plan: stream/event creation failed");
      break;
    }
    s.fused = (!(wflags & ZDT_FLAG_GENERIC) && fast_supported(s.Nf, s.Nt, s.Nr)) ? 1 : 0;
    int cflags = wflags;
    if (flags & ZDT_PLAN_AUTOTUNE) {
      rc = autotune(s, wflags);
      if (rc != ZDT_OK) break;
      cflags = path_flags(wflags, s.fused);
    }
    rc = create_ws_tuned(nullptr, s.Nf, s.Nt, s.r0, s.dr, s.Nr, nullptr, 0, cflags, ZDT_PRECISION_FP32, s.st,
                         s.fused == 1 ? s.n2 : 0, s.maxr, &s.ws);
    if (rc != ZDT_OK) break;
    s.fused = ws_path(s.ws);
    rc = hip_alloc(reinterpret_cast<void**>(&s.stats), stats_doubles(s) * sizeof(double));
    const int64_t nt = hits_tiles(s.Nf, s.Nr);
    if (rc == ZDT_OK)
      rc = hip_alloc(reinterpret_cast<void**>(&s.counts), (nt + hits_scan_scratch(nt)) * sizeof(int64_t));
    const int n2 = ws_fused_n2(s.ws);
    if (rc == ZDT_OK && n2 > 0 && s.Nf % (16 * FDR_TILE) == 0 && s.Nf / (2 * n2) <= 2048)
      rc = hip_alloc(reinterpret_cast<void**>(&s.smax), (s.Nf / FDR_RUN) * s.Nr * sizeof(float));
  }
  if (rc != ZDT_OK) {
    std::string msg = zdt_last_error();
    zdt_plan_destroy(p);
    fail(rc, "%s", msg.c_str());
    return rc;
  }
  *out = p;
  return ZDT_OK;
}

int zdt_plan_info(const zdt_plan* p, int ch, int* shape, int* path, int* n2, int* maxr) {
  if (p == nullptr) return fail(ZDT_EINVAL, "plan is NULL");
  if (ch < 0 || ch >= static_cast<int>(p->ch_shape.size())) return fail(ZDT_EINVAL, "plan: channel %d out of range", ch);
  const int si = p->ch_shape[ch];
  const Shape& s = p->shapes[si];
  if (shape) *shape = si;
  if (path) *path = s.fused;
  const int n2_default = s.Nf >= (int64_t(1) << 17) ? 1024 : 512;
  if (n2) *n2 = s.fused == 1 ? (s.n2 ? s.n2 : n2_default) : (s.fused == 2 ? n2_default : 0);
  if (maxr) *maxr = s.fused != 0 ? 0 : s.maxr;
  return ZDT_OK;
}

int zdt_plan_tuning(const zdt_plan* p, int ch, int* source, float* ms) {
  if (p == nullptr) return fail(ZDT_EINVAL, "plan is NULL");
  if (ch < 0 || ch >= static_cast<int>(p->ch_shape.size())) return fail(ZDT_EINVAL, "plan: channel %d out of range", ch);
  const Shape& s = p->shapes[p->ch_shape[ch]];
  if (source) *source = s.tune_src;
  if (ms) *ms = s.tune_ms;
  return ZDT_OK;
}

int zdt_tune_table(const char* path) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tune_path = path ? path : "";
  return ZDT_OK;
}

int zdt_plan_bind(zdt_plan* p, const float* const* specs, float* const* dests, const double* thr, int snr,
                  int64_t* hits, int64_t cap, int64_t* nhits, double* thr_out) {
  if (p == nullptr) return fail(ZDT_EINVAL, "plan is NULL");
  const int nch = static_cast<int>(p->ch_shape.size());
  if (specs == nullptr || thr == nullptr || nhits == nullptr || cap < 0 || (cap > 0 && hits == nullptr))
    return fail(ZDT_EINVAL, "plan_bind: specs, thr and nhits are required (hits when cap > 0)");
  for (int c = 0; c < nch; ++c)
    if (specs[c] == nullptr) return fail(ZDT_EINVAL, "plan_bind: specs[%d] is NULL", c);
  ZDT_TRY_HIP(hipSetDevice(p->device));
  // plan-owned FDR buffers for channels without a dest
  for (int c = 0; c < nch; ++c) {
    Shape& s = p->shapes[p->ch_shape[c]];
    if ((dests == nullptr || dests[c] == nullptr) && s.fdr == nullptr)
      ZDT_TRY(hip_alloc(reinterpret_cast<void**>(&s.fdr), static_cast<size_t>(s.Nf * s.Nr) * sizeof(float)));
  }
  ZDT_TRY_HIP(hipStreamSynchronize(p->origin));
  destroy_graph(p);
  ZDT_TRY_HIP(hipStreamBeginCapture(p->origin, hipStreamCaptureModeRelaxed));
  int rc = ZDT_OK;
  if (hipEventRecord(p->fork, p->origin) != hipSuccess) rc = fail(ZDT_EHIP, "plan: fork record");
  for (Shape& s : p->shapes) {
    if (rc != ZDT_OK) break;
    if (hipStreamWaitEvent(s.st, p->fork, 0) != hipSuccess) {
      rc = fail(ZDT_EHIP, "plan: fork wait");
      break;
    }
    for (int c : s.channels) {
      float* dest = (dests && dests[c]) ? dests[c] : s.fdr;
      rc = enqueue_channel(s, c, specs[c], dest, thr[c], snr, hits, cap, nhits, thr_out);
      if (rc != ZDT_OK) break;
    }
    if (rc == ZDT_OK && hipEventRecord(s.done, s.st) != hipSuccess) rc = fail(ZDT_EHIP, "plan: join record");
    if (rc == ZDT_OK && hipStreamWaitEvent(p->origin, s.done, 0) != hipSuccess) rc = fail(ZDT_EHIP, "plan: join wait");
  }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(p->origin, &g);
  if (rc == ZDT_OK && e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture");
  if (rc != ZDT_OK) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return rc;
  }
  p->graph = g;
  e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    destroy_graph(p);
    return hip_fail(e, "hipGraphInstantiate");
  }
  return ZDT_OK;
}

int zdt_plan_launch(zdt_plan* p) {
  if (p == nullptr) return fail(ZDT_EINVAL, "plan is NULL");
  if (p->exec == nullptr) return fail(ZDT_ESTATE, "plan_launch: bind the plan first");
  ZDT_TRY_HIP(hipSetDevice(p->device));
  // origin follows the caller's stream, and the caller's stream follows the graph
  ZDT_TRY_HIP(hipEventRecord(p->join, p->user));
  ZDT_TRY_HIP(hipStreamWaitEvent(p->origin, p->join, 0));
  ZDT_TRY_HIP(hipGraphLaunch(p->exec, p->origin));
  ZDT_TRY_HIP(hipEventRecord(p->join, p->origin));
  ZDT_TRY_HIP(hipStreamWaitEvent(p->user, p->join, 0));
  return ZDT_OK;
}

}  // extern "C"
