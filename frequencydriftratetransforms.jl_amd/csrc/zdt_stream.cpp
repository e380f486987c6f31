// Long-integration search: spectrograms of one shape streamed from host
// memory, the rate range cut into equal batches sized to an HBM budget, and
// proto-hits at an SNR threshold taken over all batches of a spectrogram.
//
// A reference user writes this loop by hand: batchrates to split the rates
// (src/batchrates.jl:22-60), estimate_memory to size the batches
// (src/zdtutils.jl:35-53), zdtfdr!(dests, ws, spec; r0 = batch starts)
// (src/zdtfdr.jl:349-362) and findprotohits(fdrs, snr; snr=true) over the
// batches (src/fdrutils.jl:148-169).  Here the copy of spectrogram i+1 to the
// device runs on its own stream while spectrogram i is searched (double
// buffered, ordered by events), and the batches' FDRs stay in HBM when the
// budget allows; otherwise a first pass over the batches gathers the
// statistics (fdrstats decomposes: mean of the first column, minimum of the
// per-column deviations) and a second pass recomputes each batch for its
// hits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "zdt_internal.h"
#include "zdt_ws.h"

struct zdt_stream {
  int device = 0;
  int64_t Nf = 0, Nt = 0, Nr = 0, Nrb = 0, nb = 0;
  float r0 = 0, dr = 0;
  bool resident = false;
  int64_t bytes = 0;
  zdt_ws* ws = nullptr;
  hipStream_t compute = nullptr, copy = nullptr;
  hipEvent_t copied[2] = {nullptr, nullptr}, used[2] = {nullptr, nullptr};
  float* spec[2] = {nullptr, nullptr};
  std::vector<float*> fdr;  // resident batch FDRs (or one reused buffer)
  int64_t* dhits = nullptr;  // device hit staging
  int64_t dcap = 0;
  double* dstats = nullptr;  // colmean, colvar (Nrb each), then generic-path scratch
};

namespace zdt {
namespace {

int salloc(zdt_stream* s, void** p, size_t bytes) {
  *p = nullptr;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(ZDT_ENOMEM, "stream: hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
  }
  s->bytes += static_cast<int64_t>(bytes);
  return ZDT_OK;
}

// device bytes of a search with batches of Nrb rates
int64_t stream_bytes(int64_t Nf, int64_t Nt, int64_t Nr, int64_t Nrb, bool resident, int flags, int64_t dcap) {
  const int64_t ws = zdt_workspace_bytes(Nf, Nt, Nrb, nullptr, 0, flags);
  if (ws < 0) return -1;
  const int64_t nb = (Nr + Nrb - 1) / Nrb;
  const int64_t fdrs = (resident ? nb : 1) * Nf * Nrb * 4;
  return ws + fdrs + 2 * Nf * Nt * 4 + dcap * 8 + (2 * Nrb + devstats_scratch_doubles(Nf, Nrb)) * 8;
}

}  // namespace
}  // namespace zdt

using namespace zdt;

extern "C" {

int zdt_host_alloc(size_t bytes, void** p) {
  if (p == nullptr) return fail(ZDT_EINVAL, "host_alloc: p is NULL");
  *p = nullptr;
  ZDT_TRY_HIP(hipHostMalloc(p, bytes == 0 ? 1 : bytes, hipHostMallocDefault));
  return ZDT_OK;
}

int zdt_host_free(void* p) {
  if (p) ZDT_TRY_HIP(hipHostFree(p));
  return ZDT_OK;
}

int zdt_stream_destroy(zdt_stream* s) {
  if (s == nullptr) return ZDT_OK;
  (void)hipSetDevice(s->device);
  if (s->compute) (void)hipStreamSynchronize(s->compute);
  if (s->copy) (void)hipStreamSynchronize(s->copy);
  if (s->ws) zdt_destroy(s->ws);
  for (float* f : s->fdr)
    if (f) (void)hipFree(f);
  for (int i = 0; i < 2; ++i) {
    if (s->spec[i]) (void)hipFree(s->spec[i]);
    if (s->copied[i]) (void)hipEventDestroy(s->copied[i]);
    if (s->used[i]) (void)hipEventDestroy(s->used[i]);
  }
  if (s->dhits) (void)hipFree(s->dhits);
  if (s->dstats) (void)hipFree(s->dstats);
  if (s->copy) (void)hipStreamDestroy(s->copy);
  if (s->compute) (void)hipStreamDestroy(s->compute);
  delete s;
  return ZDT_OK;
}

int zdt_stream_create(int64_t Nf, int64_t Nt, float r0, float dr, int64_t Nr, int64_t Nrb, int64_t hbm_bytes,
                      int flags, zdt_stream** out) {
  if (out == nullptr) return fail(ZDT_EINVAL, "stream: out is NULL");
  *out = nullptr;
  if (Nf < 2 || Nt < 1 || Nr < 1 || Nrb < 0 || hbm_bytes < 0)
    return fail(ZDT_EINVAL, "stream: bad shape Nf=%lld Nt=%lld Nr=%lld Nrb=%lld", (long long)Nf, (long long)Nt,
                (long long)Nr, (long long)Nrb);
  if (flags & ~(ZDT_FLAG_UNNORMALIZED | ZDT_FLAG_GENERIC)) return fail(ZDT_EINVAL, "stream: unsupported flags");
  if (!std::isfinite(r0) || !std::isfinite(dr)) return fail(ZDT_EINVAL, "stream: r0 and dr must be finite");
  const int64_t dcap = int64_t(1) << 20;  // device hit staging per batch pass
  int64_t budget = hbm_bytes;
  if (budget == 0) {
    size_t fr = 0, tot = 0;
    ZDT_TRY_HIP(hipMemGetInfo(&fr, &tot));
    budget = static_cast<int64_t>(fr) * 9 / 10;
  }
  // batch length: the largest that keeps every batch FDR resident, else the
  // largest that fits with one FDR buffer reused by all batches
  int64_t best = 0;
  bool resident = false;
  if (Nrb > 0) {
    best = std::min(Nrb, Nr);
    const int64_t b = stream_bytes(Nf, Nt, Nr, best, true, flags, dcap);
    resident = b >= 0 && b <= budget;
    const int64_t b1 = stream_bytes(Nf, Nt, Nr, best, false, flags, dcap);
    if (!resident && (b1 < 0 || b1 > budget))
      return fail(ZDT_ENOMEM, "stream: Nrb=%lld needs %lld bytes, budget %lld", (long long)best, (long long)b1,
                  (long long)budget);
  } else {
    for (int mode = 1; mode >= 0 && best == 0; --mode) {
      for (int64_t nrb = Nr; nrb >= 1; nrb = (nrb > 64 ? nrb - (nrb + 63) / 64 : nrb - 1)) {
        const int64_t b = stream_bytes(Nf, Nt, Nr, nrb, mode == 1, flags, dcap);
        if (b >= 0 && b <= budget) {
          best = nrb;
          resident = (mode == 1);
          break;
        }
      }
    }
    if (best == 0) return fail(ZDT_ENOMEM, "stream: no batch length fits %lld bytes", (long long)budget);
  }
  zdt_stream* s = new zdt_stream();
  s->Nf = Nf;
  s->Nt = Nt;
  s->Nr = Nr;
  s->Nrb = best;
  s->nb = (Nr + best - 1) / best;
  s->r0 = r0;
  s->dr = dr;
  s->resident = resident;
  s->dcap = dcap;
  int rc = ZDT_OK;
  if (hipGetDevice(&s->device) != hipSuccess) rc = fail(ZDT_EHIP, "stream: hipGetDevice");
  if (rc == ZDT_OK && (hipStreamCreateWithFlags(&s->compute, hipStreamNonBlocking) != hipSuccess ||
                       hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking) != hipSuccess))
    rc = fail(ZDT_EHIP, "stream: stream creation failed");
  for (int i = 0; i < 2 && rc == ZDT_OK; ++i) {
    if (hipEventCreateWithFlags(&s->copied[i], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s->used[i], hipEventDisableTiming) != hipSuccess)
      rc = fail(ZDT_EHIP, "stream: event creation failed");
    if (rc == ZDT_OK) rc = salloc(s, reinterpret_cast<void**>(&s->spec[i]), static_cast<size_t>(Nf * Nt) * 4);
  }
  if (rc == ZDT_OK) rc = zdt_create(nullptr, Nf, Nt, r0, dr, s->Nrb, nullptr, 0, flags, s->compute, &s->ws);
  if (rc == ZDT_OK) {
    int64_t wb = 0;
    (void)zdt_sizeof(s->ws, &wb);
    s->bytes += wb;
  }
  const int64_t nbuf = resident ? s->nb : 1;
  for (int64_t b = 0; b < nbuf && rc == ZDT_OK; ++b) {
    float* f = nullptr;
    rc = salloc(s, reinterpret_cast<void**>(&f), static_cast<size_t>(Nf * s->Nrb) * 4);
    if (rc == ZDT_OK) s->fdr.push_back(f);
  }
  if (rc == ZDT_OK) rc = salloc(s, reinterpret_cast<void**>(&s->dhits), static_cast<size_t>(dcap) * 8);
  if (rc == ZDT_OK)
    rc = salloc(s, reinterpret_cast<void**>(&s->dstats),
                static_cast<size_t>(2 * s->Nrb + devstats_scratch_doubles(Nf, s->Nrb)) * sizeof(double));
  if (rc != ZDT_OK) {
    std::string msg = zdt_last_error();
    zdt_stream_destroy(s);
    fail(rc, "%s", msg.c_str());
    return rc;
  }
  *out = s;
  return ZDT_OK;
}

int zdt_stream_info(const zdt_stream* s, int64_t* Nrb, int64_t* nbatch, int* resident, int64_t* bytes) {
  if (s == nullptr) return fail(ZDT_EINVAL, "stream is NULL");
  if (Nrb) *Nrb = s->Nrb;
  if (nbatch) *nbatch = s->nb;
  if (resident) *resident = s->resident ? 1 : 0;
  if (bytes) *bytes = s->bytes;
  return ZDT_OK;
}

int zdt_stream_search(zdt_stream* s, const float* const* host_specs, int nspec, double snr, int64_t* hits,
                      int64_t cap, int64_t* nhits, double* thr) {
  if (s == nullptr) return fail(ZDT_EINVAL, "stream is NULL");
  if (host_specs == nullptr || nspec < 1 || nhits == nullptr || cap < 0 || (cap > 0 && hits == nullptr))
    return fail(ZDT_EINVAL, "stream_search: need host spectrograms, nhits and (for cap > 0) hits");
  for (int i = 0; i < nspec; ++i)
    if (host_specs[i] == nullptr) return fail(ZDT_EINVAL, "stream_search: spectrogram %d is NULL", i);
  ZDT_TRY_HIP(hipSetDevice(s->device));
  const size_t spec_bytes = static_cast<size_t>(s->Nf * s->Nt) * 4;
  const int64_t Nf = s->Nf, Nrb = s->Nrb;
  auto enqueue_copy = [&](int i) -> int {
    const int slot = i & 1;
    ZDT_TRY_HIP(hipStreamWaitEvent(s->copy, s->used[slot], 0));
    ZDT_TRY_HIP(hipMemcpyAsync(s->spec[slot], host_specs[i], spec_bytes, hipMemcpyHostToDevice, s->copy));
    ZDT_TRY_HIP(hipEventRecord(s->copied[slot], s->copy));
    return ZDT_OK;
  };
  // the events start out "recorded" so the first waits pass
  for (int i = 0; i < 2; ++i) ZDT_TRY_HIP(hipEventRecord(s->used[i], s->compute));
  ZDT_TRY(enqueue_copy(0));
  int rc_trunc = ZDT_OK;
  for (int i = 0; i < nspec; ++i) {
    const int slot = i & 1;
    if (i + 1 < nspec) ZDT_TRY(enqueue_copy(i + 1));  // flies while spectrogram i is searched
    ZDT_TRY_HIP(hipStreamWaitEvent(s->compute, s->copied[slot], 0));
    ZDT_TRY(zdt_input(s->ws, s->spec[slot]));
    ZDT_TRY_HIP(hipEventRecord(s->used[slot], s->compute));
    auto r0b = [&](int64_t b) {
      return static_cast<float>(static_cast<double>(s->r0) + static_cast<double>(b * Nrb) * static_cast<double>(s->dr));
    };
    // statistics over all batches: mean of the first column, least deviation
    double m = 0.0, sd = INFINITY;
    bool nan = false;
    std::vector<double> hst(static_cast<size_t>(2 * Nrb));
    for (int64_t b = 0; b < s->nb; ++b) {
      float* f = s->fdr[s->resident ? b : 0];
      ZDT_TRY(run_devstats(s->ws, r0b(b), f, s->dstats, s->dstats + Nrb, s->dstats + 2 * Nrb));
      ZDT_TRY_HIP(hipMemcpyAsync(hst.data(), s->dstats, hst.size() * sizeof(double), hipMemcpyDeviceToHost,
                                 s->compute));
      ZDT_TRY_HIP(hipStreamSynchronize(s->compute));
      if (b == 0) m = hst[0];
      for (int64_t j = 0; j < Nrb; ++j) {
        const double v = hst[static_cast<size_t>(Nrb + j)];
        if (std::isnan(v)) nan = true;
        sd = std::min(sd, std::sqrt(std::max(v, 0.0)));
      }
    }
    const double t = snr * (nan ? NAN : sd) + m;  // fdrdenormalize (src/fdrutils.jl:104-107)
    if (thr) thr[i] = t;
    // hits in column-major order of the hcat'd batches
    int64_t total = 0;
    int64_t* out = hits ? hits + static_cast<int64_t>(i) * cap : nullptr;
    for (int64_t b = 0; b < s->nb; ++b) {
      float* f = s->fdr[s->resident ? b : 0];
      if (!s->resident) ZDT_TRY(zdt_run(s->ws, r0b(b), ZDT_WINDOW_RECT, nullptr, f));
      const float* fl[1] = {f};
      int64_t nb_hits = 0;
      const int64_t left = std::max<int64_t>(0, cap - total);
      int64_t room = std::min(left, s->dcap);
      int rc = zdt_findprotohits(fl, 1, Nf, Nrb, Nf, t, s->dhits, room, &nb_hits, s->compute);
      if (rc != ZDT_OK && rc != ZDT_ETRUNC) return rc;
      if (nb_hits > room && room < left) {
        // more hits than the staging buffer holds: grow it and extract again
        const int64_t need = std::min(nb_hits, left);
        ZDT_TRY_HIP(hipStreamSynchronize(s->compute));
        (void)hipFree(s->dhits);
        s->bytes -= s->dcap * 8;
        s->dhits = nullptr;
        s->dcap = 0;
        ZDT_TRY(salloc(s, reinterpret_cast<void**>(&s->dhits), static_cast<size_t>(need) * 8));
        s->dcap = need;
        room = need;
        rc = zdt_findprotohits(fl, 1, Nf, Nrb, Nf, t, s->dhits, room, &nb_hits, s->compute);
        if (rc != ZDT_OK && rc != ZDT_ETRUNC) return rc;
      }
      const int64_t got = std::min(nb_hits, room);
      if (got > 0) {
        std::vector<int64_t> tmp(static_cast<size_t>(got));
        ZDT_TRY_HIP(hipMemcpyAsync(tmp.data(), s->dhits, static_cast<size_t>(got) * 8, hipMemcpyDeviceToHost,
                                   s->compute));
        ZDT_TRY_HIP(hipStreamSynchronize(s->compute));
        const int64_t off = b * Nrb * Nf;
        for (int64_t h = 0; h < got; ++h) out[total + h] = tmp[static_cast<size_t>(h)] + off;
      }
      total += nb_hits;
    }
    nhits[i] = total;
    if (total > cap) rc_trunc = fail(ZDT_ETRUNC, "stream_search: spectrogram %d has %lld hits, capacity %lld", i,
                                     (long long)total, (long long)cap);
  }
  ZDT_TRY_HIP(hipStreamSynchronize(s->compute));
  return rc_trunc;
}

}  // extern "C"
