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
// hits.  Statistics, threshold and hit extraction stay on the device: the
// per-batch column statistics land side by side, one kernel forms
// thr = snr*s + m over the requested columns, and each batch's hits are
// counted, scanned and written after the previous batches' (a device-side
// running total), so the host waits once per spectrogram, and only after the
// next spectrogram's work is queued.  When Nr is not a multiple of the batch
// length the last batch's extra columns (rates past r0 + (Nr-1)*dr) are
// computed but take no part in the statistics or the hits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "zdt_internal.h"
#include "zdt_kernels.h"
#include "zdt_ws.h"

struct zdt_stream {
  int device = 0;
  int64_t Nf = 0, Nt = 0, Nr = 0, Nrb = 0, nb = 0;
  float r0 = 0, dr = 0;
  bool resident = false;
  int64_t bytes = 0;
  zdt_ws* ws = nullptr;
  hipStream_t compute = nullptr, copy = nullptr, readback = nullptr;
  hipEvent_t copied[2] = {nullptr, nullptr}, used[2] = {nullptr, nullptr}, done[2] = {nullptr, nullptr};
  float* spec[2] = {nullptr, nullptr};
  std::vector<float*> fdr;             // resident batch FDRs (or one reused buffer)
  std::vector<float*> smax;            // their run maxima (fused paths; empty otherwise)
  int64_t* dhits[2] = {nullptr, nullptr};  // device hit staging, one per spectrogram in flight
  int64_t dcap = 0;
  int64_t* counts = nullptr;  // one batch's tile counts, scan scratch, batch total, running total
  double* dstats = nullptr;   // colmean, colvar (nb*Nrb each), thr, then generic-path scratch
  void* hsmall = nullptr;     // pinned: per slot {thr (double), total (int64)}
  // zdt_stream_findhits: the hits of one spectrogram (peak, value, size, box
  // for ccap hits), staged on the device
  char* dclus = nullptr;
  int64_t ccap = 0;
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
  // the FDR buffers and their run maxima (one float per FDR_RUN cells)
  const int64_t fdrs = (resident ? nb : 1) * (Nf * Nrb * 4 + (Nf / FDR_RUN) * Nrb * 4);
  const int64_t nt = hits_tiles(Nf, Nrb);
  return ws + fdrs + 2 * Nf * Nt * 4 + 2 * dcap * 8 + (nt + hits_scan_scratch(nt) + 2) * 8 +
         (2 * nb * Nrb + 1 + devstats_scratch_doubles(Nf, Nrb)) * 8;
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
  for (float* m : s->smax)
    if (m) (void)hipFree(m);
  for (int i = 0; i < 2; ++i) {
    if (s->spec[i]) (void)hipFree(s->spec[i]);
    if (s->copied[i]) (void)hipEventDestroy(s->copied[i]);
    if (s->used[i]) (void)hipEventDestroy(s->used[i]);
  }
  if (s->readback) (void)hipStreamSynchronize(s->readback);
  for (int i = 0; i < 2; ++i) {
    if (s->dhits[i]) (void)hipFree(s->dhits[i]);
    if (s->done[i]) (void)hipEventDestroy(s->done[i]);
  }
  if (s->counts) (void)hipFree(s->counts);
  if (s->dstats) (void)hipFree(s->dstats);
  if (s->dclus) (void)hipFree(s->dclus);
  if (s->hsmall) (void)hipHostFree(s->hsmall);
  if (s->readback) (void)hipStreamDestroy(s->readback);
  if (s->copy) (void)hipStreamDestroy(s->copy);
  if (s->compute) (void)hipStreamDestroy(s->compute);
  delete s;
  return ZDT_OK;
}

int zdt_stream_create(int64_t Nf, int64_t Nt, float r0, float dr, int64_t Nr, int64_t Nrb, int64_t hbm_bytes,
                      int flags, int precision, zdt_stream** out) {
  if (out == nullptr) return fail(ZDT_EINVAL, "stream: out is NULL");
  *out = nullptr;
  if (Nf < 2 || Nt < 1 || Nr < 1 || Nrb < 0 || hbm_bytes < 0)
    return fail(ZDT_EINVAL, "stream: bad shape Nf=%lld Nt=%lld Nr=%lld Nrb=%lld", (long long)Nf, (long long)Nt,
                (long long)Nr, (long long)Nrb);
  if (flags & ~(ZDT_FLAG_UNNORMALIZED | ZDT_FLAG_GENERIC | ZDT_FLAG_PROFILE))
    return fail(ZDT_EINVAL, "stream: unsupported flags");
  if (!std::isfinite(r0) || !std::isfinite(dr)) return fail(ZDT_EINVAL, "stream: r0 and dr must be finite");
  const int64_t dcap = int64_t(1) << 20;  // initial device hit staging per spectrogram (grows)
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
                       hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking) != hipSuccess ||
                       hipStreamCreateWithFlags(&s->readback, hipStreamNonBlocking) != hipSuccess))
    rc = fail(ZDT_EHIP, "stream: stream creation failed");
  for (int i = 0; i < 2 && rc == ZDT_OK; ++i) {
    if (hipEventCreateWithFlags(&s->copied[i], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s->used[i], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s->done[i], hipEventDisableTiming) != hipSuccess)
      rc = fail(ZDT_EHIP, "stream: event creation failed");
    if (rc == ZDT_OK) rc = salloc(s, reinterpret_cast<void**>(&s->spec[i]), static_cast<size_t>(Nf * Nt) * 4);
  }
  if (rc == ZDT_OK) rc = zdt_create(nullptr, Nf, Nt, r0, dr, s->Nrb, nullptr, 0, flags, precision, s->compute,
                                      &s->ws);
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
  // run maxima let the hit count skip tiles under the threshold; the fused
  // paths' output kernel writes them (their Nf are multiples of the tile)
  if (rc == ZDT_OK && ws_path(s->ws) != 0 && Nf % (16 * FDR_TILE) == 0 && ws_fused_n2(s->ws) > 0 &&
      Nf / (2 * ws_fused_n2(s->ws)) <= 2048) {
    for (int64_t b = 0; b < nbuf && rc == ZDT_OK; ++b) {
      float* m = nullptr;
      rc = salloc(s, reinterpret_cast<void**>(&m), static_cast<size_t>((Nf / FDR_RUN) * s->Nrb) * 4);
      if (rc == ZDT_OK) s->smax.push_back(m);
    }
  }
  for (int i = 0; i < 2 && rc == ZDT_OK; ++i)
    rc = salloc(s, reinterpret_cast<void**>(&s->dhits[i]), static_cast<size_t>(dcap) * 8);
  const int64_t nt = hits_tiles(Nf, s->Nrb);
  if (rc == ZDT_OK)
    rc = salloc(s, reinterpret_cast<void**>(&s->counts), static_cast<size_t>(nt + hits_scan_scratch(nt) + 2) * 8);
  if (rc == ZDT_OK)
    rc = salloc(s, reinterpret_cast<void**>(&s->dstats),
                static_cast<size_t>(2 * s->nb * s->Nrb + 1 + devstats_scratch_doubles(Nf, s->Nrb)) * sizeof(double));
  if (rc == ZDT_OK && hipHostMalloc(&s->hsmall, 2 * 16, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    s->hsmall = nullptr;
    rc = fail(ZDT_ENOMEM, "stream: pinned staging allocation failed");
  }
  if (rc != ZDT_OK) {
    std::string msg = zdt_last_error();
    zdt_stream_destroy(s);
    fail(rc, "%s", msg.c_str());
    return rc;
  }
  *out = s;
  return ZDT_OK;
}

int zdt_stream_workspace(const zdt_stream* s, zdt_ws** ws) {
  if (s == nullptr || ws == nullptr) return fail(ZDT_EINVAL, "stream_workspace: NULL argument");
  *ws = s->ws;
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

}  // extern "C"

namespace {

// zdt_stream_findhits' outputs: per spectrogram i, hits at [i*cap ...]
struct StreamHits {
  int di, dj;
  int64_t* peak;
  float* value;
  int64_t* size;
  int64_t* box;
  int64_t cap;    // hits per spectrogram in the host arrays
  int64_t stage;  // hits the device staging holds (>= cap, >= 1)
  int64_t* nclusters;
};

// The search pipeline of zdt_stream_search.  With cl, each spectrogram's
// proto-hits (the first `cap` of them staged on the device) are clustered
// right after their extraction, while every batch's FDR is still resident,
// and hits (the host proto-hit array) may be NULL.
int stream_run(zdt_stream* s, const float* const* host_specs, int nspec, double snr, int64_t* hits, int64_t cap,
               int64_t* nhits, double* thr, const StreamHits* cl) {
  ZDT_TRY_HIP(hipSetDevice(s->device));
  const size_t spec_bytes = static_cast<size_t>(s->Nf * s->Nt) * 4;
  const int64_t Nf = s->Nf, Nrb = s->Nrb, nb = s->nb;
  // device hit staging of cap entries per spectrogram in flight
  if (cap > s->dcap) {
    ZDT_TRY_HIP(hipStreamSynchronize(s->compute));
    ZDT_TRY_HIP(hipStreamSynchronize(s->readback));
    for (int k = 0; k < 2; ++k) {
      (void)hipFree(s->dhits[k]);
      s->dhits[k] = nullptr;
    }
    s->bytes -= 2 * s->dcap * 8;
    s->dcap = 0;
    for (int k = 0; k < 2; ++k) ZDT_TRY(salloc(s, reinterpret_cast<void**>(&s->dhits[k]), static_cast<size_t>(cap) * 8));
    s->dcap = cap;
  }
  double* colmean = s->dstats;
  double* colvar = colmean + nb * Nrb;
  double* thr_dev = colvar + nb * Nrb;
  double* scratch = thr_dev + 1;
  const int64_t nt_full = hits_tiles(Nf, Nrb);
  int64_t* counts = s->counts;
  int64_t* scan_scratch = counts + nt_full;
  int64_t* batch_total = scan_scratch + hits_scan_scratch(nt_full);
  int64_t* running = batch_total + 1;

  auto enqueue_copy = [&](int i) -> int {
    const int slot = i & 1;
    ZDT_TRY_HIP(hipStreamWaitEvent(s->copy, s->used[slot], 0));
    ZDT_TRY_HIP(hipMemcpyAsync(s->spec[slot], host_specs[i], spec_bytes, hipMemcpyHostToDevice, s->copy));
    ZDT_TRY_HIP(hipEventRecord(s->copied[slot], s->copy));
    return ZDT_OK;
  };
  auto r0b = [&](int64_t b) {
    return static_cast<float>(static_cast<double>(s->r0) + static_cast<double>(b * Nrb) * static_cast<double>(s->dr));
  };
  // columns of batch b that belong to the requested rates
  auto valid = [&](int64_t b) { return std::min(Nrb, s->Nr - b * Nrb); };
  // the whole search of spectrogram i, enqueued on the compute stream
  auto enqueue_search = [&](int i) -> int {
    const int slot = i & 1;
    ZDT_TRY_HIP(hipStreamWaitEvent(s->compute, s->copied[slot], 0));
    ZDT_TRY(zdt_input(s->ws, s->spec[slot]));
    ZDT_TRY_HIP(hipEventRecord(s->used[slot], s->compute));
    // statistics of every batch, side by side: the first Nr entries are the
    // requested rates' columns
    auto smax_of = [&](int64_t b) { return s->smax.empty() ? nullptr : s->smax[s->resident ? b : 0]; };
    for (int64_t b = 0; b < nb; ++b) {
      float* f = s->fdr[s->resident ? b : 0];
      ZDT_TRY(run_devstats(s->ws, r0b(b), f, colmean + b * Nrb, colvar + b * Nrb, scratch, smax_of(b)));
    }
    ProfUse prof_(ws_profiler(s->ws));  // the hit extraction joins the workspace's kernel profile
    ZDT_TRY(launch_hits_threshold(colmean, colvar, s->Nr, snr, thr_dev, s->compute));
    // hits in column-major order of the hcat'd batches
    for (int64_t b = 0; b < nb; ++b) {
      float* f = s->fdr[s->resident ? b : 0];
      // recomputed batches rewrite their (identical) statistics, with the run maxima
      if (!s->resident)
        ZDT_TRY(run_devstats(s->ws, r0b(b), f, colmean + b * Nrb, colvar + b * Nrb, scratch, smax_of(b)));
      const int64_t nv = valid(b);
      const int64_t nt = hits_tiles(Nf, nv);
      ZDT_TRY(launch_hits_count(f, Nf, nv, Nf, 0.0, thr_dev, counts, s->compute, smax_of(b), ws_fused_n2(s->ws)));
      ZDT_TRY(launch_hits_scan(counts, nt, batch_total, scan_scratch, s->compute));
      if (cap > 0)
        ZDT_TRY(launch_hits_write_impl(f, Nf, nv, Nf, 0.0, thr_dev, counts, nt, 0, batch_total, b * Nrb * Nf,
                                       s->dhits[slot], cap, s->compute, b == 0 ? nullptr : running));
      ZDT_TRY(launch_accumulate_i64(running, batch_total, b == 0, s->compute));
    }
    char* hs = static_cast<char*>(s->hsmall) + 16 * slot;
    ZDT_TRY_HIP(hipMemcpyAsync(hs, thr_dev, sizeof(double), hipMemcpyDeviceToHost, s->compute));
    ZDT_TRY_HIP(hipMemcpyAsync(hs + 8, running, sizeof(int64_t), hipMemcpyDeviceToHost, s->compute));
    ZDT_TRY_HIP(hipEventRecord(s->done[slot], s->compute));
    return ZDT_OK;
  };
  int rc_trunc = ZDT_OK;
  // results of spectrogram i, once its work is done (the next one is queued)
  auto finish = [&](int i) -> int {
    const int slot = i & 1;
    ZDT_TRY_HIP(hipEventSynchronize(s->done[slot]));
    const char* hs = static_cast<const char*>(s->hsmall) + 16 * slot;
    double t;
    int64_t total;
    std::memcpy(&t, hs, sizeof(double));
    std::memcpy(&total, hs + 8, sizeof(int64_t));
    if (thr) thr[i] = t;
    nhits[i] = total;
    const int64_t got = hits ? std::min(total, cap) : 0;
    if (got > 0) {
      ZDT_TRY_HIP(hipMemcpyAsync(hits + static_cast<int64_t>(i) * cap, s->dhits[slot], static_cast<size_t>(got) * 8,
                                 hipMemcpyDeviceToHost, s->readback));
      ZDT_TRY_HIP(hipStreamSynchronize(s->readback));
    }
    if (total > cap)
      rc_trunc = fail(ZDT_ETRUNC, "stream_search: spectrogram %d has %lld proto-hits, capacity %lld", i,
                      (long long)total, (long long)cap);
    return ZDT_OK;
  };
  // the hits of spectrogram i, before the next spectrogram's search reuses
  // the FDR buffers: its proto-hit count is read (one synchronization), the
  // clustering runs on the compute stream and its results are copied out
  auto cluster = [&](int i) -> int {
    const int slot = i & 1;
    ZDT_TRY_HIP(hipEventSynchronize(s->done[slot]));
    int64_t total;
    std::memcpy(&total, static_cast<const char*>(s->hsmall) + 16 * slot + 8, sizeof(int64_t));
    if (total > cap) {  // not every proto-hit is staged: no hits for this spectrogram
      cl->nclusters[i] = -1;
      return ZDT_OK;
    }
    int64_t* dpeak = reinterpret_cast<int64_t*>(s->dclus);
    int64_t* dsize = dpeak + cl->stage;
    int64_t* dbox = dsize + cl->stage;
    float* dvalue = reinterpret_cast<float*>(dbox + 4 * cl->stage);
    int64_t ncl = 0;
    ProfUse prof_(ws_profiler(s->ws));  // the clustering joins the workspace's kernel profile
    const int rc = cluster_protohits_impl(s->dhits[slot], total, s->fdr.data(), static_cast<int>(nb), Nf, Nrb, Nf,
                                          cl->di, cl->dj, dpeak, dvalue, dsize, dbox, cl->cap, &ncl, nullptr,
                                          s->compute);
    if (rc != ZDT_OK && rc != ZDT_ETRUNC) return rc;  // ETRUNC: reported once the stream is done
    cl->nclusters[i] = ncl;
    const int64_t got = std::min(ncl, cl->cap);
    if (got > 0) {
      const int64_t o = static_cast<int64_t>(i) * cl->cap;
      ZDT_TRY_HIP(hipMemcpy(cl->peak + o, dpeak, static_cast<size_t>(got) * 8, hipMemcpyDeviceToHost));
      ZDT_TRY_HIP(hipMemcpy(cl->value + o, dvalue, static_cast<size_t>(got) * 4, hipMemcpyDeviceToHost));
      ZDT_TRY_HIP(hipMemcpy(cl->size + o, dsize, static_cast<size_t>(got) * 8, hipMemcpyDeviceToHost));
      ZDT_TRY_HIP(hipMemcpy(cl->box + 4 * o, dbox, static_cast<size_t>(got) * 32, hipMemcpyDeviceToHost));
    }
    return ZDT_OK;
  };
  // the events start out "recorded" so the first waits pass
  for (int i = 0; i < 2; ++i) ZDT_TRY_HIP(hipEventRecord(s->used[i], s->compute));
  ZDT_TRY(enqueue_copy(0));
  for (int i = 0; i < nspec; ++i) {
    if (i + 1 < nspec) ZDT_TRY(enqueue_copy(i + 1));  // flies while spectrogram i is searched
    ZDT_TRY(enqueue_search(i));
    if (cl) ZDT_TRY(cluster(i));
    if (i > 0) ZDT_TRY(finish(i - 1));
  }
  ZDT_TRY(finish(nspec - 1));
  ZDT_TRY_HIP(hipStreamSynchronize(s->compute));
  if (cl && rc_trunc == ZDT_OK)
    for (int i = 0; i < nspec; ++i)
      if (cl->nclusters[i] > cl->cap)
        rc_trunc = fail(ZDT_ETRUNC, "stream_findhits: spectrogram %d has %lld hits, capacity %lld", i,
                        (long long)cl->nclusters[i], (long long)cl->cap);
  return rc_trunc;
}

}  // namespace

extern "C" {

int zdt_stream_search(zdt_stream* s, const float* const* host_specs, int nspec, double snr, int64_t* hits,
                      int64_t cap, int64_t* nhits, double* thr) {
  if (s == nullptr) return fail(ZDT_EINVAL, "stream is NULL");
  if (host_specs == nullptr || nspec < 1 || nhits == nullptr || cap < 0 || (cap > 0 && hits == nullptr))
    return fail(ZDT_EINVAL, "stream_search: need host spectrograms, nhits and (for cap > 0) hits");
  for (int i = 0; i < nspec; ++i)
    if (host_specs[i] == nullptr) return fail(ZDT_EINVAL, "stream_search: spectrogram %d is NULL", i);
  return stream_run(s, host_specs, nspec, snr, hits, cap, nhits, thr, nullptr);
}

int zdt_stream_findhits(zdt_stream* s, const float* const* host_specs, int nspec, double snr, int di, int dj,
                        int64_t pcap, int64_t* peak, float* value, int64_t* size, int64_t* box, int64_t cap,
                        int64_t* nclusters, int64_t* nprotohits, double* thr) {
  if (s == nullptr) return fail(ZDT_EINVAL, "stream is NULL");
  if (host_specs == nullptr || nspec < 1 || nclusters == nullptr || nprotohits == nullptr || pcap < 1 || cap < 0 ||
      (cap > 0 && (peak == nullptr || value == nullptr || size == nullptr || box == nullptr)) || di < 0 || dj < 0)
    return fail(ZDT_EINVAL, "stream_findhits: need host spectrograms, nclusters, nprotohits, pcap >= 1, "
                            "di, dj >= 0 and (for cap > 0) the hit arrays");
  for (int i = 0; i < nspec; ++i)
    if (host_specs[i] == nullptr) return fail(ZDT_EINVAL, "stream_findhits: spectrogram %d is NULL", i);
  if (!s->resident)
    return fail(ZDT_ENOTSUP, "stream_findhits: the batches' FDRs are not resident together (a hit's peak needs "
                             "every batch); use a larger hbm_bytes or zdt_stream_search");
  // the clustering's limits (zdt_clusterprotohits): channels, hcat'd columns
  // and proto-hits each below 2^31
  if (s->Nf >= (int64_t(1) << 31) || s->nb * s->Nrb >= (int64_t(1) << 31) || pcap >= (int64_t(1) << 31))
    return fail(ZDT_ENOTSUP, "stream_findhits: 2^31 channels, columns or proto-hits and more");
  ZDT_TRY_HIP(hipSetDevice(s->device));
  const int64_t ccap = std::max<int64_t>(cap, 1);
  if (ccap > s->ccap) {
    ZDT_TRY_HIP(hipStreamSynchronize(s->compute));
    if (s->dclus) (void)hipFree(s->dclus);
    s->dclus = nullptr;
    s->bytes -= 52 * s->ccap;
    s->ccap = 0;
    ZDT_TRY(salloc(s, reinterpret_cast<void**>(&s->dclus), static_cast<size_t>(52 * ccap)));
    s->ccap = ccap;
  }
  StreamHits cl{di, dj, peak, value, size, box, cap, s->ccap, nclusters};
  return stream_run(s, host_specs, nspec, snr, nullptr, pcap, nprotohits, thr, &cl);
}

}  // extern "C"
