// Launchers of the generic ZDT kernels, the step-function kernels and the
// FDR statistics / proto-hit kernels.
#pragma once

#include "zdt_internal.h"

namespace zdt {

// Cells per tile of the proto-hit kernels: tiles are FDR_TILE consecutive
// cells of the column-major matrix (fdr_stats.hip).  The fused output kernel
// can also write the maximum of every run of FDR_RUN cells (K3Args::smax),
// FDR_TILE/FDR_RUN consecutive floats per tile.
constexpr int64_t FDR_TILE = 4096;
constexpr int64_t FDR_RUN = 32;

struct CztArgs {
  const c32* F;
  int64_t ldF;  // Nk x Nt
  c32* Ys;
  int64_t ldY;  // Nk x Nr
  int64_t Nk, Nf;
  int Nt, Nr;
  double hd;  // dr/(2 Nf): chirp(k,l) = exp(-2 pi i k l^2 hd)
  double h0;  // r0/Nf:     pre(k,t) = exp(2 pi i (k t^2 hd + k t h0))
  double hc;  // czt_reinsch: rc/Nf, rc = r0 + (Nr-1) dr/2 the centre rate
  double hz;  // czt_reinsch: (r0 - rc)/Nf
  int wkind;
  const float* wcustom;
  float scale;
};

// F -> Ys by czt_reinsch (Nt <= 64) or czt_reinsch_chunks (Nt > 64)
int launch_czt(const CztArgs& a, hipStream_t st);
// fftfdr: Ys[k,j] = sum_t F[k,t] e^{2 pi i k t rates[j]/Nf} (rates_over_nf: device, Nr doubles)
int launch_czt_rates(const c32* F, int64_t ldF, c32* Ys, int64_t ldY, int64_t Nk, int64_t Nf, int Nt,
                     const double* rates_over_nf, int Nr, hipStream_t st);
// fdshift: Y[k,t] = F[k,t] e^{2 pi i k t rate/Nf}
int launch_shift_phase(const c32* F, int64_t ldF, c32* Y, int64_t ldY, int64_t Nk, int Nt, double rate_over_nf,
                       hipStream_t st);

int launch_step_prephase(const c32* F, int64_t ldF, c32* Y, int64_t ldY, int64_t Nk, int Nt, int Nl,
                         double hd, double h0, hipStream_t st);
int launch_step_vgen(c64* V, int64_t Nk, int Nt, int Nr, int Nl, double hd, hipStream_t st);
int launch_c64_to_c32(const c64* in, c32* out, int64_t n, hipStream_t st);
int launch_step_vmul(c32* Y, const c32* V, int64_t n, hipStream_t st);
int launch_step_postphase(c32* Y, int64_t ldY, int64_t Nk, int64_t Nf, int Nr, double hd, int wkind,
                          const float* wcustom, float scale, hipStream_t st);

// FDR statistics: per-column mean and corrected variance (FP64) of an
// Nf x Nr matrix with column stride ld.  part needs 2*Nr*colstats_nchunk(Nf)
// doubles of scratch; colmean/colvar receive Nr doubles each.
int colstats_nchunk(int64_t Nf);
int launch_colstats2(const float* fdr, int64_t Nf, int64_t Nr, int64_t ld, double* part, double* colmean,
                     double* colvar, hipStream_t st);
// x = (x - m)/s in place
int launch_normalize(float* fdr, int64_t Nf, int64_t Nr, int64_t ld, double m, double s, hipStream_t st);
// proto-hits: count per tile, exclusive scan, ordered write (see fdr_stats.hip)
int64_t hits_tiles(int64_t Nf, int64_t Nr);
// thr_dev, when not NULL, holds the threshold on the device (thr is ignored)
// smax: NULL, or the matrix's run maxima as the fused output writes them
// (K3Args::smax, [Nr][N1/16][smax_n2] with N1 = Nf/(2 smax_n2); needs ld == Nf):
// tiles whose runs are all under the threshold are not read
int launch_hits_count(const float* fdr, int64_t Nf, int64_t Nr, int64_t ld, double thr, const double* thr_dev,
                      int64_t* counts, hipStream_t st, const float* smax = nullptr, int smax_n2 = 0);
// scratch: hits_scan_scratch(n) int64 words
int64_t hits_scan_scratch(int64_t n);
int launch_hits_scan(int64_t* counts, int64_t n, int64_t* total, int64_t* scratch, hipStream_t st);
// offsets: the scanned tile counts, this matrix's tiles from tile0 on; hits
// land at offsets + *pos_base (pos_base: device, NULL = 0) as linear indices
// + base_index, at most cap of them
int launch_hits_write_impl(const float* fdr, int64_t Nf, int64_t Nr, int64_t ld, double thr,
                           const double* thr_dev, const int64_t* offsets, int64_t ntiles_all, int64_t tile0,
                           const int64_t* total_dev, int64_t base_index, int64_t* hits, int64_t cap,
                           hipStream_t st, const int64_t* pos_base = nullptr);
// *acc = (init ? 0 : *acc) + *v on the device
int launch_accumulate_i64(int64_t* acc, const int64_t* v, bool init, hipStream_t st);
// *p = v by a kernel (graph-capturable, unlike a host-to-device copy)
int launch_set_double(double* p, double v, hipStream_t st);
// thr = snr * sqrt(min colvar) + colmean[0], on the device
int launch_hits_threshold(const double* colmean, const double* colvar, int64_t Nr, double snr, double* thr,
                          hipStream_t st);

// Connected components of proto-hits (hit_cluster.hip): hits are n strictly
// ascending linear indices into the hcat of nb Nf x Nr matrices (column stride
// ld); fdrs is the HOST array of the nb device pointers.  Writes the first cap
// clusters' peak index, peak value, size and box (4 per cluster), and, when
// label is not null, every proto-hit's cluster ordinal; *nclusters = count.
// Synchronizes st once.  ZDT_ETRUNC when the count exceeds cap.
int cluster_protohits_impl(const int64_t* hits, int64_t n, const float* const* fdrs, int nb, int64_t Nf,
                           int64_t Nr, int64_t ld, int di, int dj, int64_t* peak, float* value, int64_t* size,
                           int64_t* box, int64_t cap, int64_t* nclusters, int64_t* label, hipStream_t st);

}  // namespace zdt
