// Fused ZDT path for power-of-two Nf (see zdt_fast.hip).
#pragma once

#include "zdt_internal.h"

namespace zdt {

struct FastPlan;

bool fast_supported(int64_t Nf, int64_t Nt, int64_t Nr);
int64_t fast_bytes(int64_t Nf, int64_t Nt, int64_t Nr);
// n2: the strided FFT length of k1/k3 (512 or 1024; 0 = the default for Nf)
bool fast_n2_ok(int64_t Nf, int n2);
// precision: ZDT_PRECISION_* (the split modes where fast_precision_ok says so)
bool fast_precision_ok(int64_t Nf, int n2, int precision);
int fast_create(int64_t Nf, int64_t Nt, int64_t Nr, float dr, int n2, int precision, FastPlan** out);
void fast_destroy(FastPlan* p);
// input!: spectrogram (device Nf x Nt) -> the plan's input-spectrum buffer
int fast_input(FastPlan* p, const float* spec, hipStream_t st);
// the input spectrum F in natural Nk x Nt layout (for the step functions)
int fast_unpack_F(FastPlan* p, c32* F, hipStream_t st);
// one batch; dest NULL keeps the result for fast_output.  colmean/colvar
// (device, Nr doubles each), when given with a dest, receive the FDR's column
// statistics, computed by k3 as it writes the FDR.
int fast_run(FastPlan* p, float r0, int wkind, const float* window, float* dest, double scale, hipStream_t st,
             double* colmean = nullptr, double* colvar = nullptr);
int fast_output(FastPlan* p, float* dest, double scale, hipStream_t st, double* colmean = nullptr,
                double* colvar = nullptr);
// the next outputs also write the FDR's run maxima into smax (NULL: no
// longer), Nr * Nf/FDR_RUN floats (K3Args::smax); p may be NULL
void fast_set_smax(FastPlan* p, float* smax);
// the plan's strided FFT length N2 (the run maxima's row count), 0 for NULL
int fast_plan_n2(const FastPlan* p);
// doubles of k3's statistics partials (held by every fused plan)
int64_t fast_stats_doubles(int64_t Nf, int64_t Nr);

// Time steps the fused output takes (the chirp-Z runs in 64-step chunks
// above 64; the input spectrum Ft holds Nk x Nt entries)
constexpr int64_t FOUT_MAX_NT = 1024;

// Fused output for power-of-two Nf in [2^16, 2^22] and Nt <= FOUT_MAX_NT: the chirp-Z
// per bin from the input spectrum F, then the c2r through k2_c2r and k3
// (fast_output).  F is either natural (Nk x Nt, column stride ldF) or, with
// ftrans, the layout fast_out_input writes: (Nk-1) x Nt entries [t][k2][k1]
// (bin k2 + N2 k1) followed by the Nt Nyquist bins.
bool fast_out_supported(int64_t Nf, int64_t Nt, int64_t Nr);
int64_t fast_out_bytes(int64_t Nf, int64_t Nr);
// precision: as fast_create (the split modes change k1 and k3)
int fast_out_create(int64_t Nf, int64_t Nr, int precision, FastPlan** out);
// input!: spectrogram (device Nf x Nt) -> Ft (Nk x Nt entries, transposed layout)
int fast_out_input(FastPlan* p, const float* spec, int64_t Nt, c32* Ft, hipStream_t st);
// natural-layout F from Ft (for the step functions)
int fast_out_unpack_F(FastPlan* p, const c32* Ft, int64_t Nt, c32* F, hipStream_t st);
int fast_out_run(FastPlan* p, const c32* F, int64_t ldF, bool ftrans, int Nt, float r0, float dr, int wkind,
                 const float* window, float* dest, double scale, hipStream_t st, double* colmean = nullptr,
                 double* colvar = nullptr);

}  // namespace zdt
