// Workspace internals used by the plan code (not part of the C-ABI).
#pragma once

#include "zdt_internal.h"

namespace zdt {

// internal flag of create_ws_tuned: skip the fused path (the fused output or
// the generic kernels serve the shape); a bit above every public ZDT_FLAG_*
constexpr int ZDT_FLAG_NO_FUSED_INTERNAL = 0x10000;

// zdt_create with the tunables a plan's autotuner chooses: n2, the fused
// path's strided FFT length (512 or 1024; 0 = the default for Nf), and maxr,
// the generic FFTs' largest power-of-two radix (16, 8 or 4).
int create_ws_tuned(const float* spec, int64_t Nf, int64_t Nt, float r0, float dr, int64_t Nr, const int* factors,
                    int nfactors, int flags, int precision, void* stream, int n2, int maxr, zdt_ws** out);

// true when the workspace runs the fused kernels
bool ws_is_fused(const zdt_ws* ws);
// 1: fused kernels; 2: per-bin chirp-Z with the fused output; 0: generic
int ws_path(const zdt_ws* ws);
// the fused paths' strided FFT length N2 (the run maxima's row count), 0 otherwise
int ws_fused_n2(const zdt_ws* ws);

// zdtfdr! (rectangular window) into dest plus the FDR's column mean and
// corrected variance on the device (colmean, colvar: Nr doubles each), all
// enqueued on the workspace stream.  The fused paths compute the statistics
// inside the output kernel; the generic path runs colstats over dest with
// part as scratch (devstats_scratch_doubles).  smax: NULL, or (fused paths,
// ws_path != 0) Nr * Nf/FDR_RUN floats that receive the FDR's run maxima
// for launch_hits_count.
int run_devstats(zdt_ws* ws, float r0, float* dest, double* colmean, double* colvar, double* part,
                 float* smax = nullptr);
int64_t devstats_scratch_doubles(int64_t Nf, int64_t Nr);

}  // namespace zdt
