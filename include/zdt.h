/*
 * zdt.h -- C-ABI of libzdt_hip.so, the MI355X (gfx950) implementation of the
 * Chirp-Z De-Doppler Transform (ZDT) path of FrequencyDriftRateTransforms.jl.
 *
 * Each entry point replaces one method of the reference package; the citation
 * next to it names the reference function (file:line, relative to the
 * reference repository root).  All matrices are column-major, frequency
 * fastest, exactly like the Julia arrays: a spectrogram is Nf x Nt Float32,
 * an FDR matrix is Nf x Nr Float32.  Pointers named spec/dest/fdr/window/
 * hits are DEVICE pointers owned by the caller; the workspace owns every
 * other device buffer.  Every call returns ZDT_OK (0) or an error code, and
 * zdt_last_error() returns the message of the calling thread's last error
 * (the Julia host throws ErrorException with it).
 *
 * Output scale: the FDR is normalized (inverse transforms carry 1/N), i.e.
 * each element is the Fourier-interpolated drift-line sum; this is what
 * src/zdtfdr.jl:120,123,330 computes with AbstractFFTs' scaled plans and what
 * ext/CUDAFrequencyDriftRateTransformsExt.jl:101 reproduces.  Passing
 * ZDT_FLAG_UNNORMALIZED returns Nf*Nl times that (the convention of
 * test/runtests.jl:31).
 *
 * Threading: a workspace is bound to one HIP stream and is not re-entrant
 * (ext/CUDAFrequencyDriftRateTransformsExt.jl:42-45); distinct workspaces may
 * be used concurrently from different threads.
 */
#ifndef ZDT_H
#define ZDT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define ZDT_OK 0
#define ZDT_EINVAL 1   /* bad argument (shape, pointer, window kind, ...)   */
#define ZDT_ENOMEM 2   /* device allocation failed                          */
#define ZDT_EHIP 3     /* a HIP runtime call failed                         */
#define ZDT_ENOTSUP 4  /* size or factor the library does not implement     */
#define ZDT_ETRUNC 5   /* more proto-hits than the caller's capacity        */
#define ZDT_ESTATE 6   /* call order violated (e.g. output before a run)    */

/* window kinds for postprocess (src/zdtfdr.jl:297-321) */
#define ZDT_WINDOW_RECT 0    /* (n,N)->1, the default                       */
#define ZDT_WINDOW_HAMMING 1 /* 0.53836 + 0.46164*cospi(2n/N)               */
#define ZDT_WINDOW_CUSTOM 2  /* device vector of Nf/2+1 Float32 values w(k) */

/* workspace flags */
#define ZDT_FLAG_UNNORMALIZED 0x1 /* scale the FDR by Nf*Nl                  */
#define ZDT_FLAG_GENERIC 0x2      /* force the generic (unfused) kernels     */
#define ZDT_FLAG_PROFILE 0x4      /* time every kernel launch (HIP events)    */
#define ZDT_FLAG_N2_512 0x8       /* fused path: strided FFT length 512 where  *
                                   * 1024 is the default (2^17 <= Nf <= 2^19)  */
#define ZDT_FLAG_RADIX8 0x10      /* generic FFTs: radices up to 8 (16 is the  *
                                   * default; ragged plans time both)          */

/* arithmetic of the fused path's frequency-axis DFT stages (zdt_create's
 * precision argument); the chirp-Z, the phases and every other step stay FP32
 * with FP64 phase arguments in all modes */
#define ZDT_PRECISION_FP32 0       /* FP32 butterflies (the default)            */
#define ZDT_PRECISION_SPLIT_BF16 1 /* DFT stages as BF16 MFMA products of split *
                                    * hi+lo operands, FP32 accumulation         */
#define ZDT_PRECISION_SPLIT_FP16 2 /* the same with FP16 hi+lo operands         */

/* kernel paths reported by zdt_path */
#define ZDT_PATH_GENERIC 0      /* Stockham / four-step FFTs + per-bin chirp-Z  */
#define ZDT_PATH_FUSED 1        /* k1_input, k2_czt, k3_output                  */
#define ZDT_PATH_FUSED_OUTPUT 2 /* per-bin chirp-Z + the fused c2r (k2_c2r, k3) */

/* buffers exposed by zdt_buffer */
#define ZDT_BUF_F 0  /* Nk x Nt ComplexF32: input FFT (natural layout)        */
#define ZDT_BUF_Y 1  /* Nk x Nl ComplexF32: CZT work matrix (step functions)  */
#define ZDT_BUF_V 2  /* Nk x Nl ComplexF32: FFT of the chirp kernel           */

typedef struct zdt_ws zdt_ws;

/* Message of the calling thread's last failed call ("" if none). */
const char* zdt_last_error(void);

/* Library ABI version (incremented on incompatible changes). */
int zdt_version(void);

/* ---------------------------------------------------------------- sizing */

/* calcNl(Nt, Nr, factors) = nextprod(factors, Nt+Nr-1)  -- src/zdtutils.jl:7-9.
 * factors may be NULL (then (2,3,5)).  Returns -1 on bad input. */
int64_t zdt_calc_nl(int64_t Nt, int64_t Nr, const int* factors, int nfactors);

/* growNr(Nt, Nr, factors) = calcNl - Nt + 1  -- src/zdtutils.jl:17-19 */
int64_t zdt_grow_nr(int64_t Nt, int64_t Nr, const int* factors, int nfactors);

/* estimate_memory(Nf, Nt, Nr, Ni, No; factors)  -- src/zdtutils.jl:35-53
 * (the reference's formula, in bytes). */
int64_t zdt_estimate_memory(int64_t Nf, int64_t Nt, int64_t Nr, int64_t Ni,
                            int64_t No, const int* factors, int nfactors);

/* Device bytes this library would allocate for a workspace of that shape
 * (fused path; the step functions add 2*Nk*Nl*8 + scratch on first use). */
int64_t zdt_workspace_bytes(int64_t Nf, int64_t Nt, int64_t Nr,
                            const int* factors, int nfactors, int flags);

/* ------------------------------------------------------------- workspace */

/* ZDTWorkspace(spectrogram, r0, dr, Nr, factors)  -- src/zdtfdr.jl:37-75.
 * Computes Nl, allocates the device buffers, precomputes the chirp kernel and,
 * when spec is not NULL, runs input! on it.  stream is a hipStream_t (NULL =
 * the default stream).  The rates are r0 + j*dr, j < Nr, with r0/dr taken as
 * Float32 (src/zdtfdr.jl:13-14).  precision is one of ZDT_PRECISION_*; the
 * split modes exist for the fused path only (ZDT_ENOTSUP elsewhere). */
int zdt_create(const float* spec, int64_t Nf, int64_t Nt, float r0, float dr,
               int64_t Nr, const int* factors, int nfactors, int flags,
               int precision, void* stream, zdt_ws** out);

/* Frees every device buffer of the workspace (the Julia finalizer). */
int zdt_destroy(zdt_ws* ws);

/* Shape of a workspace (Base.show, src/zdtfdr.jl:89-99). Any pointer may be NULL. */
int zdt_info(const zdt_ws* ws, int64_t* Nf, int64_t* Nt, float* r0, float* dr,
             int64_t* Nr, int64_t* Nl);

/* The kernel path (ZDT_PATH_*) and precision (ZDT_PRECISION_*) a workspace
 * runs; either pointer may be NULL.  No counterpart in the reference, whose
 * only GPU path is the generic one (ext/CUDAFrequencyDriftRateTransformsExt.jl):
 * it lets hosts and tests confirm which kernels serve a shape. */
int zdt_path(const zdt_ws* ws, int* path, int* precision);

/* Device bytes held by the workspace (Base.sizeof, src/zdtfdr.jl:101-103). */
int zdt_sizeof(const zdt_ws* ws, int64_t* bytes);

/* Rebind the workspace to another stream (update_stream,
 * ext/CUDAFrequencyDriftRateTransformsExt.jl:99). */
int zdt_set_stream(zdt_ws* ws, void* stream);

/* fdrsynchronize(::Type{<:CuArray})  -- ext/CUDAFrequencyDriftRateTransformsExt.jl:26-28 */
int zdt_sync(zdt_ws* ws);

/* ------------------------------------------------------------ transform */

/* input!(workspace, spectrogram)  -- src/zdtfdr.jl:209-212.
 * spec: device Nf x Nt Float32, column stride Nf. */
int zdt_input(zdt_ws* ws, const float* spec);

/* zdtfdr!([w,] dest, workspace; r0)  -- src/zdtfdr.jl:349-396.
 * One ZDT batch at rates r0 + j*dr from the current input: preprocess!,
 * convolve!, postprocess! and, when dest is not NULL, output! into dest
 * (device Nf x Nr Float32).  With dest NULL the post-processed spectra stay in
 * the workspace for a later zdt_output (the no-dest method, :382-396).
 * window_kind is one of ZDT_WINDOW_*; window is used only for CUSTOM. */
int zdt_run(zdt_ws* ws, float r0, int window_kind, const float* window, float* dest);

/* zdtfdr!(w, dests, workspace; r0=r0s)  -- src/zdtfdr.jl:349-362: nb batches
 * from one input and one chirp kernel, dests[b] receiving the rates
 * r0s[b] + j*dr.  dests is a HOST array of nb device pointers. */
int zdt_run_batches(zdt_ws* ws, const float* r0s, int nb, int window_kind,
                    const float* window, float* const* dests);

/* output!(dest, workspace)  -- src/zdtfdr.jl:328-331 (CUDA: ext:94-103).
 * Inverse real FFT of the post-processed spectra of the last zdt_run with a
 * NULL dest (or of the step functions) into dest. */
int zdt_output(zdt_ws* ws, float* dest);

/* Step functions with the reference's materialized data flow (Y is Nk x Nl):
 * preprocess!(workspace, r0)  -- src/zdtfdr.jl:221-239
 * convolve!(workspace)        -- src/zdtfdr.jl:249-253
 * postprocess!(w, workspace)  -- src/zdtfdr.jl:297-321 */
int zdt_preprocess(zdt_ws* ws, float r0);
int zdt_convolve(zdt_ws* ws);
int zdt_postprocess(zdt_ws* ws, int window_kind, const float* window);

/* fftfdr!(fdr, workspace, rates)  -- src/fftfdr.jl:126-148 (fdshiftsum!,
 * :86-96, column by column): fdr[:, j] = irfft(sum_t F[:,t] cispi(2 k t
 * rates[j]/Nf)) for an arbitrary HOST list of nrates Float32 rates, from the
 * workspace's input spectrum F; dest is device Nf x nrates.  Always
 * normalized (the reference's scaled irfft plan).  The reference keeps F in
 * its own fftfdr_workspace (:39-59); here any ZDT workspace serves. */
int zdt_fftfdr(zdt_ws* ws, const float* rates, int64_t nrates, float* dest);

/* fdshift!(dest, workspace, rate)  -- src/fftfdr.jl:104-111: the spectrogram
 * de-dopplered at one rate, dest[:, t] = irfft(F[:,t] cispi(2 k t rate/Nf));
 * dest is device Nf x Nt. */
int zdt_fdshift(zdt_ws* ws, float rate, float* dest);

/* Device view of a workspace buffer (ZDT_BUF_*): pointer, rows, columns.
 * Y and V exist once a step function ran; F exists after zdt_input. */
int zdt_buffer(zdt_ws* ws, int which, void** ptr, int64_t* rows, int64_t* cols);

/* ------------------------------------------------------------ profiling
 * No counterpart in the reference: a measurement hook for benchmarks.  With
 * ZDT_FLAG_PROFILE every kernel the workspace launches is bracketed by a HIP
 * event pair on the workspace stream.  zdt_profile_read synchronizes and
 * aggregates the launches recorded since the last zdt_profile_reset by kernel
 * name: entry idx (0-based, < *count) yields the name, the summed duration in
 * milliseconds, the launch count and the algorithmic HBM bytes per launch. */
int zdt_profile_reset(zdt_ws* ws);
int zdt_profile_read(zdt_ws* ws, int idx, int* count, char* name, int name_len,
                     double* total_ms, int64_t* launches, double* bytes_per_launch);

/* FP32 VALU issue rate of one instruction form (0: VOP3 FMA with SGPR
 * operands, 1: VOP3 FMA, 2: VOP2 FMAC, 3: VOP2 add) with `chains` independent
 * dependency chains per lane (1, 2, 4 or 8), in lane operations per second
 * over a full-device launch.  No counterpart in the reference: a measurement
 * for the compute roofline of bench.py (bench.py --valu-calibration). */
int zdt_calib_valu(int form, int chains, double* lane_ops_per_s);

/* Cost of 8-byte LDS accesses of one pattern (0..12: reads 0-5, writes 6-12;
 * the row layouts and strides of the fused kernels' 512-point FFT passes,
 * listed in csrc/zdt_calib.hip), in nanoseconds per wavefront access and
 * CU over a full-device launch.  No counterpart in the reference: a
 * measurement behind the fused kernels' LDS layout (bench.py
 * --lds-calibration). */
int zdt_calib_lds(int pattern, double* ns_per_access);

/* ---------------------------------------------------------- statistics */

/* fdrstats(fdr) / fdrstats(fdrs)  -- src/fdrutils.jl:39-50.
 * m = mean of column 1 of fdrs[0]; s = minimum over all columns of all
 * matrices of the corrected (N-1) standard deviation.  fdrs: HOST array of nb
 * device pointers to Nf x Nr matrices with column stride ld. FP64 accumulation. */
int zdt_fdrstats(const float* const* fdrs, int nb, int64_t Nf, int64_t Nr,
                 int64_t ld, void* stream, double* m, double* s);

/* fdrnormalize!(fdr[s], m, s)  -- src/fdrutils.jl:70-90: x = (x - m)/s in place. */
int zdt_fdrnormalize(float* const* fdrs, int nb, int64_t Nf, int64_t Nr,
                     int64_t ld, double m, double s, void* stream);

/* findprotohits(fdr[s], threshold)  -- src/fdrutils.jl:148-169.
 * Every element >= threshold, in column-major order of the hcat'd batches, is
 * written to hits (device) as the 0-based linear index i + Nf*(b*Nr + j)
 * (the host turns it into the 1-based CartesianIndex (i+1, b*Nr+j+1)).  At
 * most cap indices are written; *nhits receives the total count, and the call
 * returns ZDT_ETRUNC when it exceeds cap.  hits may be NULL with cap 0 to count.
 * fdrdenormalize (snr*s + m, src/fdrutils.jl:104-119) is done by the host. */
int zdt_findprotohits(const float* const* fdrs, int nb, int64_t Nf, int64_t Nr,
                      int64_t ld, double threshold, int64_t* hits, int64_t cap,
                      int64_t* nhits, void* stream);

/* zdt_findprotohits without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (serving
 * loops and harnesses that forbid hipMalloc).  Same arguments, results and
 * errors as zdt_findprotohits; its tile counts and scan totals live in one
 * library-owned device array per GPU of ZDT_HITS_NOALLOC_MAX_TILES tiles
 * (16 MiB, held until the process ends), so nb * ceil(Nf*Nr/4096) must not
 * exceed ZDT_HITS_NOALLOC_MAX_TILES (ZDT_ENOTSUP otherwise, before any
 * launch).  Calls on one GPU share that array: calls that overlap in time on
 * the same GPU (other host threads, or other streams) race and must not be
 * issued. */
#define ZDT_HITS_NOALLOC_MAX_TILES (1LL << 21)
#define ZDT_HITS_NOALLOC_MAX_CELLS (ZDT_HITS_NOALLOC_MAX_TILES * 4096LL)
int zdt_findprotohits_noalloc(const float* const* fdrs, int nb, int64_t Nf, int64_t Nr,
                              int64_t ld, double threshold, int64_t* hits, int64_t cap,
                              int64_t* nhits, void* stream);

/* Clustering of neighbouring proto-hits into hits.  The reference stops at
 * proto-hits: findprotohits' docstring (src/fdrutils.jl:141-144) and
 * README.md:48-53 leave "clustering of neighboring proto-hits" to the next
 * stage of a search, outside the package.  Two proto-hits are neighbours when
 * their channels differ by at most di and their hcat'd rate columns by at most
 * dj (di = dj = 1: the 8 surrounding cells); a hit is one connected component.
 *
 * hits (device): n proto-hits as zdt_findprotohits writes them, strictly
 * ascending 0-based linear indices i + Nf*J into the hcat of the nb Nf x Nr
 * matrices fdrs (HOST array of device pointers, column stride ld), which
 * supply the values.  Hits are numbered in the order of their first
 * proto-hit.  For the first cap hits (device arrays of cap entries):
 *   peak[c]   linear index of the largest value (ties: the earliest proto-hit)
 *   value[c]  that value
 *   size[c]   number of proto-hits
 *   box[4c..4c+3]  first channel, last channel, first column, last column
 * label (device, n entries, or NULL): each proto-hit's hit number.
 * *nclusters receives the number of hits; ZDT_ETRUNC when it exceeds cap
 * (the first cap hits and every label are still written).  ZDT_EINVAL for an
 * index out of range or out of order (the outputs' contents are then
 * unspecified; nothing is written outside them), ZDT_ENOTSUP when n, Nf or
 * nb*Nr reaches 2^31.
 * Synchronizes the stream once (to size the outputs). */
int zdt_clusterprotohits(const int64_t* hits, int64_t n, const float* const* fdrs, int nb,
                         int64_t Nf, int64_t Nr, int64_t ld, int di, int dj, int64_t* peak,
                         float* value, int64_t* size, int64_t* box, int64_t cap,
                         int64_t* nclusters, int64_t* label, void* stream);

/* --------------------------------------------------------- ragged plans
 * Many coarse channels of uneven Nf and Nt, every one through input!, one
 * zdtfdr! batch, fdrdenormalize and findprotohits, replayed as ONE hipGraph
 * launch.  No single counterpart in the reference: it replaces the host loop
 * a reference user writes over channels around zdtfdr! (src/zdtfdr.jl:349-421),
 * fdrstats/fdrdenormalize (src/fdrutils.jl:39-50,104-119) and findprotohits
 * (src/fdrutils.jl:148-169).
 *
 * Channels with equal (Nf, Nt, r0, dr, Nr) share one workspace and run one
 * after another; distinct shapes run on parallel branches of the graph.
 * With ZDT_PLAN_AUTOTUNE, plan creation times the candidate kernel paths and
 * shapes of each workspace and keeps the fastest: the fused kernels with a
 * strided FFT length N2 of 1024 or 512 (k1/k3 tiles of 32 or 16 rows), the
 * fused output, and the generic kernels with FFT radices up to 16 or 8 (each
 * where the shape allows it).  With a tuning table (zdt_tune_table) the
 * choices persist across processes. */
typedef struct zdt_plan zdt_plan;

#define ZDT_PLAN_AUTOTUNE 0x100

/* Channel c has an Nf[c] x Nt[c] spectrogram and rates r0[c] + j*dr[c],
 * j < Nr[c].  flags: ZDT_FLAG_UNNORMALIZED, ZDT_FLAG_GENERIC, ZDT_PLAN_AUTOTUNE.
 * stream: the caller's stream that zdt_plan_launch orders against. */
int zdt_plan_ragged(int nch, const int64_t* Nf, const int64_t* Nt, const float* r0,
                    const float* dr, const int64_t* Nr, int flags, void* stream,
                    zdt_plan** out);
int zdt_plan_destroy(zdt_plan* p);

/* The choices made for channel c: shape index (channels with one shape share
 * a workspace), path (ZDT_PATH_*: 1 = fused, power-of-two Nf in [2^16, 2^20]
 * with Nt <= 16; 2 = per-bin chirp-Z with the fused output, power-of-two Nf in
 * [2^16, 2^22] with Nt <= 1024 and Nr <= 65535 where path 1 does not apply;
 * 0 = generic), n2 (the strided FFT length of paths 1 and 2) and maxr (the
 * radix cap of the generic path's FFTs; 0 on paths 1 and 2, which do not use
 * them for the transform). */
int zdt_plan_info(const zdt_plan* p, int ch, int* shape, int* path, int* n2, int* maxr);

/* Where channel c's choices came from: source 0 = defaults (no autotuning),
 * 1 = timed at plan creation, 2 = read from the tuning table; ms = the chosen
 * candidate's time for input! plus one batch (0 when not tuned). */
int zdt_plan_tuning(const zdt_plan* p, int ch, int* source, float* ms);

/* Names the tuning table (a text file, created on first write) that
 * ZDT_PLAN_AUTOTUNE plans read before timing a shape and append their
 * timings to, keyed by device (gcnArchName, CU count) and shape (Nf, Nt, Nr,
 * forced generic); NULL or "" turns the table off (the default).
 * Process-wide.  No counterpart in the reference (FFTW and cuFFT plan by
 * themselves): it keeps a plan's start-up from re-timing known shapes. */
int zdt_tune_table(const char* path);

/* Binds device buffers and captures the graph (no work is done yet).
 * specs[c]: channel c's spectrogram.  dests[c]: its Nf x Nr FDR, or NULL (or
 * dests NULL) to use a buffer owned by the plan and shared by the channels of
 * one shape.  thr[c] (host): the threshold, or the SNR when snr != 0
 * (threshold = snr*s + m from the channel's own fdrstats).  hits (device):
 * nch*cap int64, channel c's 0-based column-major linear indices at
 * hits[c*cap ...]; nhits (device, nch int64): channel c's hit count, which
 * may exceed cap (then only the first cap are written).  thr_out (device,
 * nch doubles, may be NULL) receives the thresholds used.  Host arrays are
 * read during the call only.  Binding again replaces the graph. */
int zdt_plan_bind(zdt_plan* p, const float* const* specs, float* const* dests,
                  const double* thr, int snr, int64_t* hits, int64_t cap,
                  int64_t* nhits, double* thr_out);

/* One hipGraphLaunch of the bound graph, ordered after the work already on
 * the plan's stream and before the work queued on it afterwards. */
int zdt_plan_launch(zdt_plan* p);

/* -------------------------------------------------- long-integration search
 * Spectrograms of one shape streamed from host memory (the copy of the next
 * one overlaps the search of the current one), the rates r0 + j*dr cut into
 * equal batches of Nrb rates sized to an HBM budget, and proto-hits at an SNR
 * threshold over all batches of each spectrogram.  No single counterpart in
 * the reference: it replaces the user's loop around batchrates
 * (src/batchrates.jl:22-60), estimate_memory (src/zdtutils.jl:35-53),
 * zdtfdr!(dests, ws; r0=batch starts) (src/zdtfdr.jl:349-362) and
 * findprotohits(fdrs, snr; snr=true) (src/fdrutils.jl:148-169).  Like
 * batchrates, the last batch is padded: nbatch*Nrb >= Nr rates are searched. */
typedef struct zdt_stream zdt_stream;

/* Pinned host memory (hipHostMalloc) for spectrograms that should stream
 * asynchronously; pageable memory works too but its copies do not overlap. */
int zdt_host_alloc(size_t bytes, void** p);
int zdt_host_free(void* p);

/* Decodes one bitshuffle-filtered HDF5 chunk (filter 32008) on the host: the
 * compression of the Breakthrough Listen filterbank files that the
 * reference's heavy test reads through the H5Zbitshuffle plugin
 * (test/heavytests.jl:6-14).  No counterpart in the reference itself; the
 * filterbank ingest uses it with HDF5's direct chunk reads when libhdf5 has
 * no bitshuffle plugin.  in: the stored chunk (in_len bytes); out: the
 * chunk's out_len uncompressed bytes; elem_size: the filter's cd_values[2];
 * compression: cd_values[4], 0 (bit transposition only, in blocks of
 * block_size elements, 0 = the library's default) or 2 (LZ4: the block size
 * comes from the chunk's header).  ZDT_EINVAL for a malformed chunk (every
 * read and write is bounds-checked), ZDT_ENOTSUP for other compressions. */
int zdt_bshuf_decompress(const void* in, size_t in_len, void* out, size_t out_len,
                         size_t elem_size, size_t block_size, int compression);

/* Nrb = 0 chooses the batch length: the largest that keeps every batch FDR
 * resident within hbm_bytes, else the largest that fits with one FDR buffer
 * (each batch is then computed twice: once for the statistics, once for the
 * hits).  hbm_bytes = 0 means 90% of the free device memory.
 * flags: ZDT_FLAG_UNNORMALIZED, ZDT_FLAG_GENERIC, ZDT_FLAG_PROFILE (every
 * kernel of a search, the hit extraction included, is timed on the search's
 * workspace; see zdt_stream_workspace).  precision: ZDT_PRECISION_* of the
 * search's workspace (as zdt_create). */
int zdt_stream_create(int64_t Nf, int64_t Nt, float r0, float dr, int64_t Nr,
                      int64_t Nrb, int64_t hbm_bytes, int flags, int precision, zdt_stream** out);
int zdt_stream_destroy(zdt_stream* s);
int zdt_stream_info(const zdt_stream* s, int64_t* Nrb, int64_t* nbatch, int* resident,
                    int64_t* bytes);

/* The search's workspace, owned by the stream: zdt_path, zdt_info and, with
 * ZDT_FLAG_PROFILE, zdt_profile_reset / zdt_profile_read on it.  No
 * counterpart in the reference. */
int zdt_stream_workspace(const zdt_stream* s, zdt_ws** ws);

/* Searches nspec HOST spectrograms (Nf x Nt Float32, column-major).  For
 * spectrogram i: thr[i] = snr*s + m over all its batches (may be NULL),
 * nhits[i] = its hit count, and hits[i*cap ...] its first min(nhits[i], cap)
 * 0-based linear indices i + Nf*(b*Nrb + j) (HOST arrays).  Returns
 * ZDT_ETRUNC when some spectrogram has more than cap hits. */
int zdt_stream_search(zdt_stream* s, const float* const* host_specs, int nspec, double snr,
                      int64_t* hits, int64_t cap, int64_t* nhits, double* thr);

/* The same search, ending in hits instead of proto-hits: each spectrogram's
 * proto-hits are clustered as zdt_clusterprotohits clusters them (radii di,
 * dj over the hcat'd batches), on the device, while every batch's FDR is
 * still resident, before the next spectrogram's search reuses the buffers.
 * No counterpart in the reference, which leaves the clustering to the next
 * stage of a search (src/fdrutils.jl:141-144, README.md:48-53).
 *
 * pcap: proto-hits staged per spectrogram; a spectrogram with more gets
 * nclusters[i] = -1 and no hits.  For spectrogram i (HOST arrays):
 * nprotohits[i], thr[i] (may be NULL) as zdt_stream_search gives them,
 * nclusters[i], and its first min(nclusters[i], cap) hits at peak[i*cap ...],
 * value[i*cap ...], size[i*cap ...] and box[4*i*cap ...] (0-based linear
 * peak index i + Nf*(b*Nrb + j); box: first and last channel, first and last
 * hcat'd column).  Returns ZDT_ETRUNC when some spectrogram has more than pcap
 * proto-hits or more than cap hits, ZDT_ENOTSUP for a stream whose batches
 * are not resident together (a hit's peak needs every batch).  Synchronizes
 * the host with each spectrogram once, after its proto-hits. */
int zdt_stream_findhits(zdt_stream* s, const float* const* host_specs, int nspec, double snr,
                        int di, int dj, int64_t pcap, int64_t* peak, float* value, int64_t* size,
                        int64_t* box, int64_t cap, int64_t* nclusters, int64_t* nprotohits,
                        double* thr);

#ifdef __cplusplus
}
#endif

#endif /* ZDT_H */
