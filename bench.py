"""Benchmark of the ZDT hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): one Float32
spectrogram of Nf=2^20 channels x Nt=16 time steps (Exponential(1) noise plus
16 drifting tones, synthetic, seeded), drift rates -4:1/16:4 (Nr=129, Nl=144).
One step = zdtfdr!(dest, ws, spectrogram) = input! + one ZDT batch writing the
2^20 x 129 Float32 FDR, with the spectrogram already resident in HBM.

    python bench.py [--gpus 1] [--steps 200] [--warmup 100] [--no-cpu-baseline]
    python bench.py --ragged [--channels 64] [--steps 20] [--warmup 5]
    python bench.py --long [--spectrograms 6] [--precision fp32|split-bf16|split-fp16] [--hits]
    python bench.py --cluster [--steps 20] [--warmup 5]
    python bench.py --copy-calibration | --valu-calibration | --lds-calibration

The second form times the ragged multi-channel plan (one hipGraph launch
over 64 channels of uneven Nf and Nt), the third the long-integration
search streamed from pinned host memory, the fourth the proto-hit
clustering on the workload's FDR, the last the device measurements behind
the rooflines; each prints its own JSON line.

Prints ONE JSON line.  value = FDR cells/s = Nf*Nr / (seconds per step).
The default 100 warm-up steps (~45 ms) let the shader clock settle: the
first ~60 back-to-back steps after an idle GPU run up to 17% slower (0.51 ms
falling to a steady 0.43 ms per step), and a long search runs at the
steady rate.
The roofline object covers the kernel with the largest share of the step
time, timed live with HIP events on the workspace stream (library profiling
hook) over the same number of steps run right before the timed ones: an
event pair around every launch serializes the queue, so the timed steps run
without them.  Its algorithmic bytes per launch come from the library.  The
cpu_baseline object times the FP32 CPU restatement of the reference
(oracle.zdt_oracle.RefWorkspace32) on one batch of the same workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
PKG = ROOT / "frequencydriftratetransforms.jl_amd"
sys.path.insert(0, str(PKG))

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector (non-packed VOP2 FMAs reach it too)


def k2_flops(Nf: int, Nt: int, Nr: int, N2: int = 1024) -> float:
    """Algorithmic FP32 flops of one k2_czt launch: Horner (Nt-1 complex
    multiply-adds, 8 flops, per bin and rate), the inverse DFT_N1 of every
    column and rate and the forward one of every column and time step
    (5 N log2 N each), and ~30 flops per bin and rate of c2r pre-pass,
    power recurrence and twiddles."""
    M = Nf // 2
    N1 = M // N2
    fft = 5.0 * N1 * np.log2(N1)
    return (M + 1) * Nr * (Nt - 1) * 8.0 + N2 * (Nr + Nt) * fft + (M + 1) * Nr * 30.0
NF, NT = 1 << 20, 16
R0, DR, NR = -4.0, 0.0625, 129


def step_hbm(Nf: int, Nt: int, Nr: int, dt: float) -> dict:
    """Whole-step HBM rate on the reference dataflow's byte count: spectrogram
    read, F written and read, the chirp spectra V read, Ys written and read,
    FDR written (B_alg = 4 Nf Nt + 16 Nk Nt + 8 Nk Nl + 16 Nk Nr + 4 Nf Nr,
    Nk = Nf/2+1, Nl = nextprod((2,3,5), Nt+Nr-1)).  This design moves fewer
    bytes than that (no V, no separate F round trip), so the figure is the
    reference-equivalent rate, not this path's own traffic."""
    Nk = Nf // 2 + 1
    Nl = Nt + Nr - 1
    while not _smooth(Nl):
        Nl += 1
    b = 4 * Nf * Nt + 16 * Nk * Nt + 8 * Nk * Nl + 16 * Nk * Nr + 4 * Nf * Nr
    gbs = b / dt / 1e9
    # this path's own traffic: spectrogram in and G out (k1), G in and B out
    # (k2), B in and the FDR out (k3), G and B being (Nf/2) x (Nt or Nr) complex
    moved = 12 * Nf * Nt + 12 * Nf * Nr
    return {"basis": "B_alg (reference dataflow)", "Nl": Nl, "bytes_per_step": b, "achieved": gbs,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
            "moved": {"basis": "bytes the fused kernels move (k1 + k2 + k3)", "bytes_per_step": moved,
                      "achieved": moved / dt / 1e9, "frac": moved / dt / 1e9 / HBM_PEAK_GBS}}


def _smooth(n: int) -> bool:
    for f in (2, 3, 5):
        while n % f == 0:
            n //= f
    return n == 1


def synthetic_spectrogram(Nf: int, Nt: int, seed: int = 1234) -> np.ndarray:
    """Exponential(1) noise plus 16 tones (amplitude U(10,50), rate U(-4,4))."""
    rng = np.random.default_rng(seed)
    s = rng.exponential(1.0, size=(Nt, Nf)).astype(np.float32)  # row t = time step
    trng = np.random.default_rng(seed + 7919)
    c0 = trng.uniform(0, Nf, 16)
    rate = trng.uniform(-4, 4, 16)
    amp = trng.uniform(10, 50, 16)
    for c, r, a in zip(c0, rate, amp):
        for t in range(Nt):
            s[t, int(round(c + r * t)) % Nf] += np.float32(a)
    return s  # (Nt, Nf) row-major == (Nf, Nt) column-major


def traffic_from_profiles(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (profiles/traffic_<kernel>_2p20x16x129.txt), or None.  The summary
    holds FETCH_SIZE and WRITE_SIZE in KiB per dispatch and the factors that
    map each counter onto the bytes of a plain device copy of known size."""
    p = ROOT / "profiles" / f"traffic_{kernel}_2p20x16x129.txt"
    if not p.exists():
        return None
    vals = {}
    for line in p.read_text().splitlines():
        if "=" in line and not line.lstrip().startswith("#"):
            k, v = line.split("=", 1)
            try:
                vals[k.strip()] = float(v.split()[0])
            except ValueError:
                pass
    try:
        return 1024.0 * (vals["FETCH_KIB_PER_DISPATCH"] * vals["FETCH_SCALE"]
                         + vals["WRITE_KIB_PER_DISPATCH"] * vals["WRITE_SCALE"])
    except KeyError:
        return None


def copy_calibration() -> int:
    """A plain 1 GiB device-to-device copy (three launches): run under
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to get the counters' scale."""
    import torch
    n = 1 << 28
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": 4 * n}))
    return 0


ACC_COLS = (0, 37, 64, 91, 128)  # rate columns whose FP64 oracle values the accuracy figures use


def valu_calibration() -> int:
    """FP32 VALU issue rates (zdt_calib_valu): four instruction forms with 1,
    2, 4 and 8 independent dependency chains per lane over a full-device
    launch.  Prints one JSON line, rates in TFLOP/s counting an FMA as two
    flops (an add as two as well, for comparison)."""
    import torch  # noqa: F401  (the HIP runtime the library binds to)

    from fdrt_hip import _lib
    L = _lib.lib()
    forms = ("vop3_sgpr", "vop3_vgpr", "vop2_fmac", "vop2_add")
    out = {}
    for f, name in enumerate(forms):
        out[name] = {}
        for nch in (1, 2, 4, 8):
            r = C.c_double()
            _lib.check(L.zdt_calib_valu(f, nch, C.byref(r)))
            out[name][f"chains_{nch}"] = 2.0 * r.value / 1e12
    print(json.dumps({"valu_calibration_tflops": out, "peak": FP32_PEAK_TFLOPS}))
    return 0


LDS_PATTERNS = ("read_unpadded", "read_pad8", "read_pad16", "read_pad32", "read_q64", "read_stride8_pad8",
                "write_unpadded", "write_pad8", "write_q64", "write_pass1_pad8", "write_pass1_q64",
                "write_pass2_q64", "write_pass2_pad8")


def lds_calibration() -> int:
    """Cost of 8-byte LDS accesses in the row layouts and strides of the
    fused kernels' 512-point FFT passes (zdt_calib_lds), in ns per wavefront
    access and CU.  Prints one JSON line; under rocprofv3 --pmc
    SQ_LDS_BANK_CONFLICT the kernels lds_pattern<P> give the conflict cycles
    of pattern P."""
    import torch  # noqa: F401  (the HIP runtime the library binds to)

    from fdrt_hip import _lib
    L = _lib.lib()
    out = {}
    for pat, name in enumerate(LDS_PATTERNS):
        r = C.c_double()
        _lib.check(L.zdt_calib_lds(pat, C.byref(r)))
        out[f"{pat}_{name}"] = r.value
    print(json.dumps({"lds_calibration_ns_per_wave_access_per_cu": out}))
    return 0


def cpu_baseline(gpu_cols: dict, seed: int = 1234) -> tuple:
    """One batch of the FP32 CPU restatement of the reference on one core
    (scipy.fft workers=1: a default single-threaded Julia session runs FFTW
    and the broadcasts on one thread) and, as an upper bound, with the FFTs on
    every core of the host.  As the checker it also gives the accuracy of the
    GPU FDRs (gpu_cols: precision -> the ACC_COLS columns, Nf x 5, and the
    full FDR's column sums) against the FP64 oracle and against the
    restatement's own FDR."""
    import scipy.fft as sfft
    sys.path.insert(0, str(ROOT))
    from oracle import host_oracle as H  # checker / CPU baseline only
    from oracle import zdt_oracle as Z  # checker / CPU baseline only
    s = np.asfortranarray(synthetic_spectrogram(NF, NT, seed).T)
    ws = Z.RefWorkspace32(s, R0, DR, NR)
    with sfft.set_workers(1):
        t0 = time.perf_counter()
        ws.input(s)
        ref32 = ws.zdtfdr()
        dt = time.perf_counter() - t0
    ncpu = len(os.sched_getaffinity(0))
    with sfft.set_workers(ncpu):
        t0 = time.perf_counter()
        ws.input(s)
        ws.zdtfdr()
        dt_all = time.perf_counter() - t0
    base = {"value": NF * NR / dt, "unit": "FDR cells/s", "cores": 1, "kind": "port",
            "sample": f"one zdtfdr! batch (input!+preprocess!+convolve!+postprocess!+output!) at "
                      f"2^20x16x129 by the numpy/scipy FP32 restatement, {dt:.2f} s",
            "all_cores": {"value": NF * NR / dt_all, "cores": ncpu,
                          "sample": f"the same batch with scipy.fft on {ncpu} threads (broadcasts single-threaded), "
                                    f"{dt_all:.2f} s"}}
    from fractions import Fraction
    ref = Z.fdr64(s, None, rates=[Fraction(R0) + Fraction(DR) * j for j in ACC_COLS])
    sig = H.fdrstats(ref)[1]
    acc = {"basis": f"FP64 oracle (exact phases) on rate columns {list(ACC_COLS)}; normwise = max|d|/max|fdr64|, "
                    f"cell = |d|/max(|fdr64|, s) with s = {sig:.3f} the columns' minimum std; vs_ref32 = normwise "
                    "against the FP32 restatement of the reference (Float32 phase arguments) on the same columns"}
    r32 = ref32[:, list(ACC_COLS)].astype(np.float64)
    for name, cols in gpu_cols.items():
        d = np.abs(cols.astype(np.float64) - ref)
        cell = d / np.maximum(np.abs(ref), sig)
        acc[name] = {"normwise": float(d.max() / np.abs(ref).max()), "cell_max": float(cell.max()),
                     "cell_mean": float(cell.mean()),
                     "vs_ref32": float(np.abs(cols - r32).max() / np.abs(r32).max())}
    return base, acc


def ragged_channels(nch: int, seed: int, dev):
    """The ragged workload's channels, generated on the device: Nf drawn from
    {2^18, 2^19, 2^20}, Nt from {16, 32, 60}, Exponential(1) noise plus 8
    tones (start channel uniform, rate U(-4, 4), amplitude U(10, 50)) each,
    rates -4:1/Nt:4.  Returns the shapes, the (Nf, Nt) column-major device
    spectrograms and the (r0, dr, Nr) rate ranges."""
    import torch
    rng = np.random.default_rng(seed)
    shapes = [(int(rng.choice([1 << 18, 1 << 19, 1 << 20])), int(rng.choice([16, 32, 60]))) for _ in range(nch)]
    gen = torch.Generator(device=dev).manual_seed(seed)
    specs, rates = [], []
    for Nf, Nt in shapes:
        sp = torch.empty((Nt, Nf), device=dev).exponential_(1.0, generator=gen)  # row t = time step
        c0 = rng.uniform(0, Nf, 8)
        rate = rng.uniform(-4, 4, 8)
        amp = rng.uniform(10, 50, 8)
        t = np.arange(Nt)
        ch = (np.round(c0[:, None] + rate[:, None] * t[None, :]).astype(np.int64) % Nf).ravel()
        tt = np.broadcast_to(t, (8, Nt)).ravel()
        sp[torch.from_numpy(tt).to(dev), torch.from_numpy(ch).to(dev)] += torch.from_numpy(
            np.repeat(amp, Nt).astype(np.float32)).to(dev)
        specs.append(sp.t())
        rates.append((-4.0, 1.0 / Nt, 8 * Nt + 1))
    return shapes, specs, rates


def ragged(a) -> int:
    """Many coarse channels in one plan launch: 64 channels, Nf drawn from
    {2^18, 2^19, 2^20}, Nt from {16, 32, 60}, Exponential(1) noise + 8 tones
    each, rates -4:1/Nt:4, proto-hits at SNR 10 compacted on the device.
    Prints one JSON line (aggregate FDR cells/s over all channels)."""
    import torch

    import fdrt_hip as F
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    nch = a.channels
    shapes, specs, rates = ragged_channels(nch, 4321, dev)
    t0 = time.perf_counter()
    plan = F.RaggedPlan(specs, rates, autotune=not a.no_autotune)
    plan.bind(10.0, snr=True, cap=1 << 16, keep_fdrs=False)
    torch.cuda.synchronize()
    t_plan = time.perf_counter() - t0
    for _ in range(a.warmup):
        plan.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        plan.launch()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    cells = sum(Nf * (8 * Nt + 1) for Nf, Nt in shapes)
    nh = plan.nhits()
    choice = {}
    for c, (Nf, Nt) in enumerate(shapes):
        i = plan.info(c)
        choice[f"{Nf}x{Nt}"] = i["path"] + (f" n2={i['n2']}" if i["path"] == "fused" else f" maxr={i['maxr']}")
    print(json.dumps({
        "metric": "FDR cells/s, ragged channels in one hipGraph launch",
        "value": cells / dt, "unit": "FDR cells/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "dtype": "f32", "data": "synthetic (seed 4321)",
        "config": {"workload": f"{nch} channels, Nf in {{2^18,2^19,2^20}}, Nt in {{16,32,60}}, rates -4:1/Nt:4, "
                               "SNR-10 proto-hits on device", "channels": nch, "cells": cells,
                   "shapes": choice, "plan_seconds": t_plan},
        "hits": {"total": int(nh.sum()), "max_per_channel": int(nh.max())},
    }))
    plan.close()
    return 0


def long_integration(a) -> int:
    """Long-integration search: spectrograms of Nf=2^22 x Nt=64 (Exponential(1)
    noise + 64 tones, amplitudes log-uniform 5..100) streamed from pinned host
    memory, rates -8:1/64:8 (Nr=1025) in HBM-sized batches, SNR-10 proto-hits.
    Prints one JSON line (FDR cells/s over the whole search, host copies
    included).  With --hits the search ends in hits (StreamSearch.findhits:
    each spectrogram's proto-hits clustered, 8-neighbourhood, on the device)."""
    import torch

    import fdrt_hip as F
    torch.cuda.set_device(0)
    Nf, Nt, r0, dr, Nr = 1 << 22, 64, -8.0, 1.0 / 64, 1025

    def run(stream, ss):
        if a.hits:
            h, nph, t = stream.findhits(ss, snr=10.0, cap=1 << 18)
            return [len(x.size) for x in h], [int(v) for v in nph]
        h, t = stream.search(ss, snr=10.0)
        return [len(x) for x in h], None
    rng = np.random.default_rng(2024)
    specs = []
    for _ in range(a.spectrograms):
        p = F.pinned_spectrogram(Nf, Nt)
        rows = rng.exponential(1.0, size=(Nt, Nf)).astype(np.float32)
        c0, rate = rng.uniform(0, Nf, 64), rng.uniform(-8, 8, 64)
        amp = np.exp(rng.uniform(np.log(5), np.log(100), 64)).astype(np.float32)
        for t in range(Nt):
            np.add.at(rows[t], (np.round(c0 + rate * t).astype(np.int64) % Nf), amp)
        p.t().copy_(torch.from_numpy(rows))
        specs.append(p)
    st = F.StreamSearch(Nf, Nt, (r0, dr, Nr), precision=a.precision)
    run(st, specs[:1])  # warm-up
    # one spectrogram alone: its host copy is not overlapped with anything
    t0 = time.perf_counter()
    run(st, specs[:1])
    dt1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    nhits, nproto = run(st, specs)
    dt = time.perf_counter() - t0
    n = a.spectrograms
    cells = n * Nf * Nr
    # each further spectrogram's copy overlaps the search of the one before:
    # (T(n) - T(1)) / (n - 1) is the per-spectrogram time of a long stream
    steady = (dt - dt1) / (n - 1) if n > 1 else None
    info = {"Nrb": st.Nrb, "nbatch": st.nbatch, "resident": st.resident, "device_bytes": st.nbytes, "path": st.path}
    st.close()
    # the kernels of the same search, timed one by one (HIP events around every
    # launch, hit extraction included) on a second search after the timed one
    sp = F.StreamSearch(Nf, Nt, (r0, dr, Nr), precision=a.precision, profile=True)
    run(sp, specs[:1])
    sp.reset_profile()
    nprof = min(2, n)
    run(sp, specs[:nprof])
    kt = sp.kernel_times()
    sp.close()
    ksum = sum(ms for _, ms, _ in kt) / nprof
    print(json.dumps({
        "metric": "FDR cells/s, long-integration search streamed from pinned host memory",
        "value": cells / dt, "unit": "FDR cells/s", "n_gpus": 1, "spectrograms": n,
        "ms_per_spectrogram": dt / n * 1e3, "higher_is_better": True, "dtype": "f32",
        "steady_ms_per_spectrogram": steady * 1e3 if steady else None,
        "steady_value": Nf * Nr / steady if steady else None, "single_spectrogram_ms": dt1 * 1e3,
        "kernel_sum_ms_per_spectrogram": ksum,
        "kernels": [{"name": k, "ms_per_spectrogram": ms / nprof, "launches_per_spectrogram": nl / nprof}
                    for k, ms, nl in kt],
        "data": "synthetic (seed 2024)",
        "config": {"workload": "Nf=2^22, Nt=64, rates -8:1/64:8 (Nr=1025), SNR 10"
                   + (", clustered into hits" if a.hits else ""), "precision": a.precision, **info},
        "hits" if a.hits else "protohits": nhits,
        **({"protohits": nproto} if a.hits else {}),
    }))
    return 0


def cluster_search(a) -> int:
    """Proto-hit clustering on the FDR of the 2^20 x 16 x 129 workload (the
    synthetic spectrogram through the fused path).  For three thresholds,
    SNR 8, SNR 3 and the mean (about half of the cells, one component spanning
    the matrix), the device proto-hits are made once, then
    zdt_clusterprotohits (8-neighbourhood) is timed over --steps calls; each
    call synchronizes once to size its outputs.  Prints one JSON line."""
    import torch

    import fdrt_hip as F
    from fdrt_hip import _lib
    from fdrt_hip import fdrutils as U
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    spec = torch.from_numpy(synthetic_spectrogram(NF, NT)).to(dev).t()
    ws = F.ZDTWorkspace(spec, R0, DR, NR)
    fdr = F.zdtfdr(ws)
    L = _lib.lib()
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rows = []
    for snr in (8.0, 3.0, 0.0):
        thr = F.fdrdenormalize(snr, fdr)
        buf, n, Nf, Nr, ld, arr = U._protohits_device([fdr], thr, 1 << 20)
        nc = C.c_int64()
        label = torch.empty(max(n, 1), dtype=torch.int64, device=dev)
        rc = L.zdt_clusterprotohits(buf.data_ptr(), n, arr, 1, Nf, Nr, ld, 1, 1, None, None, None, None, 0,
                                    C.byref(nc), label.data_ptr(), st)
        if rc not in (_lib.ZDT_OK, _lib.ZDT_ETRUNC):
            _lib.check(rc)
        cap = max(int(nc.value), 1)
        peak = torch.empty(cap, dtype=torch.int64, device=dev)
        value = torch.empty(cap, dtype=torch.float32, device=dev)
        size = torch.empty(cap, dtype=torch.int64, device=dev)
        box = torch.empty(4 * cap, dtype=torch.int64, device=dev)

        def call():
            _lib.check(L.zdt_clusterprotohits(buf.data_ptr(), n, arr, 1, Nf, Nr, ld, 1, 1, peak.data_ptr(),
                                              value.data_ptr(), size.data_ptr(), box.data_ptr(), cap,
                                              C.byref(nc), label.data_ptr(), st))

        for _ in range(a.warmup):
            call()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            call()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / a.steps
        rows.append({"snr": snr, "threshold": thr, "protohits": n, "hits": int(nc.value), "ms_per_call": dt * 1e3,
                     "protohits_per_s": n / dt})
        del buf, label, peak, value, size, box
    print(json.dumps({
        "metric": "proto-hits clustered per second (zdt_clusterprotohits, threshold = mean)",
        "value": rows[-1]["protohits_per_s"], "unit": "proto-hits/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "higher_is_better": True, "dtype": "int32/int64 (union-find), f32 values",
        "data": "synthetic (seed 1234)",
        "config": {"workload": "FDR of Nf=2^20, Nt=16, rates -4:1/16:4 (Nr=129), 8-neighbourhood"},
        "thresholds": rows,
    }))
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 200; --ragged 20)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps before them (default 100; --ragged 5)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--generic", action="store_true", help="force the unfused kernels")
    ap.add_argument("--fp32-only", action="store_true", help="skip the split-precision modes")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (to scale rocprof byte counters)")
    ap.add_argument("--valu-calibration", action="store_true",
                    help="only measure FP32 VALU issue rates (instruction forms x independent chains)")
    ap.add_argument("--lds-calibration", action="store_true",
                    help="only measure 8-byte LDS access costs of the FFT row layouts and strides")
    ap.add_argument("--ragged", action="store_true", help="the ragged multi-channel plan workload instead")
    ap.add_argument("--channels", type=int, default=64)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--long", action="store_true", help="the long-integration streaming search instead")
    ap.add_argument("--hits", action="store_true", help="--long: end the search in hits (proto-hits clustered)")
    ap.add_argument("--spectrograms", type=int, default=6)
    ap.add_argument("--precision", default="fp32", choices=("fp32", "split-bf16", "split-fp16"),
                    help="--long: the column DFTs' arithmetic")
    ap.add_argument("--cluster", action="store_true", help="proto-hit clustering on the 2^20 x 129 FDR instead")
    a = ap.parse_args(argv)
    if a.steps is None:
        a.steps = 20 if (a.ragged or a.cluster) else 200
    if a.warmup is None:
        a.warmup = 5 if (a.ragged or a.cluster) else 100
    if a.copy_calibration:
        return copy_calibration()
    if a.valu_calibration:
        return valu_calibration()
    if a.lds_calibration:
        return lds_calibration()
    if a.cluster:
        if a.gpus != 1:
            print("bench.py measures one GPU (N=1 only)", file=sys.stderr)
            return 2
        return cluster_search(a)
    if a.long:
        if a.gpus != 1:
            print("bench.py measures one GPU (N=1 only)", file=sys.stderr)
            return 2
        return long_integration(a)
    if a.ragged:
        if a.gpus != 1:
            print("bench.py measures one GPU (N=1 only)", file=sys.stderr)
            return 2
        return ragged(a)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU (N=1 only); the path shards across GPUs by "
              "spectrogram and by drift-rate batch, which one GPU's throughput already shows",
              file=sys.stderr)
        return 2

    import torch

    import fdrt_hip as F
    from fdrt_hip import _lib

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    spec_rows = torch.from_numpy(synthetic_spectrogram(NF, NT)).to(dev)
    spec = spec_rows.t()  # (Nf, Nt) column-major view
    base = _lib.ZDT_FLAG_GENERIC if a.generic else 0
    # workspaces through the C-ABI: one plain (the timed steps), one with the
    # profiling hook (per-kernel HIP events, for the roofline)
    L = _lib.lib()
    arr, n = _lib.int_array((2, 3, 5))
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)

    dest = F.create_fdr(spec, NR)
    total = spec.double().sum()

    def measure(precision: str) -> dict:
        """Timed steps of one precision mode: per-kernel durations from a
        workspace with the profiling hook (an event pair around every
        launch serializes the queue, ~10 us each), then the timed steps on a
        workspace without it.  Returns the timings, the kernel path the
        library chose and the FDR columns the accuracy figures use."""
        pc = _lib.PRECISIONS[precision]

        def create(flags):
            h = C.c_void_p()
            _lib.check(L.zdt_create(spec.data_ptr(), NF, NT, R0, DR, NR, arr, n, flags, pc, stream, C.byref(h)))
            return h

        hp, h = create(base), create(base | _lib.ZDT_FLAG_PROFILE)

        def step(w):
            _lib.check(L.zdt_input(w, spec.data_ptr()))
            _lib.check(L.zdt_run(w, R0, _lib.ZDT_WINDOW_RECT, None, dest.data_ptr()))

        for _ in range(a.warmup):
            step(h)
        torch.cuda.synchronize()
        _lib.check(L.zdt_profile_reset(h))
        t0 = time.perf_counter()
        for _ in range(a.steps):
            step(h)
        torch.cuda.synchronize()
        dt_prof = (time.perf_counter() - t0) / a.steps
        for _ in range(a.warmup):
            step(hp)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            step(hp)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / a.steps
        cnt = C.c_int()
        _lib.check(L.zdt_profile_read(h, -1, C.byref(cnt), None, 0, None, None, None))
        kern = []
        for i in range(cnt.value):
            name = C.create_string_buffer(64)
            ms = C.c_double()
            nl = C.c_int64()
            by = C.c_double()
            _lib.check(L.zdt_profile_read(h, i, None, name, 64, C.byref(ms), C.byref(nl), C.byref(by)))
            kern.append((name.value.decode(), ms.value, nl.value, by.value))
        kern.sort(key=lambda x: -x[1])
        path = C.c_int()
        _lib.check(L.zdt_path(hp, C.byref(path), None))
        L.zdt_destroy(h)
        L.zdt_destroy(hp)
        # sanity: the result is the drift-line sum (column sums equal the total power)
        colsum = dest.double().sum(dim=0)
        ok = bool(torch.allclose(colsum, total.expand_as(colsum), rtol=1e-5))
        cols = dest[:, list(ACC_COLS)].cpu().numpy()
        return {"dt": dt, "dt_prof": dt_prof, "kern": kern, "path": _lib.PATHS[path.value], "ok": ok, "cols": cols}

    runs = {"fp32": measure("fp32")}
    if not a.generic and not a.fp32_only:
        for prec in ("split-bf16", "split-fp16"):
            runs[prec] = measure(prec)
    if not all(r["ok"] for r in runs.values()):
        print("bench.py: result check failed", file=sys.stderr)
        return 1
    r = runs["fp32"]
    dt, dt_prof, kern = r["dt"], r["dt_prof"], r["kern"]
    dname, dms, dn, dbytes = kern[0]
    avg_s = dms / dn / 1e3
    achieved = dbytes / avg_s / 1e9

    # the VALU-bound kernel against the FP32 vector peak (algorithmic flops)
    compute = None
    k2_us = [ms / n_ * 1e3 for k, ms, n_, by in kern if k == "k2_czt"]
    if k2_us:
        tf = k2_flops(NF, NT, NR) / (k2_us[0] * 1e-6) / 1e12
        compute = {"kernel": "k2_czt", "bound": "valu", "achieved": tf, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                   "frac": tf / FP32_PEAK_TFLOPS, "flops_per_launch": k2_flops(NF, NT, NR),
                   "basis": "chirp-Z counted as complex multiply-adds (8 flops per term and rate), FFTs as "
                            "5 N log2 N; the FP32 vector peak, which non-packed VOP2 FMAs reach on gfx950"}
    out = {
        "metric": "FDR cells/s (Nf\u00d7Nrates) at 2^20ch\u00d716t\u00d7129 rates, split-BF16 vs FP32 ZDT, %HBM",
        "value": NF * NR / dt,
        "unit": "FDR cells/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Exponential(1) noise + 16 drifting tones, seed 1234)",
        "config": {"workload": "zdtfdr! input+1 batch, Nf=2^20, Nt=16, rates -4:1/16:4 (Nr=129, Nl=144)",
                   "Nf": NF, "Nt": NT, "Nr": NR, "r0": R0, "dr": DR, "path": r["path"], "precision": "fp32"},
        "roofline": {"bound": "hbm", "kernel": dname, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profiles(dname),
                     "avg_launch_us": avg_s * 1e6, "alg_bytes_per_launch": dbytes,
                     "timing": f"HIP events around each launch on the workspace stream, {a.steps} steps "
                               f"right before the timed ones ({dt_prof * 1e3:.4f} ms/step with the events)"},
        "compute": compute,
        "step_hbm": step_hbm(NF, NT, NR, dt),
        "kernels": [{"name": k, "ms_per_step": ms / a.steps, "launches_per_step": n_ / a.steps,
                     "alg_GBps": (by / (ms / n_ / 1e3) / 1e9) if ms > 0 else None} for k, ms, n_, by in kern],
    }
    if len(runs) > 1:
        # the split-precision modes on the same workload: the column FFTs of
        # input! and output! (k1, k3) as BF16/FP16 matrix-core products
        out["precision_modes"] = {
            prec: {"ms_per_step": q["dt"] * 1e3, "value": NF * NR / q["dt"], "speedup_vs_fp32": dt / q["dt"],
                   "path": q["path"],
                   "kernels": [{"name": k, "ms_per_step": ms / a.steps} for k, ms, n_, by in q["kern"]]}
            for prec, q in runs.items() if prec != "fp32"}
    if not a.no_cpu_baseline:
        out["cpu_baseline"], out["accuracy"] = cpu_baseline({p: q["cols"] for p, q in runs.items()})
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
