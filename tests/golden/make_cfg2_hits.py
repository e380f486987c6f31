"""Write cfg2_hits_snr8.json: the FP64 oracle's proto-hit set for the
headline workload, so the GPU test can compare the whole hit set without
running the 2^20 x 129 FP64 transform on the GPU box.

Workload: ``oracle.synth.cfg2(seed=1234)`` (2^20 channels x 16 steps,
Exponential(1) noise + 16 tones), rates -4:1/16:4 (Nr = 129).  The oracle is
``oracle.zdt_oracle.fdr64`` (exact-phase FP64 definition); the statistics are
``oracle.host_oracle.fdrstats`` (src/fdrutils.jl:39-50) and the hits
``findall(>=(thr))`` at thr = 8 s + m (src/fdrutils.jl:104-107,148-154).

Stored: thr64 and its (m, s); every hit as 1-based (i, j) in column-major
order; every cell with |fdr64 - thr64| <= BAND as (i, j, fdr64 - thr64), so
the test can form the tie band for its own measured epsilon (which must not
exceed BAND); the FDR's max |value| (normwise scale); and, for the tone rates
nearest each injected tone, nothing else -- the test recomputes sampled
columns with fdr64 itself.

Run (from the repository root): python tests/golden/make_cfg2_hits.py
(about a minute and ~6 GB of host memory).
"""
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import host_oracle as H  # noqa: E402
from oracle import synth  # noqa: E402
from oracle import zdt_oracle as Z  # noqa: E402

SEED, SNR = 1234, 8.0
R0, DR, NR = -4.0, 0.0625, 129
BAND = 0.07  # absolute; max |FDR| is ~633: normwise 1e-4 (the split-precision bar) of it is 0.063


def main() -> int:
    s = synth.cfg2(seed=SEED)
    Nf, Nt = s.shape
    ref = Z.fdr64(s, R0, DR, NR)
    m, sd = H.fdrstats(ref)
    thr = H.fdrdenormalize(SNR, m, sd)
    hits = H.findprotohits(ref, thr)
    d = ref - thr
    bi, bj = np.nonzero(np.abs(d) <= BAND)
    order = np.lexsort((bi, bj))  # column-major
    band = [[int(bi[q]) + 1, int(bj[q]) + 1, float(d[bi[q], bj[q]])] for q in order]
    out = {
        "workload": f"oracle.synth.cfg2(seed={SEED}): Nf={Nf}, Nt={Nt}, rates {R0}:{DR}:{-R0} (Nr={NR})",
        "oracle": "oracle.zdt_oracle.fdr64 (FP64, exact phases); oracle.host_oracle.fdrstats/findprotohits",
        "snr": SNR, "m": m, "s": sd, "thr64": thr, "band_abs": BAND,
        "max_abs_fdr64": float(np.max(np.abs(ref))),
        "colsum_fdr64": float(ref[:, 0].sum()),
        "hits": hits.tolist(),
        "band": band,
    }
    path = Path(__file__).with_name("cfg2_hits_snr8.json")
    path.write_text(json.dumps(out, separators=(",", ":")) + "\n")
    print(f"{path}: {len(hits)} hits, {len(band)} band cells, thr64={thr:.9g}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
