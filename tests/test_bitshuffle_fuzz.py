"""Memory safety of the bitshuffle chunk decoder (csrc/bitshuffle.cpp) on
damaged input, host only.

The decoder reads chunks straight from files, so a corrupt chunk must be
refused without a read or write outside the buffers.  tests/bshuf_fuzz.cpp
decodes valid LZ4-mode containers (built here with the helpers of
test_bitshuffle.py, with liblz4's compressor when the system has it, so that
blocks hold real matches) and then thousands of damaged copies of them.  It is
built with g++ under AddressSanitizer + UndefinedBehaviorSanitizer, and under
ThreadSanitizer for the threaded block decode.  A sanitizer report or a write
past the output fails the test.  The tests are skipped where g++ cannot build
with the sanitizer."""
import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from test_bitshuffle import LZ4, container, lz4, lz4_literals

HERE = Path(__file__).parent
CSRC = HERE.parent / "frequencydriftratetransforms.jl_amd" / "csrc"

CASES = [  # (element size, elements, block size in elements, 0 = default; real LZ4 matches or literals only)
    (4, 5005, 0, True), (4, 1003, 256, True), (2, 4099, 64, False), (1, 777, 128, True), (8, 600, 0, False),
    (4, 4096 * 3 + 3, 0, True),
]


def write_cases(path):
    rng = np.random.default_rng(5)
    with open(path, "wb") as f:
        f.write(np.int64(len(CASES)).tobytes())
        for E, n, bs, matches in CASES:
            if E == 4:
                raw = rng.exponential(1.0, n).astype(np.float32).tobytes()
            else:  # few distinct bytes, so that LZ4 finds matches
                raw = rng.integers(0, 4, n * E, dtype=np.uint8).tobytes()
            c = container(raw, E, bs, lz4 if (matches and LZ4 is not None) else lz4_literals)
            f.write(np.array([E, len(raw), len(c)], np.int64).tobytes())
            f.write(c)
            f.write(raw)


def build(tmp_path, sanitize):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    exe = tmp_path / f"fuzz_{sanitize.split(',')[0]}"
    cmd = [gxx, "-O1", "-g", "-std=c++17", f"-fsanitize={sanitize}", "-fno-sanitize-recover=all",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{CSRC}", str(HERE / "bshuf_fuzz.cpp"),
           str(CSRC / "bitshuffle.cpp"), "-o", str(exe), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip(f"g++ cannot build with -fsanitize={sanitize}: {r.stderr[-300:]}")
    return exe


@pytest.mark.parametrize("sanitize,iters", [("address,undefined", 30000), ("thread", 4000)])
def test_damaged_chunks_stay_in_bounds(tmp_path, sanitize, iters):
    exe = build(tmp_path, sanitize)
    cases = tmp_path / "cases.bin"
    write_cases(cases)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([str(exe), str(cases), str(iters), "32008"], capture_output=True, text=True, env=env,
                       timeout=300)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
    assert "Sanitizer" not in r.stderr, r.stderr[-3000:]
    acc, ref = (int(w) for w in r.stdout.split()[1::2])
    # both outcomes occur: damaged literals still decode, damaged structure is refused
    assert acc + ref == iters and acc > 0 and ref > 0
