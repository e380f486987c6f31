"""The host-side C-ABI checks of test_abi.py (symbol exports, sizing against
the oracle, argument validation, the batchrates mirror) run again in the GPU
suite, so the GPU run covers the whole boundary: calcNl/growNr/
estimate_memory (src/zdtutils.jl:7-53) and batchrates (src/batchrates.jl:22-60).
"""
import pytest

from test_abi import *  # noqa: F401,F403  (the same test functions and fixture)

pytestmark = pytest.mark.gpu
