"""fdrt_hip: Python host mirror of FrequencyDriftRateTransforms.jl's ZDT path
over libzdt_hip.so (MI355X / gfx950).

Names follow the reference's exports (src/FrequencyDriftRateTransforms.jl:3-24)
restricted to the ZDT path; a trailing underscore stands for Julia's ``!``.
The same C-ABI is bound from Julia by the ``julia/ZDTHip`` package and its
extension of FrequencyDriftRateTransforms.
"""
import torch  # noqa: F401  (loads the HIP runtime libzdt_hip.so binds to)

from ._lib import ZDTError, lib
from .fdrutils import (Hits, clusterprotohits, create_fdr, fdrdenormalize, fdrnormalize, fdrnormalize_,
                       fdrstats, findhits, findprotohits)
from .fftfdr import (FftFdrWorkspace, fdshift, fdshift_, fdshiftsum, fdshiftsum_, fftfdr, fftfdr_,
                     fftfdr_workspace, fftfdr_workspace_)
from .filterbank import FilterbankError, FilterbankFile, read_spectrogram, write_filterbank
from .plan import RaggedPlan, set_tune_table
from .stream import StreamSearch, pinned_spectrogram
from .zdt import (ZDTWorkspace, convolve_, fdrsynchronize, input_, output_, postprocess_,
                  preprocess_, zdtfdr, zdtfdr_)
from .zdtutils import (RateRange, batchrates, calcNl, driftrates, estimate_memory, growNr,
                       workspace_bytes)

__all__ = [
    "ZDTWorkspace", "input_", "output_", "preprocess_", "convolve_", "postprocess_",
    "zdtfdr", "zdtfdr_", "calcNl", "growNr", "estimate_memory", "driftrates", "batchrates",
    "RateRange", "workspace_bytes", "create_fdr", "fdrstats", "fdrsynchronize",
    "fdrnormalize_", "fdrnormalize", "fdrdenormalize", "findprotohits", "clusterprotohits", "findhits", "Hits", "ZDTError", "lib",
    "RaggedPlan", "set_tune_table", "FftFdrWorkspace", "fftfdr_workspace", "fftfdr_workspace_", "fftfdr", "fftfdr_",
    "fdshift", "fdshift_", "fdshiftsum", "fdshiftsum_", "StreamSearch", "pinned_spectrogram",
    "FilterbankFile", "FilterbankError", "read_spectrogram", "write_filterbank",
]
