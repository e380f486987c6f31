"""
    FrequencyDriftRateTransformsZDTHipExt

Loaded when both FrequencyDriftRateTransforms and ZDTHip are: adds the
package's own methods for `HipMatrix{Float32}` spectrograms and FDRs and for
the workspace they create, so user code written against the package runs on
MI355X unchanged:

    using FrequencyDriftRateTransforms, ZDTHip
    spec = HipMatrix(spectrogram)                 # Matrix{Float32} -> device
    ws   = ZDTWorkspace(spec, -4:1//16:4)         # a ZDTHip.HipZDTWorkspace
    fdr  = zdtfdr(ws)                              # HipMatrix{Float32}, Nf x Nr
    hits = findprotohits(fdr, 8; snr=true)         # Vector{CartesianIndex{2}}

It plays the part of the package's CUDA extension
(ext/CUDAFrequencyDriftRateTransformsExt.jl:1-28, which specialises
`plan_ffts!`, `output!` and `fdrsynchronize` on `CuArray`): here the whole
ZDT runs in libzdt_hip.so, so the entry points themselves get methods.  Every
method below is more specific than the package's generic one in the
workspace or array argument, which keeps dispatch unambiguous.  Pure host
functions (`calcNl`, `growNr`, `estimate_memory`, `batchrates`,
`fdrnormalize(::Number, m, s)`, `fdrdenormalize(snr, m, s)`) are the
package's own and are not touched.
"""
module FrequencyDriftRateTransformsZDTHipExt

import FrequencyDriftRateTransforms: ZDTWorkspace, input!, preprocess!, convolve!, postprocess!, output!,
                                     zdtfdr!, zdtfdr, driftrates, fdrstats, fdrnormalize!, findprotohits,
                                     fdrsynchronize, fftfdr_workspace, fftfdr_workspace!, fftfdr!, fftfdr,
                                     fdshift!, fdshift, fdshiftsum!, fdshiftsum
import FrequencyDriftRateTransforms: fdrdenormalize
using ZDTHip: ZDTHip, HipMatrix, HipZDTWorkspace, HipFdrs

const W = Union{Function,Symbol,Val}

# ---------------------------------------------------------------- workspace (src/zdtfdr.jl:37-87)

# ZDTWorkspace(spectrogram, r0, δr, Nr[, factors]; output_aligned) for a device
# spectrogram; the package's range form (src/zdtfdr.jl:80-87) forwards here.
# `precision` selects the arithmetic of the column DFTs (:fp32 by default).
function ZDTWorkspace(spectrogram::HipMatrix{Float32}, r0::Real, δr::Real, Nr::Integer,
                      factors::Union{Tuple,AbstractVector}=(2, 3, 5); output_aligned=false,
                      precision::Symbol=:fp32, unnormalized::Bool=false)
    HipZDTWorkspace(spectrogram, r0, δr, Nr, factors; precision, unnormalized)
end

"driftrates(ws, r0=ws.r0) -- src/zdtutils.jl:62-64"
driftrates(ws::HipZDTWorkspace, r0=ws.r0) = range(r0, step=ws.δr, length=ws.Nr)

# ---------------------------------------------------------------- step functions (src/zdtfdr.jl:209-331)

input!(ws::HipZDTWorkspace, spectrogram::HipMatrix{Float32}) = ZDTHip._input!(ws, spectrogram)

preprocess!(ws::HipZDTWorkspace, r0::Float32) = ZDTHip._preprocess!(ws, r0)
preprocess!(ws::HipZDTWorkspace, r0::Real) = ZDTHip._preprocess!(ws, Float32(r0))
preprocess!(ws::HipZDTWorkspace) = ZDTHip._preprocess!(ws, ws.r0)

convolve!(ws::HipZDTWorkspace) = ZDTHip._convolve!(ws)

postprocess!(w::Function, ws::HipZDTWorkspace) = ZDTHip._postprocess!(w, ws)
postprocess!(::Val{:hamming}, ws::HipZDTWorkspace) = ZDTHip._postprocess!(:hamming, ws)
postprocess!(::Val{:rect}, ws::HipZDTWorkspace) = ZDTHip._postprocess!(:rect, ws)
postprocess!(::Val{S}, ws::HipZDTWorkspace) where {S} = error("unsupported window type ($S)")
postprocess!(w::Symbol, ws::HipZDTWorkspace) = postprocess!(Val(w), ws)
postprocess!(ws::HipZDTWorkspace) = ZDTHip._postprocess!(:rect, ws)

output!(dest::HipMatrix{Float32}, ws::HipZDTWorkspace) = ZDTHip._output!(dest, ws)

# ---------------------------------------------------------------- zdtfdr!/zdtfdr (src/zdtfdr.jl:349-421)
# One fused batch per (dest, r0) instead of the step functions: the device
# never materialises Y unless a step function asks for it.

function zdtfdr!(w::W, dests, ws::HipZDTWorkspace, spectrogram=nothing; r0=ws.r0)
    spectrogram === nothing || input!(ws, spectrogram)
    for (dest, rate) in zip(dests, Iterators.cycle(r0))
        dest isa HipMatrix{Float32} && size(dest) == (ws.Nf, ws.Nr) || error("dest has unexpected type or size")
        ZDTHip._run!(ws, w, rate, dest.ptr)
    end
    dests
end
zdtfdr!(dests, ws::HipZDTWorkspace, spectrogram=nothing; r0=ws.r0) = zdtfdr!(Val(:rect), dests, ws, spectrogram; r0)

function zdtfdr!(w::W, dest::AbstractMatrix{<:Real}, ws::HipZDTWorkspace, spectrogram=nothing; r0::Real=ws.r0)
    zdtfdr!(w, (dest,), ws, spectrogram; r0)
    dest
end
zdtfdr!(dest::AbstractMatrix{<:Real}, ws::HipZDTWorkspace, spectrogram=nothing; r0::Real=ws.r0) =
    zdtfdr!(Val(:rect), dest, ws, spectrogram; r0)

# no dest: the spectra stay in the workspace for output!
function zdtfdr!(w::W, ws::HipZDTWorkspace, spectrogram=nothing; r0::Real=ws.r0)
    spectrogram === nothing || input!(ws, spectrogram)
    ZDTHip._run!(ws, w, r0, Ptr{Float32}(C_NULL))
    nothing
end
zdtfdr!(ws::HipZDTWorkspace, spectrogram=nothing; r0::Real=ws.r0) = zdtfdr!(Val(:rect), ws, spectrogram; r0)

function zdtfdr(w::W, ws::HipZDTWorkspace, spectrogram=nothing; r0::Real=ws.r0)
    zdtfdr!(w, HipMatrix{Float32}(undef, ws.Nf, ws.Nr), ws, spectrogram; r0)
end
zdtfdr(ws::HipZDTWorkspace, spectrogram=nothing; r0::Real=ws.r0) = zdtfdr(Val(:rect), ws, spectrogram; r0)

# ---------------------------------------------------------------- statistics and hits (src/fdrutils.jl)
# create_fdr needs nothing here: the package's method calls similar(), which
# HipMatrix implements.  fdrdenormalize(snr, fdr[s]) reaches fdrstats below.

fdrstats(fdr::HipMatrix{Float32}) = ZDTHip._stats((fdr,))
fdrstats(fdrs::HipFdrs) = ZDTHip._stats(fdrs)

fdrnormalize!(fdr::HipMatrix{Float32}, m, s) = (ZDTHip._normalize!((fdr,), m, s); fdr)
fdrnormalize!(fdr::HipMatrix{Float32}) = fdrnormalize!(fdr, fdrstats(fdr)...)
fdrnormalize!(fdrs::HipFdrs, m, s) = ZDTHip._normalize!(fdrs, m, s)
fdrnormalize!(fdrs::HipFdrs) = fdrnormalize!(fdrs, fdrstats(fdrs)...)

function findprotohits(fdr::HipMatrix{Float32}, threshold; snr::Bool=false)
    findprotohits((fdr,), threshold; snr)
end
function findprotohits(fdrs::HipFdrs, threshold; snr::Bool=false)
    snr && (threshold = fdrdenormalize(threshold, fdrs))
    ZDTHip._hits(fdrs, threshold)
end

"fdrsynchronize -- ext/CUDAFrequencyDriftRateTransformsExt.jl:26-28"
fdrsynchronize(::Type{<:HipMatrix}) = ZDTHip.device_synchronize()
fdrsynchronize(ws::HipZDTWorkspace) = ZDTHip._sync(ws)

# ---------------------------------------------------------------- Fourier-shift de-doppler (src/fftfdr.jl)
# any ZDT workspace holds the input spectrum these need

fftfdr_workspace(spectrogram::HipMatrix{Float32}; bunaligned=true) = HipZDTWorkspace(spectrogram, 0f0, 1f0, 1)
fftfdr_workspace!(ws::HipZDTWorkspace, spectrogram::HipMatrix{Float32}; bunaligned=true) = input!(ws, spectrogram)
fftfdr_workspace!(::Nothing, spectrogram::HipMatrix{Float32}; bunaligned=true) = fftfdr_workspace(spectrogram)

fftfdr!(fdr::HipMatrix{Float32}, ws::HipZDTWorkspace, rates) = ZDTHip._fftfdr!(fdr, ws, rates)
fftfdr(ws::HipZDTWorkspace, rates) = fftfdr!(HipMatrix{Float32}(undef, ws.Nf, length(rates)), ws, rates)

fdshiftsum!(dest::HipMatrix{Float32}, ws::HipZDTWorkspace, rate) = fftfdr!(dest, ws, (rate,))
fdshiftsum(ws::HipZDTWorkspace, rate) = fftfdr(ws, (rate,))

fdshift!(dest::HipMatrix{Float32}, ws::HipZDTWorkspace, rate) = ZDTHip._fdshift!(dest, ws, rate)
fdshift(ws::HipZDTWorkspace, rate) = fdshift!(HipMatrix{Float32}(undef, ws.Nf, ws.Nt), ws, rate)

end # module FrequencyDriftRateTransformsZDTHipExt
