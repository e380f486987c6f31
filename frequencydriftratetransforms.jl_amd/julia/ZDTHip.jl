"""
    ZDTHip

Julia host module for libzdt_hip.so, the MI355X (gfx950) implementation of the
Chirp-Z De-Doppler Transform path of FrequencyDriftRateTransforms.jl.  It keeps
the package's call signatures (`ZDTWorkspace`, `input!`, `preprocess!`,
`convolve!`, `postprocess!`, `output!`, `zdtfdr!`, `zdtfdr`, `findprotohits`,
`fdrstats`, `fdrnormalize!`, `fdrdenormalize`, `fdrsynchronize`, `calcNl`,
`growNr`, `estimate_memory`, `driftrates`, `batchrates`) and reaches the
device through `ccall` only: device memory is a small `HipMatrix` handle
(pointer + size) managed through the HIP runtime's C API, so no GPU package is
needed.  Every library call returns a status; a non-zero status throws
`ErrorException(zdt_last_error())`.

Reference interfaces replaced: src/zdtfdr.jl:10-421, src/zdtutils.jl:7-64,
src/batchrates.jl:22-60, src/fdrutils.jl:11-169,
ext/CUDAFrequencyDriftRateTransformsExt.jl:26-103.
"""
module ZDTHip

export HipMatrix, ZDTWorkspace, input!, output!, preprocess!, convolve!, postprocess!,
       zdtfdr, zdtfdr!, calcNl, growNr, estimate_memory, driftrates, batchrates,
       create_fdr, fdrstats, fdrsynchronize, fdrnormalize!, fdrnormalize, fdrdenormalize,
       findprotohits, RaggedPlan, bind!, launch!, planhits,
       fftfdr_workspace, fftfdr_workspace!, fftfdr, fftfdr!, fdshift, fdshift!, fdshiftsum, fdshiftsum!,
       StreamSearch, search

const LIBZDT = get(ENV, "ZDT_LIB", joinpath(@__DIR__, "..", "lib", "libzdt_hip.so"))
const LIBHIP = "libamdhip64.so"

const ZDT_OK = 0
const ZDT_ETRUNC = 5
const WINDOW_RECT = 0
const WINDOW_HAMMING = 1
const WINDOW_CUSTOM = 2
const FLAG_UNNORMALIZED = 0x1

zdt_error() = unsafe_string(ccall((:zdt_last_error, LIBZDT), Cstring, ()))
check(rc::Integer) = rc == ZDT_OK ? nothing : error(zdt_error())

# ----------------------------------------------------------------- device memory

"""
    HipMatrix{T}(undef, m, n) / HipMatrix(A::Matrix)

Column-major device matrix: the stand-in for a `CuMatrix`.  `Array(H)` copies
back to the host.  Memory is released by a finalizer (`hipFree`).
"""
mutable struct HipMatrix{T} <: AbstractMatrix{T}
    ptr::Ptr{T}
    dims::NTuple{2,Int}
    function HipMatrix{T}(::UndefInitializer, m::Integer, n::Integer) where {T}
        p = Ref{Ptr{Cvoid}}(C_NULL)
        nbytes = max(m * n * sizeof(T), 1)
        rc = ccall((:hipMalloc, LIBHIP), Cint, (Ref{Ptr{Cvoid}}, Csize_t), p, nbytes)
        rc == 0 || error("hipMalloc($nbytes) failed ($rc)")
        H = new{T}(Ptr{T}(p[]), (Int(m), Int(n)))
        finalizer(H) do h
            ccall((:hipFree, LIBHIP), Cint, (Ptr{Cvoid},), h.ptr)
        end
        return H
    end
end

function HipMatrix(A::AbstractMatrix{T}) where {T}
    H = HipMatrix{T}(undef, size(A)...)
    B = Matrix{T}(A)
    rc = ccall((:hipMemcpy, LIBHIP), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Cint),
               H.ptr, B, sizeof(B), 1)  # hipMemcpyHostToDevice
    rc == 0 || error("hipMemcpy H2D failed ($rc)")
    return H
end

Base.size(H::HipMatrix) = H.dims
Base.getindex(::HipMatrix, ::Int...) = error("HipMatrix elements live on the device; use Array(H)")
Base.pointer(H::HipMatrix) = H.ptr

function Base.Array(H::HipMatrix{T}) where {T}
    A = Matrix{T}(undef, H.dims...)
    rc = ccall((:hipMemcpy, LIBHIP), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Cint),
               A, H.ptr, sizeof(A), 2)  # hipMemcpyDeviceToHost
    rc == 0 || error("hipMemcpy D2H failed ($rc)")
    return A
end

# ----------------------------------------------------------------- sizing (src/zdtutils.jl)

_factors(f) = Cint[Int(x) for x in f]

"calcNl(Nt, Nr, factors=(2,3,5)): nextprod(factors, Nt+Nr-1) -- src/zdtutils.jl:7-9"
function calcNl(Nt, Nr, factors=(2, 3, 5))
    f = _factors(factors)
    v = ccall((:zdt_calc_nl, LIBZDT), Int64, (Int64, Int64, Ptr{Cint}, Cint), Nt, Nr, f, length(f))
    v < 0 && error("calcNl: invalid arguments")
    return Int(v)
end

"growNr(Nt, Nr, factors=(2,3,5)) -- src/zdtutils.jl:17-19"
growNr(Nt, Nr, factors=(2, 3, 5)) = calcNl(Nt, Nr, factors) - Nt + 1

"estimate_memory(Nf, Nt, Nr, Ni=1, No=1; factors) -- src/zdtutils.jl:35-53"
function estimate_memory(Nf, Nt, Nr, Ni=1, No=1; factors=(2, 3, 5))
    f = _factors(factors)
    ccall((:zdt_estimate_memory, LIBZDT), Int64,
          (Int64, Int64, Int64, Int64, Int64, Ptr{Cint}, Cint), Nf, Nt, Nr, Ni, No, f, length(f))
end

"batchrates(Nt, δhzps, r1, r2=-r1; Nrb=360) -- src/batchrates.jl:22-60"
function batchrates(Nt::Integer, δhzps, r1, r2=-r1; Nrb=360)
    if r2 < r1
        r1, r2 = r2, r1
    end
    nc1 = round(Int, r1 / δhzps)
    nc2 = round(Int, r2 / δhzps)
    Nrspan = abs(nc2 - nc1) + 1
    Nrb = growNr(Nt, min(Nrb, Nrspan))
    Nr = cld(Nrspan, Nrb) * Nrb
    Nrbonus = Nr - Nrspan
    nc1 -= round(Int, sign(δhzps) * Nrbonus / 2, RoundFromZero)
    nc2 += round(Int, sign(δhzps) * Nrbonus / 2, RoundToZero)
    rs = range(nc1 // (Nt - 1), stop=nc2 // (Nt - 1), length=Nr)
    @assert length(rs) % Nrb == 0
    collect(Iterators.partition(rs, Nrb))
end

# ----------------------------------------------------------------- workspace

"""
    ZDTWorkspace(spectrogram::HipMatrix{Float32}, rates[, factors]; unnormalized=false)
    ZDTWorkspace(spectrogram::HipMatrix{Float32}, r0, δr, Nr[, factors]; unnormalized=false)

src/zdtfdr.jl:37-87.  Allocates the device buffers and runs `input!`.
"""
mutable struct ZDTWorkspace
    handle::Ptr{Cvoid}
    Nf::Int
    Nt::Int
    r0::Float32
    δr::Float32
    Nr::Int
    Nl::Int
    factors::Tuple
    stream::Ptr{Cvoid}
    window::Union{Nothing,HipMatrix{Float32}}
end

function ZDTWorkspace(spec::HipMatrix{Float32}, r0::Real, δr::Real, Nr::Integer,
                      factors=(2, 3, 5); unnormalized::Bool=false, stream::Ptr{Cvoid}=C_NULL)
    Nf, Nt = size(spec)
    f = _factors(factors)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    flags = unnormalized ? FLAG_UNNORMALIZED : 0
    check(ccall((:zdt_create, LIBZDT), Cint,
                (Ptr{Float32}, Int64, Int64, Cfloat, Cfloat, Int64, Ptr{Cint}, Cint, Cint, Cint, Ptr{Cvoid},
                 Ref{Ptr{Cvoid}}),
                spec.ptr, Nf, Nt, Float32(r0), Float32(δr), Nr, f, length(f), flags, Cint(0), stream, h))
    nl = Ref{Int64}(0)
    check(ccall((:zdt_info, LIBZDT), Cint,
                (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Int64}, Ref{Int64}),
                h[], C_NULL, C_NULL, C_NULL, C_NULL, C_NULL, nl))
    ws = ZDTWorkspace(h[], Nf, Nt, Float32(r0), Float32(δr), Nr, Int(nl[]), Tuple(factors), stream, nothing)
    finalizer(ws) do w
        ccall((:zdt_destroy, LIBZDT), Cint, (Ptr{Cvoid},), w.handle)
    end
    return ws
end

function ZDTWorkspace(spec::HipMatrix{Float32}, rates::AbstractRange, factors=(2, 3, 5); kw...)
    ZDTWorkspace(spec, first(rates), step(rates), length(rates), factors; kw...)
end

function Base.show(io::IO, ws::ZDTWorkspace)
    print(io, "ZDTWorkspace(Nf=", ws.Nf, ",Nt=", ws.Nt, ",r0=", ws.r0, ",δr=", ws.δr,
          ",Nr=", ws.Nr, ",Nl=", ws.Nl, ",", ws.factors, ")")
end

function Base.sizeof(ws::ZDTWorkspace)
    b = Ref{Int64}(0)
    check(ccall((:zdt_sizeof, LIBZDT), Cint, (Ptr{Cvoid}, Ref{Int64}), ws.handle, b))
    Int(b[])
end

"driftrates(ws, r0=ws.r0) -- src/zdtutils.jl:62-64"
driftrates(ws::ZDTWorkspace, r0=ws.r0) = range(r0, step=ws.δr, length=ws.Nr)

"fdrsynchronize -- ext/CUDAFrequencyDriftRateTransformsExt.jl:26-28"
fdrsynchronize(ws::ZDTWorkspace) = check(ccall((:zdt_sync, LIBZDT), Cint, (Ptr{Cvoid},), ws.handle))
fdrsynchronize(::Type{<:HipMatrix}) = ccall((:hipDeviceSynchronize, LIBHIP), Cint, ())

# window -> (kind, device pointer); src/zdtfdr.jl:297-321
function _window(ws::ZDTWorkspace, w)
    w isa Val && (w = typeof(w).parameters[1])
    if w === :rect
        return WINDOW_RECT, Ptr{Float32}(C_NULL)
    elseif w === :hamming
        return WINDOW_HAMMING, Ptr{Float32}(C_NULL)
    elseif w isa Symbol
        error("unsupported window type ($w)")
    elseif w isa Function
        v = Float32[w(n, ws.Nf) for n in 0:ws.Nf÷2]
        ws.window = HipMatrix(reshape(v, :, 1))
        return WINDOW_CUSTOM, ws.window.ptr
    end
    error("unsupported window argument $(typeof(w))")
end

"input!(workspace, spectrogram) -- src/zdtfdr.jl:209-212"
function input!(ws::ZDTWorkspace, spec::HipMatrix{Float32})
    size(spec) == (ws.Nf, ws.Nt) || error("input spectrogram has unexpected size")
    check(ccall((:zdt_input, LIBZDT), Cint, (Ptr{Cvoid}, Ptr{Float32}), ws.handle, spec.ptr))
end

"preprocess!(workspace, r0=workspace.r0) -- src/zdtfdr.jl:221-239"
preprocess!(ws::ZDTWorkspace, r0::Real=ws.r0) =
    check(ccall((:zdt_preprocess, LIBZDT), Cint, (Ptr{Cvoid}, Cfloat), ws.handle, Float32(r0)))

"convolve!(workspace) -- src/zdtfdr.jl:249-253"
convolve!(ws::ZDTWorkspace) = check(ccall((:zdt_convolve, LIBZDT), Cint, (Ptr{Cvoid},), ws.handle))

"postprocess!([w,] workspace) -- src/zdtfdr.jl:297-321"
function postprocess!(w, ws::ZDTWorkspace)
    kind, ptr = _window(ws, w)
    check(ccall((:zdt_postprocess, LIBZDT), Cint, (Ptr{Cvoid}, Cint, Ptr{Float32}), ws.handle, kind, ptr))
end
postprocess!(ws::ZDTWorkspace) = postprocess!(:rect, ws)

"output!(dest, workspace) -- src/zdtfdr.jl:328-331"
function output!(dest::HipMatrix{Float32}, ws::ZDTWorkspace)
    size(dest) == (ws.Nf, ws.Nr) || error("dest has unexpected size")
    check(ccall((:zdt_output, LIBZDT), Cint, (Ptr{Cvoid}, Ptr{Float32}), ws.handle, dest.ptr))
    return dest
end

"zdtfdr!([w,] dests, workspace[, spectrogram]; r0) -- src/zdtfdr.jl:349-362"
function zdtfdr!(w::Union{Function,Symbol,Val}, dests, ws::ZDTWorkspace, spectrogram=nothing; r0=ws.r0)
    spectrogram === nothing || input!(ws, spectrogram)
    kind, ptr = _window(ws, w)
    for (dest, rate) in zip(dests, Iterators.cycle(r0))
        size(dest) == (ws.Nf, ws.Nr) || error("dest has unexpected size")
        check(ccall((:zdt_run, LIBZDT), Cint, (Ptr{Cvoid}, Cfloat, Cint, Ptr{Float32}, Ptr{Float32}),
                    ws.handle, Float32(rate), kind, ptr, dest.ptr))
    end
    return dests
end
zdtfdr!(dests, ws::ZDTWorkspace, spectrogram=nothing; r0=ws.r0) = zdtfdr!(:rect, dests, ws, spectrogram; r0)

function zdtfdr!(w::Union{Function,Symbol,Val}, dest::HipMatrix{Float32}, ws::ZDTWorkspace, spectrogram=nothing;
                 r0::Real=ws.r0)
    zdtfdr!(w, (dest,), ws, spectrogram; r0)
    return dest
end
zdtfdr!(dest::HipMatrix{Float32}, ws::ZDTWorkspace, spectrogram=nothing; r0::Real=ws.r0) =
    zdtfdr!(:rect, dest, ws, spectrogram; r0)

# no dest: the result stays in the workspace for output! (src/zdtfdr.jl:382-396)
function zdtfdr!(w::Union{Function,Symbol,Val}, ws::ZDTWorkspace, spectrogram=nothing; r0::Real=ws.r0)
    spectrogram === nothing || input!(ws, spectrogram)
    kind, ptr = _window(ws, w)
    check(ccall((:zdt_run, LIBZDT), Cint, (Ptr{Cvoid}, Cfloat, Cint, Ptr{Float32}, Ptr{Float32}),
                ws.handle, Float32(r0), kind, ptr, C_NULL))
    return nothing
end
zdtfdr!(ws::ZDTWorkspace, spectrogram=nothing; r0::Real=ws.r0) = zdtfdr!(:rect, ws, spectrogram; r0)

"zdtfdr([w,] workspace[, spectrogram]; r0) -- src/zdtfdr.jl:412-421"
function zdtfdr(w::Union{Function,Symbol,Val}, ws::ZDTWorkspace, spectrogram=nothing; r0::Real=ws.r0)
    dest = HipMatrix{Float32}(undef, ws.Nf, ws.Nr)
    zdtfdr!(w, dest, ws, spectrogram; r0)
end
zdtfdr(ws::ZDTWorkspace, spectrogram=nothing; r0::Real=ws.r0) = zdtfdr(:rect, ws, spectrogram; r0)

# ----------------------------------------------------------------- statistics (src/fdrutils.jl)

"create_fdr(spectrogram, Nr_or_rates) -- src/fdrutils.jl:11-26"
create_fdr(spec::HipMatrix{Float32}, Nr::Integer) = HipMatrix{Float32}(undef, size(spec, 1), Nr)
create_fdr(spec::HipMatrix{Float32}, rates) = create_fdr(spec, length(rates))

_ptrs(fdrs) = Ptr{Float32}[f.ptr for f in fdrs]

"fdrstats(fdr) / fdrstats(fdrs) -- src/fdrutils.jl:39-50"
fdrstats(fdr::HipMatrix{Float32}) = fdrstats((fdr,))
function fdrstats(fdrs)
    Nf, Nr = size(first(fdrs))
    m = Ref{Cdouble}(0)
    s = Ref{Cdouble}(0)
    check(ccall((:zdt_fdrstats, LIBZDT), Cint,
                (Ptr{Ptr{Float32}}, Cint, Int64, Int64, Int64, Ptr{Cvoid}, Ref{Cdouble}, Ref{Cdouble}),
                _ptrs(fdrs), length(fdrs), Nf, Nr, Nf, C_NULL, m, s))
    (m[], s[])
end

fdrnormalize(x::Number, m, s) = (x - m) / s

"fdrnormalize!(fdr[s][, m, s]) -- src/fdrutils.jl:65-90"
fdrnormalize!(fdr::HipMatrix{Float32}, m, s) = (fdrnormalize!((fdr,), m, s); fdr)
fdrnormalize!(fdr::HipMatrix{Float32}) = fdrnormalize!(fdr, fdrstats(fdr)...)
function fdrnormalize!(fdrs, m, s)
    Nf, Nr = size(first(fdrs))
    check(ccall((:zdt_fdrnormalize, LIBZDT), Cint,
                (Ptr{Ptr{Float32}}, Cint, Int64, Int64, Int64, Cdouble, Cdouble, Ptr{Cvoid}),
                _ptrs(fdrs), length(fdrs), Nf, Nr, Nf, m, s, C_NULL))
    fdrs
end
fdrnormalize!(fdrs) = fdrnormalize!(fdrs, fdrstats(fdrs)...)

"fdrdenormalize(snr, m, s) / fdrdenormalize(snr, fdr[s]) -- src/fdrutils.jl:104-119"
fdrdenormalize(snr, m, s) = snr * s + m
fdrdenormalize(snr, fdrs) = fdrdenormalize(snr, fdrstats(fdrs)...)

"findprotohits(fdr[s], threshold; snr=false) -- src/fdrutils.jl:148-169"
findprotohits(fdr::HipMatrix{Float32}, threshold; snr::Bool=false) = findprotohits((fdr,), threshold; snr)
function findprotohits(fdrs, threshold; snr::Bool=false)
    snr && (threshold = fdrdenormalize(threshold, fdrs))
    Nf, Nr = size(first(fdrs))
    n = Ref{Int64}(0)
    ptrs = _ptrs(fdrs)
    # count, then fetch the 0-based linear indices i + Nf*(b*Nr + j)
    rc = ccall((:zdt_findprotohits, LIBZDT), Cint,
               (Ptr{Ptr{Float32}}, Cint, Int64, Int64, Int64, Cdouble, Ptr{Int64}, Int64, Ref{Int64}, Ptr{Cvoid}),
               ptrs, length(fdrs), Nf, Nr, Nf, threshold, C_NULL, 0, n, C_NULL)
    rc in (ZDT_OK, ZDT_ETRUNC) || check(rc)
    H = HipMatrix{Int64}(undef, max(n[], 1), 1)
    check(ccall((:zdt_findprotohits, LIBZDT), Cint,
                (Ptr{Ptr{Float32}}, Cint, Int64, Int64, Int64, Cdouble, Ptr{Int64}, Int64, Ref{Int64}, Ptr{Cvoid}),
                ptrs, length(fdrs), Nf, Nr, Nf, threshold, H.ptr, n[], n, C_NULL))
    lin = n[] == 0 ? Int64[] : vec(Array(H))[1:n[]]
    [CartesianIndex(Int(l % Nf) + 1, Int(l ÷ Nf) + 1) for l in lin]
end

# ----------------------------------------------------------------- Fourier-shift de-doppler

"fftfdr_workspace(spectrogram) -- src/fftfdr.jl:39-59 (any ZDT workspace holds the input spectrum)"
fftfdr_workspace(spec::HipMatrix{Float32}; bunaligned=true) = ZDTWorkspace(spec, 0f0, 1f0, 1)

"fftfdr_workspace!(workspace, spectrogram) -- src/fftfdr.jl:69-84"
function fftfdr_workspace!(ws::ZDTWorkspace, spec::HipMatrix{Float32}; bunaligned=true)
    size(spec) == (ws.Nf, ws.Nt) || error("input spectrogram has unexpected size")
    input!(ws, spec)
    ws
end
fftfdr_workspace!(::Nothing, spec::HipMatrix{Float32}; bunaligned=true) = fftfdr_workspace(spec)

"fftfdr!(fdr, workspace, rates) -- src/fftfdr.jl:126-134"
function fftfdr!(fdr::HipMatrix{Float32}, ws::ZDTWorkspace, rates)
    r = Float32.(collect(rates))
    size(fdr) == (ws.Nf, length(r)) || error("fdr must be $(ws.Nf) x $(length(r))")
    check(ccall((:zdt_fftfdr, LIBZDT), Cint, (Ptr{Cvoid}, Ptr{Cfloat}, Int64, Ptr{Float32}),
                ws.handle, r, length(r), fdr.ptr))
    fdr
end

"fftfdr(workspace, rates) -- src/fftfdr.jl:143-148"
fftfdr(ws::ZDTWorkspace, rates) = fftfdr!(HipMatrix{Float32}(undef, ws.Nf, length(rates)), ws, rates)

"fdshiftsum!(dest, workspace, rate) / fdshiftsum(workspace, rate) -- src/fftfdr.jl:86-102"
fdshiftsum!(dest::HipMatrix{Float32}, ws::ZDTWorkspace, rate) = fftfdr!(dest, ws, (rate,))
fdshiftsum(ws::ZDTWorkspace, rate) = fftfdr(ws, (rate,))

"fdshift!(dest, workspace, rate) -- src/fftfdr.jl:104-111"
function fdshift!(dest::HipMatrix{Float32}, ws::ZDTWorkspace, rate)
    size(dest) == (ws.Nf, ws.Nt) || error("dest must be $(ws.Nf) x $(ws.Nt)")
    check(ccall((:zdt_fdshift, LIBZDT), Cint, (Ptr{Cvoid}, Cfloat, Ptr{Float32}), ws.handle, Float32(rate), dest.ptr))
    dest
end

"fdshift(workspace, rate) -- src/fftfdr.jl:113-118"
fdshift(ws::ZDTWorkspace, rate) = fdshift!(HipMatrix{Float32}(undef, ws.Nf, ws.Nt), ws, rate)

# ----------------------------------------------------------------- ragged plans

const PLAN_AUTOTUNE = 0x100

"""
    RaggedPlan(specs::Vector{HipMatrix{Float32}}, rates::Vector; autotune=true)

Many coarse channels of uneven Nf and Nt through `input!`, one `zdtfdr!`
batch, `fdrdenormalize` and `findprotohits`, replayed as one hipGraph launch
(include/zdt.h, zdt_plan_*).  `rates[c]` is a range (`r0:δr:r1`) or an
`(r0, δr, Nr)` tuple.  No counterpart in the reference: it replaces the
user's loop over channels around src/zdtfdr.jl:349-421 and
src/fdrutils.jl:148-169.
"""
mutable struct RaggedPlan
    handle::Ptr{Cvoid}
    specs::Vector{HipMatrix{Float32}}
    Nf::Vector{Int64}
    Nr::Vector{Int64}
    cap::Int
    hits::Union{Nothing,HipMatrix{Int64}}
    nhits::Union{Nothing,HipMatrix{Int64}}
    dests::Vector{HipMatrix{Float32}}
end

_rate_params(r::Tuple) = (Float32(r[1]), Float32(r[2]), Int64(r[3]))
_rate_params(r::AbstractRange) = (Float32(first(r)), Float32(step(r)), Int64(length(r)))

function RaggedPlan(specs::Vector{HipMatrix{Float32}}, rates::Vector; autotune::Bool=true,
                    unnormalized::Bool=false, stream::Ptr{Cvoid}=C_NULL)
    n = length(specs)
    n == length(rates) || error("need one rate range per spectrogram")
    p = map(_rate_params, rates)
    Nf = Int64[size(s, 1) for s in specs]
    Nt = Int64[size(s, 2) for s in specs]
    flags = (autotune ? PLAN_AUTOTUNE : 0) | (unnormalized ? FLAG_UNNORMALIZED : 0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:zdt_plan_ragged, LIBZDT), Cint,
                (Cint, Ptr{Int64}, Ptr{Int64}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Int64}, Cint, Ptr{Cvoid},
                 Ref{Ptr{Cvoid}}),
                n, Nf, Nt, Float32[x[1] for x in p], Float32[x[2] for x in p], Int64[x[3] for x in p],
                flags, stream, h))
    P = RaggedPlan(h[], specs, Nf, Int64[x[3] for x in p], 0, nothing, nothing, HipMatrix{Float32}[])
    finalizer(P) do q
        ccall((:zdt_plan_destroy, LIBZDT), Cint, (Ptr{Cvoid},), q.handle)
    end
    return P
end

"""
    bind!(plan, thresholds; snr=true, cap=4096)

Allocates one FDR per channel and the hit buffers, and captures the graph.
`thresholds` is one value or one per channel (SNRs when `snr`).
"""
function bind!(P::RaggedPlan, thresholds; snr::Bool=true, cap::Integer=4096)
    n = length(P.specs)
    thr = Float64.(thresholds isa Number ? fill(thresholds, n) : thresholds)
    P.dests = [HipMatrix{Float32}(undef, P.Nf[c], P.Nr[c]) for c in 1:n]
    P.cap = cap
    P.hits = HipMatrix{Int64}(undef, max(n * cap, 1), 1)
    P.nhits = HipMatrix{Int64}(undef, n, 1)
    check(ccall((:zdt_plan_bind, LIBZDT), Cint,
                (Ptr{Cvoid}, Ptr{Ptr{Float32}}, Ptr{Ptr{Float32}}, Ptr{Cdouble}, Cint, Ptr{Int64}, Int64,
                 Ptr{Int64}, Ptr{Cdouble}),
                P.handle, [s.ptr for s in P.specs], [d.ptr for d in P.dests], thr, snr, P.hits.ptr, cap,
                P.nhits.ptr, C_NULL))
    P
end

"launch!(plan): one hipGraphLaunch of the bound plan."
launch!(P::RaggedPlan) = (check(ccall((:zdt_plan_launch, LIBZDT), Cint, (Ptr{Cvoid},), P.handle)); P)

"planhits(plan, c): channel c's proto-hits as `CartesianIndex`es (like `findprotohits`)."
function planhits(P::RaggedPlan, c::Integer)
    n = vec(Array(P.nhits))[c]
    n <= P.cap || error("channel $c: $n hits exceed capacity $(P.cap)")
    lin = vec(Array(P.hits))[(c-1)*P.cap+1:(c-1)*P.cap+n]
    Nf = P.Nf[c]
    [CartesianIndex(Int(l % Nf) + 1, Int(l ÷ Nf) + 1) for l in lin]
end

# ----------------------------------------------------------------- long-integration search

"""
    StreamSearch(Nf, Nt, rates; Nrb=0, hbm_bytes=0)

Spectrograms of one shape streamed from host memory, rates cut into equal
HBM-sized batches (include/zdt.h, zdt_stream_*).  Replaces the loop around
`batchrates` (src/batchrates.jl:22-60), `zdtfdr!(dests, ws; r0)`
(src/zdtfdr.jl:349-362) and `findprotohits(fdrs, snr; snr=true)`
(src/fdrutils.jl:148-169).
"""
mutable struct StreamSearch
    handle::Ptr{Cvoid}
    Nf::Int
    Nt::Int
    Nrb::Int
    nbatch::Int
end

function StreamSearch(Nf::Integer, Nt::Integer, rates; Nrb::Integer=0, hbm_bytes::Integer=0)
    r0, δr, Nr = _rate_params(rates)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:zdt_stream_create, LIBZDT), Cint,
                (Int64, Int64, Cfloat, Cfloat, Int64, Int64, Int64, Cint, Ref{Ptr{Cvoid}}),
                Nf, Nt, r0, δr, Nr, Nrb, hbm_bytes, 0, h))
    nrb, nb = Ref{Int64}(0), Ref{Int64}(0)
    check(ccall((:zdt_stream_info, LIBZDT), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}, Ptr{Cint}, Ptr{Int64}),
                h[], nrb, nb, C_NULL, C_NULL))
    S = StreamSearch(h[], Nf, Nt, nrb[], nb[])
    finalizer(S) do q
        ccall((:zdt_stream_destroy, LIBZDT), Cint, (Ptr{Cvoid},), q.handle)
    end
    S
end

"""
    search(S, specs::Vector{Matrix{Float32}}, snr; cap=2^20) -> (hits, thresholds)

`hits[i]` lists spectrogram i's proto-hits as `CartesianIndex`es over the
hcat'd batches, like `findprotohits(fdrs, snr; snr=true)`.
"""
function search(S::StreamSearch, specs::Vector{Matrix{Float32}}, snr; cap::Integer=2^20)
    n = length(specs)
    all(size(s) == (S.Nf, S.Nt) for s in specs) || error("spectrograms must be $(S.Nf) x $(S.Nt)")
    hits = Vector{Int64}(undef, max(n * cap, 1))
    nh = zeros(Int64, n)
    thr = zeros(Float64, n)
    GC.@preserve specs begin
        rc = ccall((:zdt_stream_search, LIBZDT), Cint,
                   (Ptr{Cvoid}, Ptr{Ptr{Float32}}, Cint, Cdouble, Ptr{Int64}, Int64, Ptr{Int64}, Ptr{Cdouble}),
                   S.handle, [pointer(s) for s in specs], n, snr, hits, cap, nh, thr)
    end
    check(rc)
    out = [[CartesianIndex(Int(l % S.Nf) + 1, Int(l ÷ S.Nf) + 1) for l in hits[(i-1)*cap+1:(i-1)*cap+nh[i]]]
           for i in 1:n]
    out, thr
end

end # module ZDTHip
