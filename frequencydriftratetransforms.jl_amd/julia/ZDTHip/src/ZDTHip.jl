"""
    ZDTHip

Device side of FrequencyDriftRateTransforms.jl's Chirp-Z De-Doppler Transform
path on MI355X (gfx950): the `HipMatrix` device array and thin `ccall`
wrappers over the C-ABI of libzdt_hip.so (include/zdt.h).  It defines no
function of the package; `using FrequencyDriftRateTransforms, ZDTHip` loads
the extension `FrequencyDriftRateTransformsZDTHipExt`, which adds the
package's own methods (`ZDTWorkspace`, `input!`, `zdtfdr!`, `findprotohits`,
...) for `HipMatrix` spectrograms, as the package's CUDA extension does for
`CuArray` (ext/CUDAFrequencyDriftRateTransformsExt.jl:1-28).

The objects with no counterpart in the package (`RaggedPlan`, `StreamSearch`)
live here.  Every library call returns a status; a non-zero status throws
`ErrorException(zdt_last_error())`.
"""
module ZDTHip

export HipMatrix, HipZDTWorkspace, zdtpath, zdtprecision, clusterprotohits, findhits,
       RaggedPlan, bind!, launch!, planhits, plantuning, tune_table!, StreamSearch, search

const LIBZDT = get(ENV, "ZDT_LIB", joinpath(@__DIR__, "..", "..", "..", "lib", "libzdt_hip.so"))
const LIBHIP = "libamdhip64.so"

const ZDT_OK = 0
const ZDT_ETRUNC = 5
const WINDOW_RECT = 0
const WINDOW_HAMMING = 1
const WINDOW_CUSTOM = 2
const FLAG_UNNORMALIZED = 0x1
const PLAN_AUTOTUNE = 0x100
const PRECISIONS = Dict(:fp32 => 0, :split_bf16 => 1, :split_fp16 => 2)
const PATHS = Dict(0 => :generic, 1 => :fused, 2 => :fused_output)

zdt_error() = unsafe_string(ccall((:zdt_last_error, LIBZDT), Cstring, ()))
check(rc::Integer) = rc == ZDT_OK ? nothing : error(zdt_error())

# ----------------------------------------------------------------- device memory

"""
    HipMatrix{T}(undef, m, n) / HipMatrix(A::AbstractMatrix)

Column-major device matrix, the counterpart of a `CuMatrix` for this path.
`Array(H)` copies back to the host; `similar` allocates another one (which
makes the package's `create_fdr`, src/fdrutils.jl:11-26, work unchanged).
Memory is released by a finalizer (`hipFree`).
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
Base.similar(::HipMatrix, ::Type{T}, dims::Dims{2}) where {T} = HipMatrix{T}(undef, dims...)

function Base.Array(H::HipMatrix{T}) where {T}
    A = Matrix{T}(undef, H.dims...)
    rc = ccall((:hipMemcpy, LIBHIP), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Cint),
               A, H.ptr, sizeof(A), 2)  # hipMemcpyDeviceToHost
    rc == 0 || error("hipMemcpy D2H failed ($rc)")
    return A
end

device_synchronize() = ccall((:hipDeviceSynchronize, LIBHIP), Cint, ())

# ----------------------------------------------------------------- workspace handle

"""
    HipZDTWorkspace(spectrogram::HipMatrix{Float32}, r0, δr, Nr[, factors];
                    unnormalized=false, precision=:fp32, stream=C_NULL)

The library's workspace (zdt_create): device buffers, chirp kernel and, from
`spectrogram`, the input spectrum.  `precision` is `:fp32`, `:split_bf16` or
`:split_fp16` (the split modes run the column DFTs of input!/output! as
matrix-core products; power-of-two Nf in [2^17, 2^22]).  The package's
`ZDTWorkspace(spectrogram::HipMatrix{Float32}, ...)` returns one of these.
"""
mutable struct HipZDTWorkspace
    handle::Ptr{Cvoid}
    Nf::Int
    Nt::Int
    r0::Float32
    δr::Float32
    Nr::Int
    Nl::Int
    factors::Union{Tuple,AbstractVector}
    stream::Ptr{Cvoid}
    window::Union{Nothing,HipMatrix{Float32}}
end

_factors(f) = Cint[Int(x) for x in f]

function HipZDTWorkspace(spec::HipMatrix{Float32}, r0::Real, δr::Real, Nr::Integer,
                         factors::Union{Tuple,AbstractVector}=(2, 3, 5); unnormalized::Bool=false,
                         precision::Symbol=:fp32, stream::Ptr{Cvoid}=C_NULL)
    Nf, Nt = size(spec)
    f = _factors(factors)
    haskey(PRECISIONS, precision) || error("precision must be one of $(sort(collect(keys(PRECISIONS))))")
    h = Ref{Ptr{Cvoid}}(C_NULL)
    flags = unnormalized ? FLAG_UNNORMALIZED : 0
    check(ccall((:zdt_create, LIBZDT), Cint,
                (Ptr{Float32}, Int64, Int64, Cfloat, Cfloat, Int64, Ptr{Cint}, Cint, Cint, Cint, Ptr{Cvoid},
                 Ref{Ptr{Cvoid}}),
                spec.ptr, Nf, Nt, Float32(r0), Float32(δr), Nr, f, length(f), flags, PRECISIONS[precision],
                stream, h))
    nl = Ref{Int64}(0)
    check(ccall((:zdt_info, LIBZDT), Cint,
                (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Cfloat}, Ptr{Cfloat}, Ptr{Int64}, Ref{Int64}),
                h[], C_NULL, C_NULL, C_NULL, C_NULL, C_NULL, nl))
    ws = HipZDTWorkspace(h[], Nf, Nt, Float32(r0), Float32(δr), Nr, Int(nl[]), factors, stream, nothing)
    finalizer(ws) do w
        ccall((:zdt_destroy, LIBZDT), Cint, (Ptr{Cvoid},), w.handle)
    end
    return ws
end

function _path(ws::HipZDTWorkspace)
    p, q = Ref{Cint}(0), Ref{Cint}(0)
    check(ccall((:zdt_path, LIBZDT), Cint, (Ptr{Cvoid}, Ref{Cint}, Ref{Cint}), ws.handle, p, q))
    p[], q[]
end

"zdtpath(ws): the kernel path serving the workspace (:fused, :fused_output or :generic)."
zdtpath(ws::HipZDTWorkspace) = PATHS[_path(ws)[1]]

"zdtprecision(ws): the arithmetic of its column DFTs (:fp32, :split_bf16 or :split_fp16)."
zdtprecision(ws::HipZDTWorkspace) = Dict(v => k for (k, v) in PRECISIONS)[_path(ws)[2]]

function Base.show(io::IO, ws::HipZDTWorkspace)
    print(io, "HipZDTWorkspace(Nf=", ws.Nf, ",Nt=", ws.Nt, ",r0=", ws.r0, ",δr=", ws.δr,
          ",Nr=", ws.Nr, ",Nl=", ws.Nl, ",", ws.factors, ")")
end

function Base.sizeof(ws::HipZDTWorkspace)
    b = Ref{Int64}(0)
    check(ccall((:zdt_sizeof, LIBZDT), Cint, (Ptr{Cvoid}, Ref{Int64}), ws.handle, b))
    Int(b[])
end

# window -> (kind, device pointer); the window forms of src/zdtfdr.jl:297-321
function _window(ws::HipZDTWorkspace, w)
    w isa Val && (w = typeof(w).parameters[1])
    if w === :rect
        return WINDOW_RECT, Ptr{Float32}(C_NULL)
    elseif w === :hamming
        return WINDOW_HAMMING, Ptr{Float32}(C_NULL)
    elseif w isa Symbol
        error("unsupported window type ($w)")
    elseif w isa Function
        # evaluated on the host in Float64 (as Julia promotes w's result with
        # the ComplexF32 phase), stored as Float32
        v = Float32[w(n, ws.Nf) for n in 0:ws.Nf÷2]
        ws.window = HipMatrix(reshape(v, :, 1))
        return WINDOW_CUSTOM, ws.window.ptr
    end
    error("unsupported window argument $(typeof(w))")
end

function _input!(ws::HipZDTWorkspace, spec::HipMatrix{Float32})
    size(spec) == (ws.Nf, ws.Nt) || error("input spectrogram has unexpected size")
    check(ccall((:zdt_input, LIBZDT), Cint, (Ptr{Cvoid}, Ptr{Float32}), ws.handle, spec.ptr))
    ws
end

_preprocess!(ws::HipZDTWorkspace, r0::Float32) =
    check(ccall((:zdt_preprocess, LIBZDT), Cint, (Ptr{Cvoid}, Cfloat), ws.handle, r0))

_convolve!(ws::HipZDTWorkspace) = check(ccall((:zdt_convolve, LIBZDT), Cint, (Ptr{Cvoid},), ws.handle))

function _postprocess!(w, ws::HipZDTWorkspace)
    kind, ptr = _window(ws, w)
    check(ccall((:zdt_postprocess, LIBZDT), Cint, (Ptr{Cvoid}, Cint, Ptr{Float32}), ws.handle, kind, ptr))
end

function _output!(dest::HipMatrix{Float32}, ws::HipZDTWorkspace)
    size(dest) == (ws.Nf, ws.Nr) || error("dest has unexpected size")
    check(ccall((:zdt_output, LIBZDT), Cint, (Ptr{Cvoid}, Ptr{Float32}), ws.handle, dest.ptr))
    dest
end

# one fused batch: zdt_run (dest may be C_NULL: the spectra stay for output!)
function _run!(ws::HipZDTWorkspace, w, r0::Real, dest::Ptr{Float32})
    kind, ptr = _window(ws, w)
    check(ccall((:zdt_run, LIBZDT), Cint, (Ptr{Cvoid}, Cfloat, Cint, Ptr{Float32}, Ptr{Float32}),
                ws.handle, Float32(r0), kind, ptr, dest))
end

_sync(ws::HipZDTWorkspace) = check(ccall((:zdt_sync, LIBZDT), Cint, (Ptr{Cvoid},), ws.handle))

function _fftfdr!(fdr::HipMatrix{Float32}, ws::HipZDTWorkspace, rates)
    r = Float32.(collect(rates))
    size(fdr) == (ws.Nf, length(r)) || error("fdr must be $(ws.Nf) x $(length(r))")
    check(ccall((:zdt_fftfdr, LIBZDT), Cint, (Ptr{Cvoid}, Ptr{Cfloat}, Int64, Ptr{Float32}),
                ws.handle, r, length(r), fdr.ptr))
    fdr
end

function _fdshift!(dest::HipMatrix{Float32}, ws::HipZDTWorkspace, rate)
    size(dest) == (ws.Nf, ws.Nt) || error("dest must be $(ws.Nf) x $(ws.Nt)")
    check(ccall((:zdt_fdshift, LIBZDT), Cint, (Ptr{Cvoid}, Cfloat, Ptr{Float32}), ws.handle, Float32(rate),
                dest.ptr))
    dest
end

# ----------------------------------------------------------------- statistics and hits

const HipFdrs = Union{Tuple{Vararg{HipMatrix{Float32}}},AbstractVector{<:HipMatrix{Float32}}}

_ptrs(fdrs) = Ptr{Float32}[f.ptr for f in fdrs]

function _stats(fdrs)
    Nf, Nr = size(first(fdrs))
    m = Ref{Cdouble}(0)
    s = Ref{Cdouble}(0)
    check(ccall((:zdt_fdrstats, LIBZDT), Cint,
                (Ptr{Ptr{Float32}}, Cint, Int64, Int64, Int64, Ptr{Cvoid}, Ref{Cdouble}, Ref{Cdouble}),
                _ptrs(fdrs), length(fdrs), Nf, Nr, Nf, C_NULL, m, s))
    (m[], s[])
end

function _normalize!(fdrs, m, s)
    Nf, Nr = size(first(fdrs))
    check(ccall((:zdt_fdrnormalize, LIBZDT), Cint,
                (Ptr{Ptr{Float32}}, Cint, Int64, Int64, Int64, Cdouble, Cdouble, Ptr{Cvoid}),
                _ptrs(fdrs), length(fdrs), Nf, Nr, Nf, m, s, C_NULL))
    fdrs
end

# every element >= threshold, column-major over the hcat'd matrices, as 0-based
# linear indices i + Nf*(b*Nr + j) left on the device: (H, n)
function _hits_device(fdrs, threshold)
    Nf, Nr = size(first(fdrs))
    n = Ref{Int64}(0)
    ptrs = _ptrs(fdrs)
    rc = ccall((:zdt_findprotohits, LIBZDT), Cint,
               (Ptr{Ptr{Float32}}, Cint, Int64, Int64, Int64, Cdouble, Ptr{Int64}, Int64, Ref{Int64}, Ptr{Cvoid}),
               ptrs, length(fdrs), Nf, Nr, Nf, threshold, C_NULL, 0, n, C_NULL)
    rc in (ZDT_OK, ZDT_ETRUNC) || check(rc)
    H = HipMatrix{Int64}(undef, max(n[], 1), 1)
    check(ccall((:zdt_findprotohits, LIBZDT), Cint,
                (Ptr{Ptr{Float32}}, Cint, Int64, Int64, Int64, Cdouble, Ptr{Int64}, Int64, Ref{Int64}, Ptr{Cvoid}),
                ptrs, length(fdrs), Nf, Nr, Nf, threshold, H.ptr, n[], n, C_NULL))
    H, n[]
end

_cartesian(lin, Nf) = [CartesianIndex(Int(l % Nf) + 1, Int(l ÷ Nf) + 1) for l in lin]

# the same as 1-based CartesianIndex{2}
function _hits(fdrs, threshold)
    H, n = _hits_device(fdrs, threshold)
    _cartesian(n == 0 ? Int64[] : vec(Array(H))[1:n], size(first(fdrs), 1))
end

# ----------------------------------------------------------------- hits from proto-hits

"""
    clusterprotohits(fdr[s], hijs; di=1, dj=1) -> (; peak, value, size, box, labels)
    findhits(fdr[s], threshold; snr=false, di=1, dj=1) -> (hijs, hits)

Group proto-hits into hits (include/zdt.h, `zdt_clusterprotohits`).  Two
proto-hits are neighbours when their channels differ by at most `di` and their
(hcat'd) rate columns by at most `dj`; a hit is a connected group.  Hits are
numbered by their first proto-hit in column-major order:

- `peak[h]`: the `CartesianIndex` of hit h's largest value `value[h]` (ties to
  the earlier proto-hit);
- `size[h]`: its number of proto-hits;
- `box[h]`: (first channel, last channel, first column, last column);
- `labels[p]`: the hit of proto-hit p.

`findhits` is `findprotohits` followed by the clustering, with the proto-hits
kept on the device in between.  No counterpart in the package: it stops at
proto-hits and leaves this step to the next stage of a search
(src/fdrutils.jl:141-144, README.md:48-53).
"""
function clusterprotohits(fdrs::HipFdrs, hijs::AbstractVector{<:CartesianIndex{2}}; di::Integer=1, dj::Integer=1)
    Nf = size(first(fdrs), 1)
    lin = Int64[(h[1] - 1) + Nf * (h[2] - 1) for h in hijs]
    H = HipMatrix(reshape(isempty(lin) ? Int64[0] : lin, :, 1))
    _cluster(fdrs, H, length(lin), di, dj)
end
clusterprotohits(fdr::HipMatrix{Float32}, hijs; kw...) = clusterprotohits((fdr,), hijs; kw...)

function findhits(fdrs::HipFdrs, threshold; snr::Bool=false, di::Integer=1, dj::Integer=1)
    if snr  # the package's fdrdenormalize: snr * s + m (src/fdrutils.jl:104-119)
        m, s = _stats(fdrs)
        threshold = threshold * s + m
    end
    H, n = _hits_device(fdrs, threshold)
    hijs = _cartesian(n == 0 ? Int64[] : vec(Array(H))[1:n], size(first(fdrs), 1))
    hijs, _cluster(fdrs, H, n, di, dj)
end
findhits(fdr::HipMatrix{Float32}, threshold; kw...) = findhits((fdr,), threshold; kw...)

function _cluster(fdrs, H::HipMatrix{Int64}, n::Integer, di::Integer, dj::Integer)
    Nf, Nr = size(first(fdrs))
    ptrs = _ptrs(fdrs)
    nc = Ref{Int64}(0)
    rc = ccall((:zdt_clusterprotohits, LIBZDT), Cint,
               (Ptr{Int64}, Int64, Ptr{Ptr{Float32}}, Cint, Int64, Int64, Int64, Cint, Cint,
                Ptr{Int64}, Ptr{Float32}, Ptr{Int64}, Ptr{Int64}, Int64, Ref{Int64}, Ptr{Int64}, Ptr{Cvoid}),
               H.ptr, n, ptrs, length(fdrs), Nf, Nr, Nf, di, dj,
               C_NULL, C_NULL, C_NULL, C_NULL, 0, nc, C_NULL, C_NULL)
    rc in (ZDT_OK, ZDT_ETRUNC) || check(rc)
    c = max(nc[], 1)
    P = HipMatrix{Int64}(undef, c, 1)
    V = HipMatrix{Float32}(undef, c, 1)
    S = HipMatrix{Int64}(undef, c, 1)
    B = HipMatrix{Int64}(undef, 4, c)
    L = HipMatrix{Int64}(undef, max(n, 1), 1)
    check(ccall((:zdt_clusterprotohits, LIBZDT), Cint,
                (Ptr{Int64}, Int64, Ptr{Ptr{Float32}}, Cint, Int64, Int64, Int64, Cint, Cint,
                 Ptr{Int64}, Ptr{Float32}, Ptr{Int64}, Ptr{Int64}, Int64, Ref{Int64}, Ptr{Int64}, Ptr{Cvoid}),
                H.ptr, n, ptrs, length(fdrs), Nf, Nr, Nf, di, dj,
                P.ptr, V.ptr, S.ptr, B.ptr, nc[], nc, L.ptr, C_NULL))
    h = nc[]
    box = Array(B)
    (; peak = _cartesian(vec(Array(P))[1:h], Nf),
       value = vec(Array(V))[1:h],
       size = Int.(vec(Array(S))[1:h]),
       box = [Tuple(Int.(box[:, q]) .+ 1) for q in 1:h],
       labels = n == 0 ? Int[] : Int.(vec(Array(L))[1:n]) .+ 1)
end

# ----------------------------------------------------------------- ragged plans

"""
    RaggedPlan(specs::Vector{HipMatrix{Float32}}, rates::Vector; autotune=true)

Many coarse channels of uneven Nf and Nt through `input!`, one `zdtfdr!`
batch, `fdrdenormalize` and `findprotohits`, replayed as one hipGraph launch
(include/zdt.h, zdt_plan_*).  `rates[c]` is a range (`r0:δr:r1`) or an
`(r0, δr, Nr)` tuple.  No counterpart in the package: it replaces the
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

"""
    plantuning(plan, c) -> (source, ms)

Where channel c's kernel choice came from (`:default`, `:timed` or `:table`)
and the chosen candidate's time for input! plus one batch (zdt_plan_tuning).
"""
function plantuning(P::RaggedPlan, c::Integer)
    src, ms = Ref{Cint}(0), Ref{Cfloat}(0)
    check(ccall((:zdt_plan_tuning, LIBZDT), Cint, (Ptr{Cvoid}, Cint, Ref{Cint}, Ref{Cfloat}), P.handle, c - 1, src, ms))
    (Dict(0 => :default, 1 => :timed, 2 => :table)[src[]], ms[])
end

"tune_table!(path): the tuning table autotuned plans read and append to (`nothing` turns it off)."
tune_table!(path::Union{Nothing,AbstractString}) =
    check(ccall((:zdt_tune_table, LIBZDT), Cint, (Cstring,), path === nothing ? C_NULL : path))

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
    StreamSearch(Nf, Nt, rates; Nrb=0, hbm_bytes=0, precision=:fp32)

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

function StreamSearch(Nf::Integer, Nt::Integer, rates; Nrb::Integer=0, hbm_bytes::Integer=0,
                      precision::Symbol=:fp32)
    r0, δr, Nr = _rate_params(rates)
    haskey(PRECISIONS, precision) || error("precision must be one of $(sort(collect(keys(PRECISIONS))))")
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:zdt_stream_create, LIBZDT), Cint,
                (Int64, Int64, Cfloat, Cfloat, Int64, Int64, Int64, Cint, Cint, Ref{Ptr{Cvoid}}),
                Nf, Nt, r0, δr, Nr, Nrb, hbm_bytes, 0, PRECISIONS[precision], h))
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

"""
    findhits(S, specs::Vector{Matrix{Float32}}, snr; di=1, dj=1, cap=2^16, pcap=2^22)
        -> (hits, nprotohits, thresholds)

The search ending in hits (`zdt_stream_findhits`): spectrogram i's proto-hits
clustered on the device while its batches' FDRs are resident, as
`clusterprotohits` clusters them.  `hits[i]` has the fields of
`clusterprotohits`' result, with empty `labels`.  Needs a resident stream.
"""
function findhits(S::StreamSearch, specs::Vector{Matrix{Float32}}, snr; di::Integer=1, dj::Integer=1,
                  cap::Integer=2^16, pcap::Integer=2^22)
    n = length(specs)
    all(size(s) == (S.Nf, S.Nt) for s in specs) || error("spectrograms must be $(S.Nf) x $(S.Nt)")
    c1 = max(cap, 1)
    peak = zeros(Int64, n * c1)
    value = zeros(Float32, n * c1)
    sz = zeros(Int64, n * c1)
    box = zeros(Int64, 4 * n * c1)
    ncl = zeros(Int64, n)
    nph = zeros(Int64, n)
    thr = zeros(Float64, n)
    GC.@preserve specs begin
        rc = ccall((:zdt_stream_findhits, LIBZDT), Cint,
                   (Ptr{Cvoid}, Ptr{Ptr{Float32}}, Cint, Cdouble, Cint, Cint, Int64, Ptr{Int64}, Ptr{Float32},
                    Ptr{Int64}, Ptr{Int64}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Cdouble}),
                   S.handle, [pointer(s) for s in specs], n, snr, di, dj, pcap, peak, value, sz, box, cap, ncl,
                   nph, thr)
    end
    check(rc)
    hits = map(1:n) do i
        h = Int(ncl[i])
        o = (i - 1) * cap
        bx = reshape(box[4o+1:4(o+h)], 4, h)
        (; peak = _cartesian(peak[o+1:o+h], S.Nf),
           value = value[o+1:o+h],
           size = Int.(sz[o+1:o+h]),
           box = [Tuple(Int.(bx[:, q]) .+ 1) for q in 1:h],
           labels = Int[])
    end
    hits, nph, thr
end

end # module ZDTHip
