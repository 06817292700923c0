# LHLFactorizationHIP: the reference's public names (src/LHLFactorization.jl:55-56)
# over liblhl_hip.so's C ABI (include/lhl_hip.h), through ccall.  Host arrays go
# through the library's staging copies (on_device = 0); for AMDGPU.jl ROCArrays
# pass on_device = 1 and the device pointer (same HIP device as the library's).
# Not run here: no Julia toolchain is installed.  The Python mirror
# (lhl.jl_amd/lhl_hip) calls the same symbols and is what the tests run.
module LHLFactorizationHIP

using LinearAlgebra

const liblhl = joinpath(@__DIR__, "..", "lib", "liblhl_hip.so")
const LHL_F64, LHL_F32, LHL_C64, LHL_C32 = Cint(0), Cint(1), Cint(2), Cint(3)
const LHL_PREC_WORKING, LHL_PREC_REFINE64 = Cint(0), Cint(1)
const LHL_NOT_CONVERGED = Cint(6)

dtype_code(::Type{Float64}) = LHL_F64
dtype_code(::Type{Float32}) = LHL_F32
dtype_code(::Type{ComplexF64}) = LHL_C64     # a complex J / complex workspace (interleaved)
dtype_code(::Type{ComplexF32}) = LHL_C32
const LHLElt = Union{Float32,Float64,ComplexF32,ComplexF64}

function check(st::Cint, h::Ptr{Cvoid})
    st == 0 && return nothing
    msg = unsafe_string(ccall((:lhl_last_error, liblhl), Cstring, (Ptr{Cvoid},), h))
    st == 1 && throw(ArgumentError(msg))
    st == 2 && throw(DimensionMismatch(msg))
    st == LHL_NOT_CONVERGED && (@warn msg; return nothing)   # outputs are written
    error(msg)
end

# LHLShift{TG,Tr} (src:84-116).  `ptr` is C_NULL for ws.shift (the handle's own
# state) and for an unbound LHLShift{TG}(n) until its first lhl_shift!.
mutable struct LHLShift{TG,Tr}
    ptr::Ptr{Cvoid}
    ws::Any                          # the workspace it serves (keeps it alive)
    own::Bool                        # ws.shift
    n::Int
    version::Int                     # bumped by every lhl_shift!
    cached::Int                      # the version the host mirrors hold
    Gt::Matrix{TG}; rdiag::Vector{TG}; swap::Vector{Bool}; resid::Vector{TG}
end
LHLShift{TG}(n::Integer) where {TG} =
    LHLShift{TG,real(TG)}(C_NULL, nothing, false, n, 0, -1,
                          zeros(TG, n, n), zeros(TG, n), falses(n), zeros(TG, n))

mutable struct LHLWorkspace{T,Tr,TG}
    h::Ptr{Cvoid}
    n::Int
    shift::LHLShift{TG,Tr}
end

function LHLWorkspace{T}(n::Integer; shift::Type = T) where {T<:LHLElt}
    shift === T || shift === Complex{real(T)} ||
        throw(ArgumentError("shift type must be $T or $(Complex{real(T)}), got $shift"))   # src:167-169
    r = Ref{Ptr{Cvoid}}(C_NULL)
    st = ccall((:lhl_create, liblhl), Cint, (Ref{Ptr{Cvoid}}, Int64, Cint, Cint),
               r, n, dtype_code(T), shift <: Complex)
    check(st, r[])
    sh = LHLShift{shift}(n); sh.own = true
    ws = LHLWorkspace{T,real(T),shift}(r[], n, sh)
    sh.ws = ws
    finalizer(w -> ccall((:lhl_destroy, liblhl), Cint, (Ptr{Cvoid},), w.h), ws)
end

function LHLShift{TG}(ws::LHLWorkspace) where {TG}
    sh = LHLShift{TG}(ws.n)
    _bind!(sh, ws)
    return sh
end

function _bind!(sh::LHLShift{TG}, ws::LHLWorkspace{T}) where {TG,T}
    sh.ws === ws && return sh
    (TG === T || TG === Complex{real(T)}) ||
        throw(ArgumentError("an LHLShift{$TG} cannot serve a workspace of element type $T"))
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:lhl_shift_create, liblhl), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}, Cint),
                ws.h, r, TG <: Complex), ws.h)
    sh.ptr, sh.ws, sh.n = r[], ws, ws.n
    finalizer(s -> s.ptr == C_NULL || ccall((:lhl_shift_destroy, liblhl), Cint, (Ptr{Cvoid},), s.ptr), sh)
    return sh
end

function lhl_reduce!(ws::LHLWorkspace{T}, J::StridedMatrix{TJ}, balance::Bool = true,
                     thread = Val(true)) where {T,TJ<:LHLElt}
    n = LinearAlgebra.checksquare(J)
    st = GC.@preserve J ccall((:lhl_reduce, liblhl), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Cint, Cint, Cint),
               ws.h, pointer(J), n, max(stride(J, 2), n), dtype_code(TJ), balance, 0)
    check(st, ws.h)
    ws.n = n
    ws.shift.n = n
    ws.shift.version += 1            # the held shift is unloaded by a new reduction
    return ws
end
lhl!(ws::LHLWorkspace, J; balance::Bool = true, thread = Val(true)) = lhl_reduce!(ws, J, balance, thread)
lhl(J::AbstractMatrix{T}; balance::Bool = true, shift::Type = T, thread = Val(true)) where {T} =
    lhl_reduce!(LHLWorkspace{T}(LinearAlgebra.checksquare(J); shift), J, balance, thread)

_shptr(sh::LHLShift) = sh.own ? C_NULL : sh.ptr

function _shift!(sh::LHLShift{TG}, ws::LHLWorkspace, σ::Number, τ::Number) where {TG}
    sh.own || _bind!(sh, ws)
    info = Ref{Int64}(0)
    st = ccall((:lhl_shift, liblhl), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Float64, Float64, Float64, Float64, Ref{Int64}),
               ws.h, _shptr(sh), real(σ), imag(σ), real(τ), imag(τ), info)
    check(st, ws.h)                  # a complex σ/τ on a real state is ArgumentError (src:2466-2473)
    sh.n = ws.n
    sh.version += 1
    return info[]
end
lhl_shift!(ws::LHLWorkspace, σ, τ) = (_shift!(ws.shift, ws, σ, τ); ws)
lhl_shift!(sh::LHLShift, ws::LHLWorkspace, σ, τ) = (_shift!(sh, ws, σ, τ); sh)

# The shift's scalar fields (no device copy) and its array fields (host mirrors,
# refreshed once per loaded shift so repeated reads return the same arrays).
function _shift_info(sh::LHLShift{TG}) where {TG}
    ws = sh.ws
    info = Ref{Int64}(0); sg = zeros(2); ta = zeros(2); loaded = Ref{Cint}(0)
    check(ccall((:lhl_get_shift_info, liblhl), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ref{Int64}, Ptr{Float64}, Ptr{Float64}, Ref{Cint}),
                ws.h, _shptr(sh), info, sg, ta, loaded), ws.h)
    conv(v) = TG <: Complex ? TG(v[1], v[2]) : TG(v[1])
    return Int(info[]), conv(sg), conv(ta)
end
function _refresh!(sh::LHLShift{TG}) where {TG}
    sh.cached == sh.version && return sh
    n = sh.n
    size(sh.Gt) == (n, n) || (sh.Gt = zeros(TG, n, n); sh.rdiag = zeros(TG, n);
                              sh.swap = falses(n); sh.resid = zeros(TG, n))
    U = zeros(TG, n, n); l = zeros(TG, max(n - 1, 1)); sw = zeros(UInt8, n)
    check(ccall((:lhl_get_shift_state, liblhl), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{TG}, Ptr{TG}, Ptr{UInt8}, Ptr{TG},
                 Ptr{Float64}, Ptr{Float64}, Ptr{Int64}),
                sh.ws.h, _shptr(sh), U, l, sw, sh.rdiag, C_NULL, C_NULL, C_NULL), sh.ws.h)
    sh.Gt .= transpose(U)            # Gt: U transposed with the multipliers on the superdiagonal
    for k in 1:n-1
        sh.Gt[k, k+1] = l[k]
    end
    sh.swap .= sw .!= 0
    sh.cached = sh.version
    return sh
end

const _SHIFT_FIELDS = (:Gt, :rdiag, :swap, :resid, :σ, :τ, :info)
function Base.getproperty(sh::LHLShift, s::Symbol)
    s === :info && return _shift_info(sh)[1]
    s === :σ && return _shift_info(sh)[2]
    s === :τ && return _shift_info(sh)[3]
    s in (:Gt, :rdiag, :swap) && return getfield(_refresh!(sh), s)
    return getfield(sh, s)
end
Base.propertynames(::LHLShift) = (fieldnames(LHLShift)..., :σ, :τ, :info)
function Base.getproperty(ws::LHLWorkspace, s::Symbol)          # src:182-191
    s in _SHIFT_FIELDS && return getproperty(getfield(ws, :shift), s)
    return getfield(ws, s)
end
Base.propertynames(::LHLWorkspace) = (fieldnames(LHLWorkspace)..., _SHIFT_FIELDS...)

function _ldiv!(x::StridedVecOrMat{S}, sh::LHLShift, ws::LHLWorkspace) where {S}
    R = real(S)
    st = GC.@preserve x ccall((:lhl_ldiv, liblhl), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Int64, Cint, Cint, Cint),
               ws.h, _shptr(sh), pointer(x), size(x, 1), size(x, 2), max(stride(x, 2), size(x, 1)),
               dtype_code(R), S <: Complex, 0)
    check(st, ws.h)
    return x
end
lhl_ldiv!(x, ws::LHLWorkspace) = _ldiv!(x, ws.shift, ws)
lhl_ldiv!(x, sh::LHLShift, ws::LHLWorkspace) = _ldiv!(x, sh, ws)

# A: a dense operator (Matrix), or `nothing` for sigma*I + tau*J with the resident FP64 J.
# `tol` > 0 (an extension) stops once every column reaches it.
function _refine!(x::StridedVecOrMat{S}, A, b, sh::LHLShift, ws::LHLWorkspace, steps::Integer;
                  tol = 0.0) where {S}
    R = real(S)
    W = A === nothing ? nothing : convert(Matrix{S}, A)
    bb = convert(Array{S}, b)
    iters = Ref{Cint}(0)
    st = GC.@preserve x W bb ccall((:lhl_refine, liblhl), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Int64, Int64,
                Cint, Cint, Cint, Cint, Float64, Ref{Cint}),
               ws.h, _shptr(sh), pointer(x), W === nothing ? C_NULL : pointer(W),
               W === nothing ? 0 : stride(W, 2), pointer(bb), size(x, 1), size(x, 2),
               max(stride(x, 2), size(x, 1)), dtype_code(R), S <: Complex, 0, steps, tol, iters)
    check(st, ws.h)
    return x
end
lhl_refine!(x, A, b, ws::LHLWorkspace, steps::Int; tol = 0.0) = _refine!(x, A, b, ws.shift, ws, steps; tol)
lhl_refine!(x, A, b, sh::LHLShift, ws::LHLWorkspace, steps::Int; tol = 0.0) =
    _refine!(x, A, b, sh, ws, steps; tol)

for (jname, cname) in ((:applyZ!, :lhl_applyZ), (:applyZinv!, :lhl_applyZinv))
    @eval function $jname(x::StridedVecOrMat{S}, ws::LHLWorkspace) where {S}
        st = GC.@preserve x ccall(($(QuoteNode(cname)), liblhl), Cint,
                   (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64, Int64, Cint, Cint, Cint),
                   ws.h, pointer(x), size(x, 1), size(x, 2), max(stride(x, 2), size(x, 1)),
                   dtype_code(real(S)), S <: Complex, 0)
        check(st, ws.h)
        return x
    end
end

# Many shifts at once: shifts[s] = (σ, τ); the RHS of shift s are the columns
# rhs_offsets[s]+1 : rhs_offsets[s+1] of B (Julia 1-based ranges over 0-based offsets).
# The library runs the shifts in equal chunks whose packed U's stay under 64 GiB (or a
# third of the free device memory); ENV["LHL_SHIFT_POOL_BYTES"] lowers that bound for a
# process that shares the GPU.  A tolerance not reached in max_iter rounds is a warning,
# with X, info and iters written.
function shift_solve_batched!(X::Matrix{S}, ws::LHLWorkspace, σ::AbstractVector, τ::AbstractVector,
                              B::Matrix{S}, rhs_offsets::Vector{Int64};
                              refine64::Bool = true, tol = 1e-12, max_iter = 30) where {S<:Union{Float64,ComplexF64}}
    ns = length(σ)
    sig = collect(Iterators.flatten((real(s), imag(s)) for s in σ))
    ta = collect(Iterators.flatten((real(t), imag(t)) for t in τ))
    info = zeros(Int64, ns); iters = zeros(Cint, ns)
    st = ccall((:lhl_shift_solve_batched, liblhl), Cint,
               (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Int64}, Ptr{Cvoid}, Int64,
                Ptr{Cvoid}, Int64, Cint, Cint, Float64, Cint, Ptr{Int64}, Ptr{Cint}),
               ws.h, ns, sig, ta, S <: Complex, rhs_offsets, B, stride(B, 2), X, stride(X, 2), 0,
               refine64 ? LHL_PREC_REFINE64 : LHL_PREC_WORKING, tol, max_iter, info, iters)
    check(st, ws.h)
    return X, info, iters
end

export lhl, lhl!, lhl_reduce!, lhl_shift!, lhl_ldiv!, lhl_refine!, applyZ!, applyZinv!,
       LHLWorkspace, LHLShift, shift_solve_batched!
end
