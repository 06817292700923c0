# LinearSolve.jl adapter: LHLFactorization / update_gamma! over a SciMLOperators
# WOperator (README.md:57-75, src/LHLFactorization.jl:39-41), on the GPU library.
#
# The operator's concrete form is W = J - M/γ (transform = true, the split
# README.md:57 names) or W = γJ - M; the mass matrix must be λI, so that
# W = σI + τJ with (σ, τ) = (-λ/γ, 1) or (-λ, γ).  init reduces J once; a new γ
# only reloads the shift (O(n²)) at the next solve!; a complex γ (a Radau stage)
# loads a complex LHLShift held beside the real one, on the same reduction.
# Parity unpinned: LinearSolve.jl and SciMLOperators.jl are not in the reference.
# The Python mirror lhl_hip/linsolve.py implements the same mapping and is what
# the tests run (no Julia is installed here).
module LHLFactorizationHIPLinearSolveExt

using LinearAlgebra, LinearSolve, SciMLBase, SciMLOperators
using LHLFactorizationHIP
const H = LHLFactorizationHIP

struct HIPLHLFactorization <: LinearSolve.AbstractFactorization
    refine::Int          # lhl_refine! rounds per solve against σI + τJ (0: none)
    tol::Float64         # stop refining once every column is below it (0: fixed rounds)
end
HIPLHLFactorization(; refine::Integer = 0, tol::Real = 0.0) = HIPLHLFactorization(refine, tol)

_mass_scalar(M::UniformScaling) = M.λ
_mass_scalar(M::Number) = M
function _mass_scalar(M::AbstractMatrix)
    λ = M[1, 1]
    M == λ * I || throw(ArgumentError("HIPLHLFactorization: the mass matrix must be a multiple of I"))
    return λ
end

function _shift_of(W::SciMLOperators.WOperator)
    λ = _mass_scalar(W.mass_matrix)
    iszero(W.gamma) && throw(ArgumentError("HIPLHLFactorization: gamma must be nonzero"))
    return W.transform ? (-λ / W.gamma, one(W.gamma)) : (-λ, W.gamma)
end

mutable struct HIPLHLCache
    ws::H.LHLWorkspace
    csh::Union{Nothing,H.LHLShift}     # the complex stage's state, made on first use
    loaded::Dict{Bool,Any}             # (σ, τ) held by the real / complex state
end

_denseJ(W) = Matrix{Float64}(convert(AbstractMatrix, W.J))

function LinearSolve.init_cacheval(alg::HIPLHLFactorization, W::SciMLOperators.WOperator, b, u, Pl, Pr,
                                   maxiters::Int, abstol, reltol, verbose, assumptions)
    return HIPLHLCache(H.lhl(_denseJ(W)), nothing, Dict{Bool,Any}())
end

function SciMLBase.solve!(cache::LinearSolve.LinearCache, alg::HIPLHLFactorization; kwargs...)
    c = cache.cacheval
    W = cache.A
    if cache.isfresh                      # a new operator: reduce its J (O(n³))
        H.lhl!(c.ws, _denseJ(W))
        empty!(c.loaded)
        cache.isfresh = false
    end
    σ, τ = _shift_of(W)
    cplx = σ isa Complex || τ isa Complex
    sh = c.ws.shift
    if cplx
        c.csh === nothing && (c.csh = H.LHLShift{ComplexF64}(c.ws))
        sh = c.csh
    end
    if get(c.loaded, cplx, nothing) != (σ, τ)   # O(n²): the reduction is reused
        cplx ? H.lhl_shift!(sh, c.ws, σ, τ) : H.lhl_shift!(c.ws, σ, τ)
        c.loaded[cplx] = (σ, τ)
    end
    copyto!(cache.u, cache.b)
    cplx ? H.lhl_ldiv!(cache.u, sh, c.ws) : H.lhl_ldiv!(cache.u, c.ws)
    if alg.refine > 0 && sh.info == 0
        cplx ? H.lhl_refine!(cache.u, nothing, cache.b, sh, c.ws, alg.refine; tol = alg.tol) :
               H.lhl_refine!(cache.u, nothing, cache.b, c.ws, alg.refine; tol = alg.tol)
    end
    rc = sh.info == 0 ? SciMLBase.ReturnCode.Success : SciMLBase.ReturnCode.Failure
    return SciMLBase.build_linear_solution(alg, cache.u, nothing, cache; retcode = rc)
end

# update_gamma!(cache, γ): move the shift; the next solve! reloads it in O(n²).
function update_gamma!(cache::LinearSolve.LinearCache, γ)
    cache.A.gamma = γ
    return cache
end

export HIPLHLFactorization, update_gamma!
end
