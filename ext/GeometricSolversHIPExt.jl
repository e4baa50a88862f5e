# GeometricSolversHIPExt — the R2 batched solve of GeometricSolvers.jl on MI355X through libgsolve_hip.
#
# A package extension (weakdep AMDGPU) that adds more specific methods of `solve_batch!` for device
# batches of the registered problem families. Each method keeps the signature, the host validation
# and the per-system semantics of the generic method (src/nonlinear/batched.jl:257-332) and replaces
# only its launch (`_launch!` → `_solve_kernel!`, batched.jl:171-195) with one `ccall` into the
# C ABI of include/gsolve.h. A family, method or option this library does not serve falls through
# to the generic KernelAbstractions path.
#
# The families below are ordinary `NonlinearProblem` residuals with analytic Jacobians, written in
# the operation order the device functions follow; on the host they are what `solve` runs, so the
# reference's own R1 `solve` of the same problem is the oracle of each device solve.
#
# Project.toml of the reference would gain:
#   [weakdeps]   AMDGPU = "21141c5a-9bdb-4563-92ae-f87d6854732e"
#   [extensions] GeometricSolversHIPExt = "AMDGPU"
#   [compat]     AMDGPU = "2.12"
module GeometricSolversHIPExt

using AMDGPU: ROCVector, ROCArray, stream
using GeometricSolvers
using GeometricSolvers: Options, SolverStatus, ReturnCode, L2Norm, MaxNorm, Static, Backtracking, DogLeg,
                        Newton, NoRefinement, Refinement, LUFactorization, NonlinearProblem, PerSystem,
                        BatchParams, _validate_batch, _validate_batch_problem, _validate_workgroupsize,
                        resolve_options, _tally_return_codes, ToReal
using StaticArrays: SVector, SMatrix, @SMatrix
import GeometricSolvers: solve_batch!

const libgsolve = get(ENV, "GSOLVE_HIP_LIB", "libgsolve_hip")

# ---- the ABI structs (include/gsolve.h) -------------------------------------------------------

struct GsOptionsF64
    f_abstol::Float64
    f_reltol::Float64
    x_abstol::Float64
    x_reltol::Float64
    min_iterations::Int32
    max_iterations::Int32
    max_stalls::Int32
    norm::Int32
    g_tol::Float64
    kappa::Float64
    roundoff::UInt8
    rate_test::UInt8
    pad::NTuple{6, UInt8}
end

struct GsMethod
    linsolve::Int32
    factor_precision::Int32
    globalization::Int32
    refinement::Int32
    static_alpha::Float64
    bt_c1::Float64
    bt_p::Float64
    bt_maxiter::Int32
    ref_maxiter::Int32
    ref_eta::Float64
    regularization::Float64
    lm::NTuple{6, Float64}
    best::Int32
    reuse::Int32
    dl_radius::Float64
    lm_scaling::Int32
    pad::Int32
end

struct GsFamily
    id::Int32
    n::Int32
    m::Int32
    p_len::Int32
    shared::NTuple{8, Float64}
end

# SolverStatus{Float64} crosses the ABI as itself: its C layout is gs_status_f64.
function __init__()
    sizeof(SolverStatus{Float64}) == 72 || error("SolverStatus{Float64} is not 72 bytes")
    fieldoffset(SolverStatus{Float64}, 4) == 8 || error("SolverStatus{Float64}.fnorm is not at 8")
    fieldoffset(SolverStatus{Float64}, 10) == 32 || error("SolverStatus{Float64}.rank is not at 32")
    fieldoffset(SolverStatus{Float64}, 14) == 64 || error("SolverStatus{Float64}.contracted is not at 64")
    sizes = zeros(Int64, 4)
    ccall((:gs_abi_sizes, libgsolve), Cvoid, (Ptr{Int64},), sizes)
    sizes == [72, sizeof(GsOptionsF64), sizeof(GsMethod), sizeof(GsFamily)] ||
        error("libgsolve_hip ABI sizes $(sizes) do not match this extension")
    return nothing
end

# ---- the registered families ------------------------------------------------------------------

abstract type HIPFamily end

"""
    HenonHeilesMidpoint()

The implicit-midpoint step of Hénon–Heiles, `x = (q₁, q₂, p₁, p₂)`: with `p = (h, x_k)` (a
`BatchParams(h, X_k)`), `F(x) = x − x_k − h f((x_k + x)/2)`, `f = (p₁, p₂, −q₁ − 2q₁q₂,
−q₂ − q₁² + q₂²)`, and `J = I − (h/2) Df`.
"""
struct HenonHeilesMidpoint <: HIPFamily end

function (::HenonHeilesMidpoint)(x::SVector{4, T}, p) where {T}
    h, xk = p
    m = (xk + x) / 2
    f = SVector{4, T}(m[3], m[4], -m[1] - 2 * m[1] * m[2], -m[2] - m[1]^2 + m[2]^2)
    return x - xk - h * f
end

function hh_midpoint_jacobian(x::SVector{4, T}, p) where {T}
    h, xk = p
    q1 = (xk[1] + x[1]) / 2
    q2 = (xk[2] + x[2]) / 2
    c = h / 2
    o, z = one(T), zero(T)
    return @SMatrix [o z -c z;
                     z o z -c;
                     c*(1 + 2 * q2) c*(2 * q1) o z;
                     c*(2 * q1) c*(1 - 2 * q2) z o]
end

"""
    TanhLinear3()

`F(x, (A, b)) = A x + 0.1 tanh.(x) − b`, `J = A + Diagonal(0.1 (1 − tanh.(x)²))`, `n = 3`.
"""
struct TanhLinear3 <: HIPFamily end

(::TanhLinear3)(x::SVector{3, T}, p) where {T} = p[1] * x + T(0.1) * tanh.(x) - p[2]
function tanh_linear3_jacobian(x::SVector{3, T}, p) where {T}
    t = tanh.(x)
    return p[1] + SMatrix{3, 3, T}(ntuple(k -> (k - 1) % 4 == 0 ? T(0.1) * (1 - t[(k - 1) ÷ 3 + 1]^2) : zero(T), 9))
end

struct Coupled2 <: HIPFamily end
(::Coupled2)(x::SVector{2, T}, p) where {T} = SVector{2, T}(x[1] + x[2]^2 / T(8) - p[1], x[2] + x[1]^2 / T(8) - p[2])
coupled2_jacobian(x::SVector{2, T}, _) where {T} = @SMatrix [one(T) x[2]/T(4); x[1]/T(4) one(T)]

struct ScalarSquare <: HIPFamily end
(::ScalarSquare)(x::T, p) where {T <: Number} = x^2 - T(p)
scalar_square_jacobian(x, _) = 2 * x

"""The `NonlinearProblem` of a registered family, with its analytic Jacobian."""
family_problem(f::HenonHeilesMidpoint) = NonlinearProblem(f, hh_midpoint_jacobian)
family_problem(f::TanhLinear3) = NonlinearProblem(f, tanh_linear3_jacobian)
family_problem(f::Coupled2) = NonlinearProblem(f, coupled2_jacobian)
family_problem(f::ScalarSquare) = NonlinearProblem(f, scalar_square_jacobian)

family_id(::HenonHeilesMidpoint) = Int32(5)
family_id(::TanhLinear3) = Int32(4)
family_id(::Coupled2) = Int32(1)
family_id(::ScalarSquare) = Int32(3)
family_n(::HenonHeilesMidpoint) = 4
family_n(::TanhLinear3) = 3
family_n(::Coupled2) = 2
family_n(::ScalarSquare) = 1

# The parameter records of a batch: (device pointer, byte stride, shared scalars, roots to keep).
function family_records(::HenonHeilesMidpoint, p::BatchParams, n)
    Xk = p.per_system::ROCVector{SVector{4, Float64}}
    return pointer(Xk), Int64(32), (Float64(p.shared),), Xk
end
function family_records(::TanhLinear3, p::PerSystem, n)
    # records of (SMatrix{3,3} column-major, SVector{3}) are 96 bytes each
    d = p.data::ROCVector{Tuple{SMatrix{3, 3, Float64, 9}, SVector{3, Float64}}}
    return pointer(d), Int64(96), (), d
end
function family_records(::Coupled2, p::PerSystem, n)
    d = p.data::ROCVector{SVector{2, Float64}}
    return pointer(d), Int64(16), (), d
end
function family_records(::Coupled2, p::SVector{2, Float64}, n)
    d = ROCArray([p])
    return pointer(d), Int64(0), (), d
end
function family_records(::ScalarSquare, p::PerSystem, n)
    d = p.data::ROCVector{Float64}
    return pointer(d), Int64(8), (), d
end

# ---- the method and options as ABI structs ----------------------------------------------------

device_method(m) = nothing  # not served: the generic method runs
# `reuse = k` is served (the held factor of newton.jl:1781-1874); `keep_factor` keeps a factor across
# the solves of one state, and each system of a batch solves once from a fresh state.
function device_method(m::Newton{<:LUFactorization{F}, G, Nothing, RF, Float64, K}) where {
        F, G, RF, K <: Union{Nothing, Int32}}
    F === Nothing || F === Float32 || return nothing
    G <: Union{Static, Backtracking, DogLeg} || return nothing
    RF <: Union{NoRefinement, Refinement} || return nothing
    g = m.globalization
    bt = g isa Backtracking ? g : Backtracking()
    rf = m.refinement isa Refinement ? m.refinement : Refinement()
    glob = g isa Backtracking ? Int32(1) : g isa DogLeg ? Int32(3) : Int32(0)
    radius = g isa DogLeg && g.radius !== nothing ? Float64(g.radius) : 0.0
    return GsMethod(Int32(0), F === Float32 ? Int32(1) : Int32(0), glob,
        m.refinement isa Refinement ? Int32(1) : Int32(0), g isa Static ? Float64(g.α) : 1.0,
        Float64(bt.c₁), Float64(bt.p), Int32(bt.maxiter), Int32(rf.maxiter), Float64(rf.η),
        Float64(m.regularization), ntuple(_ -> 0.0, 6), m.best === Val(true) ? Int32(1) : Int32(0),
        m.reuse === nothing ? Int32(0) : m.reuse, radius, Int32(0), Int32(0))
end

function device_options(o::Options{Float64})
    return GsOptionsF64(o.f_abstol, o.f_reltol, o.x_abstol, o.x_reltol, o.min_iterations,
        o.max_iterations, o.max_stalls, o.norm isa MaxNorm ? Int32(1) : Int32(0), o.g_tol, o.kappa,
        UInt8(o.roundoff), UInt8(o.rate_test), ntuple(_ -> 0x00, 6))
end

# ---- the launch -------------------------------------------------------------------------------

function gs_error(rc)
    msg = unsafe_string(ccall((:gs_last_error, libgsolve), Cstring, ()))
    rc == -2 && throw(DimensionMismatch(msg))
    rc in (-1, -3) && throw(ArgumentError(msg))
    error("libgsolve_hip: $msg")
end

function hip_launch!(X, S, codes, fam::HIPFamily, cm::GsMethod, p, opt::Options{Float64}, order, ws)
    P, stride, shared, root = family_records(fam, p, length(X))
    cf = GsFamily(family_id(fam), Int32(family_n(fam)), Int32(family_n(fam)), Int32(0),
        ntuple(k -> k <= length(shared) ? shared[k] : 0.0, 8))
    co = device_options(opt)
    order_dev = order === nothing ? nothing : ROCArray(Int64.(collect(order)))
    GC.@preserve root order_dev begin
        rc = ccall((:gs_solve_batch_f64, libgsolve), Cint,
            (Ref{GsMethod}, Ref{GsFamily}, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Ptr{Cvoid},
                Ptr{Cvoid}, Ref{GsOptionsF64}, Int32, Ptr{Cvoid}),
            cm, cf, length(X), pointer(X), P, stride, S === nothing ? C_NULL : pointer(S),
            codes === nothing ? C_NULL : pointer(codes),
            order_dev === nothing ? C_NULL : pointer(order_dev), co, Int32(something(ws, 0)),
            stream().stream)
    end
    rc == 0 || gs_error(rc)
    return nothing
end

# The status form (batched.jl:257-287) for a device batch of a registered family.
function solve_batch!(X::ROCVector{SVector{N, Float64}}, S::ROCVector{SolverStatus{Float64}},
        prob::NonlinearProblem{<:HIPFamily}, method::Newton, p = nothing;
        order::Union{Nothing, AbstractVector{<:Integer}} = nothing,
        options::Union{Nothing, Options} = nothing,
        workgroupsize::Union{Nothing, Integer} = nothing, kwargs...) where {N}
    cm = device_method(Adapt_real(method))
    cm === nothing && return invoke(solve_batch!,
        Tuple{AbstractVector, AbstractVector{SolverStatus{Float64}}, NonlinearProblem, Any, Any},
        X, S, prob, method, p; order, options, workgroupsize, kwargs...)
    _validate_batch(X, p, order)
    _validate_batch_problem(prob)
    length(S) == length(X) || throw(DimensionMismatch(
        "status vector length $(length(S)) does not match batch length $(length(X))"))
    ws = _validate_workgroupsize(workgroupsize)
    opt = resolve_options(Float64, options, kwargs)
    isempty(X) && return S
    hip_launch!(X, S, nothing, prob.F, cm, p, opt, order, ws)
    return S
end

# The counts form (batched.jl:289-307).
function solve_batch!(X::ROCVector{SVector{N, Float64}}, prob::NonlinearProblem{<:HIPFamily},
        method::Newton, p = nothing;
        order::Union{Nothing, AbstractVector{<:Integer}} = nothing,
        options::Union{Nothing, Options} = nothing,
        workgroupsize::Union{Nothing, Integer} = nothing, kwargs...) where {N}
    cm = device_method(Adapt_real(method))
    cm === nothing && return invoke(solve_batch!,
        Tuple{AbstractVector, NonlinearProblem, Any, Any},
        X, prob, method, p; order, options, workgroupsize, kwargs...)
    _validate_batch(X, p, order)
    _validate_batch_problem(prob)
    ws = _validate_workgroupsize(workgroupsize)
    opt = resolve_options(Float64, options, kwargs)
    isempty(X) && return _tally_return_codes(ReturnCode[])
    codes = ROCArray{ReturnCode}(undef, length(X))
    hip_launch!(X, nothing, codes, prob.F, cm, p, opt, order, ws)
    return _tally_return_codes(Array(codes))
end

# Every float of the method in Float64, as `_prepare_batch` adapts it (batched.jl:162).
Adapt_real(m) = GeometricSolvers.Adapt.adapt(ToReal{Float64}(), m)

end # module
