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
                        BatchParams, LevenbergMarquardt, LeastSquaresProblem, Deflated, StationarityProblem,
                        Reparametrized,
                        _validate_batch,
                        _validate_batch_problem, _validate_workgroupsize, resolve_options,
                        _tally_return_codes, ToReal
using StaticArrays: SVector, SMatrix, @SMatrix
using LinearAlgebra: Diagonal
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
    data::Ptr{Cvoid}   # the family's own device data (A of DenseSin), C_NULL for the others
end
GsFamily(id, n, m, p_len, shared) = GsFamily(id, n, m, p_len, shared, C_NULL)

# SolverStatus{Float64} crosses the ABI as itself: its C layout is gs_status_f64.
function __init__()
    sizeof(SolverStatus{Float64}) == 72 || error("SolverStatus{Float64} is not 72 bytes")
    fieldoffset(SolverStatus{Float64}, 4) == 8 || error("SolverStatus{Float64}.fnorm is not at 8")
    fieldoffset(SolverStatus{Float64}, 10) == 32 || error("SolverStatus{Float64}.rank is not at 32")
    fieldoffset(SolverStatus{Float64}, 14) == 64 || error("SolverStatus{Float64}.contracted is not at 64")
    ccall((:gs_version, libgsolve), Int32, ()) == 2 || error("libgsolve_hip is not ABI version 2")
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

"""
    FpuGl3(β = 1.0)

The Gauss–Legendre s = 3 stage system of a 16-site FPU-β chain with fixed ends, `n = 48`, in the
stage increments `z = (z₁, z₂, z₃)` (16 each) of the position-only (Nyström) collocation form:
`F_a = z_a − h c_a p − h² Σ_b (A²)_ab g(q + z_b)`, `g(Q)_m = V′(Q_{m+1} − Q_m) − V′(Q_m − Q_{m−1})`,
`V′(r) = r + β r³`, `Q_0 = Q_17 = 0`. With `p = (h, qp)` (a `BatchParams(h, QP)`, `qp = (q, p)`
of 32 entries). The tableau constants are the device's, rounded once from 50-digit values.
"""
struct FpuGl3 <: HIPFamily
    β::Float64
end
FpuGl3() = FpuGl3(1.0)

const GL3_C = (0x1.cda042f0236e1p-4, 0x1.0000000000000p-1, 0x1.c64bf7a1fb924p-1)
const GL3_A2 = (0x1.6c16c16c16c17p-7, -0x1.0f9922995a9d1p-7, 0x1.ce741a1dacb5ep-9,
                0x1.a370da4233f7dp-4, 0x1.c71c71c71c71cp-6, -0x1.537f6b3fb1445p-8,
                0x1.bfe2a15ea5befp-3, 0x1.4f8d7b6829931p-3, 0x1.6c16c16c16c17p-7)  # (A²)_ab, row-major

# the bond r_s = Q_s − Q_{s−1} of stage b (b = 0..2, s = 0..16; the ends at zero)
@inline function gl3_bond(Q, b, s, ::Type{T}) where {T}
    hi = s < 16 ? Q[16b + s + 1] : zero(T)
    lo = s > 0 ? Q[16b + s] : zero(T)
    return hi - lo
end

function (f::FpuGl3)(z::SVector{48, T}, p) where {T}
    h, qp = p
    β = T(f.β)
    Q = SVector{48, T}(ntuple(l -> qp[(l - 1) % 16 + 1] + z[l], 48))
    dφ(r) = r + β * ((r * r) * r)
    g = SVector{48, T}(ntuple(l -> dφ(gl3_bond(Q, (l - 1) ÷ 16, (l - 1) % 16 + 1, T)) -
                                   dφ(gl3_bond(Q, (l - 1) ÷ 16, (l - 1) % 16, T)), 48))
    h2 = h * h
    return SVector{48, T}(ntuple(48) do l
        a, m = (l - 1) ÷ 16, (l - 1) % 16
        t = T(GL3_A2[3a + 1]) * g[m + 1]
        t = t + T(GL3_A2[3a + 2]) * g[16 + m + 1]
        t = t + T(GL3_A2[3a + 3]) * g[32 + m + 1]
        (z[l] - (h * T(GL3_C[a + 1])) * qp[16 + m + 1]) - h2 * t
    end)
end

# J = I − h² (A² ⊗ H), H the tridiagonal force Jacobian of each stage with stiffnesses
# K_s = 1 + 3β r_s²: the entry (a·16 + m, b·16 + k) is δ − (h² (A²)_ab) H_b(m, k).
function fpu_gl3_jacobian(f::FpuGl3, z::SVector{48, T}, p) where {T}
    h, qp = p
    Q = SVector{48, T}(ntuple(l -> qp[(l - 1) % 16 + 1] + z[l], 48))
    K(b, s) = one(T) + (3 * T(f.β)) * (gl3_bond(Q, b, s, T)^2)
    h2 = h * h
    return SMatrix{48, 48, T}(ntuple(48 * 48) do idx
        i, j = (idx - 1) % 48, (idx - 1) ÷ 48
        a, m, b, k = i ÷ 16, i % 16, j ÷ 16, j % 16
        H = k == m ? -(K(b, m) + K(b, m + 1)) : k + 1 == m ? K(b, m) : k == m + 1 ? K(b, m + 1) : zero(T)
        (i == j ? one(T) : zero(T)) - (h2 * T(GL3_A2[3a + b + 1])) * H
    end)
end

"""
    DenseSin(A)

The dense device problem of the reference's benchmarks (`TenEResidual`,
`benchmark/device_problems.jl:147-158`): `F(x, b) = A x + sin.(x)/10 − b`, the system's own `b` as
its parameter (a `PerSystem` of `SVector{n, T}`), `J = A + Diagonal(cos.(x)/10)`. `A` (n × n,
served on the device for n = 2, 4, 6, 8, 12, 16, 24, 32, 48, 64) is kept as an `SMatrix` for the
generic path and as column-major device matrices in Float64 and rounded to Float32 (the
`tenE_matrix(n, Float32)` of a Float32 batch), which the library reads through `gs_family.data`.
"""
struct DenseSin{N, L} <: HIPFamily
    A::SMatrix{N, N, Float64, L}
    Ad::ROCArray{Float64, 2}
    Ad32::ROCArray{Float32, 2}
end
function DenseSin(A::AbstractMatrix)
    n = size(A, 1)
    size(A, 2) == n || throw(DimensionMismatch("A must be square"))
    S = SMatrix{n, n, Float64}(A)
    return DenseSin(S, ROCArray(Matrix(S)), ROCArray(Float32.(Matrix(S))))
end
(f::DenseSin{N})(x::SVector{N, T}, b) where {N, T} = SMatrix{N, N, T}(f.A) * x .+ sin.(x) ./ 10 .- b
dense_sin_jacobian(f::DenseSin{N}, x::SVector{N, T}, _) where {N, T} =
    SMatrix{N, N, T}(f.A) + Diagonal(cos.(x) ./ 10)

"""The `NonlinearProblem` of a registered family, with its analytic Jacobian."""
family_problem(f::HenonHeilesMidpoint) = NonlinearProblem(f, hh_midpoint_jacobian)
family_problem(f::TanhLinear3) = NonlinearProblem(f, tanh_linear3_jacobian)
family_problem(f::Coupled2) = NonlinearProblem(f, coupled2_jacobian)
family_problem(f::ScalarSquare) = NonlinearProblem(f, scalar_square_jacobian)
family_problem(f::FpuGl3) = NonlinearProblem(f, (z, p) -> fpu_gl3_jacobian(f, z, p))
family_problem(f::DenseSin) = NonlinearProblem(f, (x, b) -> dense_sin_jacobian(f, x, b))

family_id(::HenonHeilesMidpoint) = Int32(5)
family_id(::TanhLinear3) = Int32(4)
family_id(::Coupled2) = Int32(1)
family_id(::ScalarSquare) = Int32(3)
family_id(::FpuGl3) = Int32(6)
family_id(::DenseSin) = Int32(10)
family_n(::HenonHeilesMidpoint) = 4
family_n(::TanhLinear3) = 3
family_n(::Coupled2) = 2
family_n(::ScalarSquare) = 1
family_n(::FpuGl3) = 48
family_n(::DenseSin{N}) where {N} = N

# The family's own device data (`gs_family.data`) for a batch in R: the matrix of DenseSin, none for
# the others.
family_data(::HIPFamily, ::Type) = C_NULL
family_data(f::DenseSin, ::Type{R}) where {R} = R === Float32 ? pointer(f.Ad32) : pointer(f.Ad)

# The parameter records of a batch in the working precision R: (device pointer, byte stride, shared
# scalars, roots to keep). A record is the isbits element of the device array (its size is the
# stride); a shared record travels as a one-element array at stride 0.
function family_records(::HenonHeilesMidpoint, p::BatchParams, ::Type{R}) where {R}
    Xk = p.per_system::ROCVector{SVector{4, R}}
    return pointer(Xk), Int64(sizeof(SVector{4, R})), (Float64(p.shared),), Xk
end
function family_records(::TanhLinear3, p::PerSystem, ::Type{R}) where {R}
    # records of (SMatrix{3,3} column-major, SVector{3}): 12 entries each
    d = p.data::ROCVector{Tuple{SMatrix{3, 3, R, 9}, SVector{3, R}}}
    return pointer(d), Int64(12 * sizeof(R)), (), d
end
function family_records(::Coupled2, p::PerSystem, ::Type{R}) where {R}
    d = p.data::ROCVector{SVector{2, R}}
    return pointer(d), Int64(2 * sizeof(R)), (), d
end
function family_records(::Coupled2, p::SVector{2}, ::Type{R}) where {R}
    d = ROCArray([SVector{2, R}(p)])
    return pointer(d), Int64(0), (), d
end
function family_records(::ScalarSquare, p::PerSystem, ::Type{R}) where {R}
    d = p.data::ROCVector{R}
    return pointer(d), Int64(sizeof(R)), (), d
end
function family_records(::DenseSin{N}, p::PerSystem, ::Type{R}) where {N, R}
    d = p.data::ROCVector{SVector{N, R}}
    return pointer(d), Int64(N * sizeof(R)), (), d
end
function family_records(f::FpuGl3, p::BatchParams, ::Type{Float64})
    QP = p.per_system::ROCVector{SVector{32, Float64}}
    return pointer(QP), Int64(256), (Float64(p.shared), f.β), QP
end

# What the library serves of a family beyond the method checks: the n = 48 kernel takes Static,
# Backtracking (with or without `reuse`) or DogLeg, in Float64 only; anything else falls through
# to the generic path.
served(::HIPFamily, cm, ::Type{R}) where {R} = true
served(::FpuGl3, cm, ::Type{R}) where {R} = R === Float64 && cm.globalization in (0, 1, 3)
# the dense family in Float64 or Float32: n = 2, 4, 6, 8 (and 12 in Float32) on the lane-per-system
# kernels, n = 12, 16, 24, 32, 48, 64 on the dense kernel; Static, Backtracking (with or without
# `reuse`) and DogLeg
served(::DenseSin{N}, cm, ::Type{R}) where {N, R} = (R === Float64 || R === Float32) &&
    N in (2, 4, 6, 8, 12, 16, 24, 32, 48, 64) && cm.globalization in (0, 1, 3)

# ---- the method and options as ABI structs ----------------------------------------------------

device_method(m) = nothing  # not served: the generic method runs
# `reuse = k` is served (the held factor of newton.jl:1781-1874); `keep_factor` keeps a factor across
# the solves of one state, and each system of a batch solves once from a fresh state.
function device_method(m::Newton{<:LUFactorization{F}, G, Nothing, RF, R, K}) where {
        F, G, RF, R <: Union{Float32, Float64}, K <: Union{Nothing, Int32}}
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

# Newton(; globalization = LevenbergMarquardt(...)) for the least-squares batch: the stacked Givens
# step (levenbergmarquardt.jl:359-482), which the reference takes for every linear method except an
# SVD factor without scaling (its one-SVD step rounds differently, and there is no other R2 form of
# a LeastSquaresProblem to fall through to).
function device_lm_method(m::Newton)
    g = m.globalization
    g isa LevenbergMarquardt || throw(ArgumentError(
        "the device least-squares batch takes Newton(; globalization = LevenbergMarquardt())"))
    m.reuse === nothing || throw(ArgumentError("Newton with a LevenbergMarquardt globalization takes no reuse"))
    m.jacobian === nothing || throw(ArgumentError(
        "the device least-squares batch uses the family's analytic Jacobian: pass Newton without a jacobian"))
    m.linsolve isa GeometricSolvers.SVDFactorization && g.scaling === Val(:none) && throw(ArgumentError(
        "the device least-squares batch takes the stacked Givens step: use a linear method other than SVDFactorization() or scaling = :diagonal"))
    lm = (Float64(g.μ₀), Float64(g.μ_min), Float64(g.μ_max), Float64(g.μ_decrease), Float64(g.μ_increase),
        Float64(g.ρ_accept))
    return GsMethod(Int32(0), Int32(0), Int32(2), Int32(0), 1.0, 1e-4, 0.5, Int32(100), Int32(5),
        sqrt(eps(Float64)), 0.0, lm, Int32(0), Int32(0), 0.0,
        g.scaling === Val(:diagonal) ? Int32(1) : Int32(0), Int32(0))
end

# Options{R} as the ABI's doubles (an Options{Float32} passes through exactly).
function device_options(o::Options{R}) where {R}
    return GsOptionsF64(Float64(o.f_abstol), Float64(o.f_reltol), Float64(o.x_abstol), Float64(o.x_reltol),
        o.min_iterations, o.max_iterations, o.max_stalls, o.norm isa MaxNorm ? Int32(1) : Int32(0),
        Float64(o.g_tol), Float64(o.kappa), UInt8(o.roundoff), UInt8(o.rate_test), ntuple(_ -> 0x00, 6))
end

# ---- the launch -------------------------------------------------------------------------------

function gs_error(rc)
    msg = unsafe_string(ccall((:gs_last_error, libgsolve), Cstring, ()))
    rc == -2 && throw(DimensionMismatch(msg))
    rc in (-1, -3) && throw(ArgumentError(msg))
    error("libgsolve_hip: $msg")
end

function hip_launch!(X, S, codes, fam::HIPFamily, cm::GsMethod, p, opt::Options{R}, order, ws) where {R}
    P, stride, shared, root = family_records(fam, p, R)
    cf = GsFamily(family_id(fam), Int32(family_n(fam)), Int32(family_n(fam)), Int32(0),
        ntuple(k -> k <= length(shared) ? shared[k] : 0.0, 8), family_data(fam, R))
    co = device_options(opt)
    order_dev = order === nothing ? nothing : ROCArray(Int64.(collect(order)))
    Sp = S === nothing ? C_NULL : pointer(S)
    Cp = codes === nothing ? C_NULL : pointer(codes)
    Op = order_dev === nothing ? C_NULL : pointer(order_dev)
    GC.@preserve root order_dev fam begin
        rc = if R === Float32
            # the Float32 twin (SolverStatus{Float32}, Float32 records and iterates)
            ccall((:gs_solve_batch_f32, libgsolve), Cint,
                (Ref{GsMethod}, Ref{GsFamily}, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Ptr{Cvoid},
                    Ptr{Cvoid}, Ref{GsOptionsF64}, Int32, Ptr{Cvoid}),
                cm, cf, length(X), pointer(X), P, stride, Sp, Cp, Op, co, Int32(something(ws, 0)),
                stream().stream)
        else
            ccall((:gs_solve_batch_f64, libgsolve), Cint,
                (Ref{GsMethod}, Ref{GsFamily}, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Ptr{Cvoid},
                    Ptr{Cvoid}, Ref{GsOptionsF64}, Int32, Ptr{Cvoid}),
                cm, cf, length(X), pointer(X), P, stride, Sp, Cp, Op, co, Int32(something(ws, 0)),
                stream().stream)
        end
    end
    rc == 0 || gs_error(rc)
    return nothing
end

# The status form (batched.jl:257-287) for a device batch of a registered family, in Float64 or
# Float32 (the working precision of the iterates; S holds SolverStatus of the same precision).
function solve_batch!(X::ROCVector{SVector{N, T}}, S::ROCVector{SolverStatus{T}},
        prob::NonlinearProblem{<:HIPFamily}, method::Newton, p = nothing;
        order::Union{Nothing, AbstractVector{<:Integer}} = nothing,
        options::Union{Nothing, Options} = nothing,
        workgroupsize::Union{Nothing, Integer} = nothing, kwargs...) where {N, T <: Union{Float32, Float64}}
    cm = device_method(Adapt_real(method, T))
    (cm === nothing || !served(prob.F, cm, T)) && return invoke(solve_batch!,
        Tuple{AbstractVector, AbstractVector{SolverStatus{T}}, NonlinearProblem, Any, Any},
        X, S, prob, method, p; order, options, workgroupsize, kwargs...)
    _validate_batch(X, p, order)
    _validate_batch_problem(prob)
    length(S) == length(X) || throw(DimensionMismatch(
        "status vector length $(length(S)) does not match batch length $(length(X))"))
    ws = _validate_workgroupsize(workgroupsize)
    opt = resolve_options(T, options, kwargs)
    isempty(X) && return S
    hip_launch!(X, S, nothing, prob.F, cm, p, opt, order, ws)
    return S
end

# The counts form (batched.jl:289-307).
function solve_batch!(X::ROCVector{SVector{N, T}}, prob::NonlinearProblem{<:HIPFamily},
        method::Newton, p = nothing;
        order::Union{Nothing, AbstractVector{<:Integer}} = nothing,
        options::Union{Nothing, Options} = nothing,
        workgroupsize::Union{Nothing, Integer} = nothing, kwargs...) where {N, T <: Union{Float32, Float64}}
    cm = device_method(Adapt_real(method, T))
    (cm === nothing || !served(prob.F, cm, T)) && return invoke(solve_batch!,
        Tuple{AbstractVector, NonlinearProblem, Any, Any},
        X, prob, method, p; order, options, workgroupsize, kwargs...)
    _validate_batch(X, p, order)
    _validate_batch_problem(prob)
    ws = _validate_workgroupsize(workgroupsize)
    opt = resolve_options(T, options, kwargs)
    isempty(X) && return _tally_return_codes(ReturnCode[])
    codes = ROCArray{ReturnCode}(undef, length(X))
    hip_launch!(X, nothing, codes, prob.F, cm, p, opt, order, ws)
    return _tally_return_codes(Array(codes))
end

# Every float of the method in the working precision T, as `_prepare_batch` adapts it (batched.jl:162).
Adapt_real(m, ::Type{T} = Float64) where {T} = GeometricSolvers.Adapt.adapt(ToReal{T}(), m)

# ---- the Deflated batch (deflation.jl:714-760) ------------------------------------------------

struct GsDeflation
    nroots::Int32
    pad::Int32
    power::Float64
    shift::Float64
    roots::NTuple{32, Float64}
end

# The roots as the flat record of gs_deflation, or `nothing` where the device form does not hold
# them (more than 4 roots or 32 entries, a power that is not an integer in 1:64); such a Deflated,
# and every method the device form does not take (best = true, a Float32 batch), runs the
# reference's generic batch, which also raises the reference's own errors.
function device_deflation(def::Deflated, n::Int)
    nr = length(def.roots)
    (1 <= nr <= 4 && nr * n <= 32) || return nothing
    p = Float64(def.power)
    (isinteger(p) && 1 <= p <= 64) || return nothing
    flat = Float64[]
    for r in def.roots
        length(r) == n || return nothing
        append!(flat, Float64.(collect(r)))
    end
    return GsDeflation(Int32(nr), Int32(0), p, Float64(def.shift),
        ntuple(k -> k <= length(flat) ? flat[k] : 0.0, 32))
end

served_deflated(fam::HIPFamily, cm, ::Type{T}) where {T} = !(fam isa Union{FpuGl3, DenseSin}) && T === Float64 &&
    cm !== nothing && cm.globalization in (0, 1, 3) && (cm.reuse == 0 || cm.globalization != 3) && cm.best == 0

function hip_launch_deflated!(X, S, codes, def::Deflated, cd::GsDeflation, cm::GsMethod, p, opt, order, ws)
    fam = def.prob.F
    P, stride, shared, root = family_records(fam, p, Float64)
    cf = GsFamily(family_id(fam), Int32(family_n(fam)), Int32(family_n(fam)), Int32(0),
        ntuple(k -> k <= length(shared) ? shared[k] : 0.0, 8))
    order_dev = order === nothing ? nothing : ROCArray(Int64.(collect(order)))
    GC.@preserve root order_dev begin
        rc = ccall((:gs_solve_batch_deflated_f64, libgsolve), Cint,
            (Ref{GsMethod}, Ref{GsFamily}, Ref{GsDeflation}, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Int64,
                Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{GsOptionsF64}, Int32, Ptr{Cvoid}),
            cm, cf, cd, length(X), pointer(X), P, stride, S === nothing ? C_NULL : pointer(S),
            codes === nothing ? C_NULL : pointer(codes), order_dev === nothing ? C_NULL : pointer(order_dev),
            device_options(opt), Int32(something(ws, 0)), stream().stream)
    end
    rc == 0 || gs_error(rc)
    return nothing
end

# The status form of the Deflated batch.
function solve_batch!(X::ROCVector{SVector{N, T}}, S::ROCVector{SolverStatus{T}},
        def::Deflated{<:NonlinearProblem{<:HIPFamily}}, method::Newton, p = nothing;
        order::Union{Nothing, AbstractVector{<:Integer}} = nothing,
        options::Union{Nothing, Options} = nothing,
        workgroupsize::Union{Nothing, Integer} = nothing, kwargs...) where {N, T <: Union{Float32, Float64}}
    cm = device_method(Adapt_real(method, T))
    cd = device_deflation(def, N)
    (cd === nothing || !served_deflated(def.prob.F, cm, T)) && return invoke(solve_batch!,
        Tuple{AbstractVector, AbstractVector{SolverStatus{T}}, Deflated{<:NonlinearProblem}, Any, Any},
        X, S, def, method, p; order, options, workgroupsize, kwargs...)
    _validate_batch(X, p, order)
    length(S) == length(X) || throw(DimensionMismatch(
        "status vector length $(length(S)) does not match batch length $(length(X))"))
    ws = _validate_workgroupsize(workgroupsize)
    opt = resolve_options(T, options, kwargs)
    isempty(X) && return S
    hip_launch_deflated!(X, S, nothing, def, cd, cm, p, opt, order, ws)
    return S
end

# The counts form of the Deflated batch.
function solve_batch!(X::ROCVector{SVector{N, T}}, def::Deflated{<:NonlinearProblem{<:HIPFamily}},
        method::Newton, p = nothing;
        order::Union{Nothing, AbstractVector{<:Integer}} = nothing,
        options::Union{Nothing, Options} = nothing,
        workgroupsize::Union{Nothing, Integer} = nothing, kwargs...) where {N, T <: Union{Float32, Float64}}
    cm = device_method(Adapt_real(method, T))
    cd = device_deflation(def, N)
    (cd === nothing || !served_deflated(def.prob.F, cm, T)) && return invoke(solve_batch!,
        Tuple{AbstractVector, Deflated{<:NonlinearProblem}, Any, Any},
        X, def, method, p; order, options, workgroupsize, kwargs...)
    _validate_batch(X, p, order)
    ws = _validate_workgroupsize(workgroupsize)
    opt = resolve_options(T, options, kwargs)
    isempty(X) && return _tally_return_codes(ReturnCode[])
    codes = ROCArray{ReturnCode}(undef, length(X))
    hip_launch_deflated!(X, nothing, codes, def, cd, cm, p, opt, order, ws)
    return _tally_return_codes(Array(codes))
end

# ---- the Reparametrized batch (reparametrized.jl:981-1031) -------------------------------------

"""
    LinearChart(B::SMatrix{N, N, Float64}, c::SVector{N, Float64})
    PolarChart()

The registered charts of a device `Reparametrized` batch: `x = B y + c` (`∂g = B`), and the polar
chart `x = (r cos θ, r sin θ)` of the reference's tests (n = 2). Both are callables `g(y, p)`, so a
`Reparametrized(prob, LinearChart(B, c))` is also what the reference's generic path runs.
"""
struct LinearChart{N, L}
    B::SMatrix{N, N, Float64, L}
    c::SVector{N, Float64}
end
(g::LinearChart)(y, p) = g.B * y + g.c
struct PolarChart end
(::PolarChart)(y, p) = SVector(y[1] * cos(y[2]), y[1] * sin(y[2]))

struct GsChart
    kind::Int32
    pad::Int32
    B::NTuple{16, Float64}
    c::NTuple{4, Float64}
end

device_chart(g::LinearChart{N}) where {N} = GsChart(Int32(1), Int32(0),
    ntuple(k -> k <= N * N ? g.B[(k - 1) ÷ N + 1, (k - 1) % N + 1] : 0.0, 16),
    ntuple(k -> k <= N ? g.c[k] : 0.0, 4))
device_chart(::PolarChart) = GsChart(Int32(2), Int32(0), ntuple(_ -> 0.0, 16), ntuple(_ -> 0.0, 4))

const HIPReparam = Reparametrized{<:NonlinearProblem{<:HIPFamily}, <:Union{LinearChart, PolarChart}}

# A Float32 batch takes the linear chart only (its B and c are used as Float32, as ToReal{Float32}
# adapts them); the polar chart in Float32 runs the reference's generic path.
served_reparam(rep, cm, ::Type{T} = Float64) where {T} = !rep.covariant &&
    !(rep.prob.F isa Union{FpuGl3, DenseSin}) && cm !== nothing &&
    cm.globalization in (0, 1, 3) && (cm.reuse == 0 || cm.globalization != 3) &&
    (!(rep.g isa PolarChart) || (T === Float64 && family_n(rep.prob.F) == 2))

function hip_launch_reparam!(Y, S, codes, rep, cm::GsMethod, p, opt::Options{R}, order, ws) where {R}
    fam = rep.prob.F
    P, stride, shared, root = family_records(fam, p, R)
    cf = GsFamily(family_id(fam), Int32(family_n(fam)), Int32(family_n(fam)), Int32(0),
        ntuple(k -> k <= length(shared) ? shared[k] : 0.0, 8))
    order_dev = order === nothing ? nothing : ROCArray(Int64.(collect(order)))
    Sp = S === nothing ? C_NULL : pointer(S)
    Cp = codes === nothing ? C_NULL : pointer(codes)
    Op = order_dev === nothing ? C_NULL : pointer(order_dev)
    GC.@preserve root order_dev begin
        rc = if R === Float32
            ccall((:gs_solve_batch_reparametrized_f32, libgsolve), Cint,
                (Ref{GsMethod}, Ref{GsFamily}, Ref{GsChart}, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Int64,
                    Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{GsOptionsF64}, Int32, Ptr{Cvoid}),
                cm, cf, device_chart(rep.g), length(Y), pointer(Y), P, stride, Sp, Cp, Op,
                device_options(opt), Int32(something(ws, 0)), stream().stream)
        else
            ccall((:gs_solve_batch_reparametrized_f64, libgsolve), Cint,
                (Ref{GsMethod}, Ref{GsFamily}, Ref{GsChart}, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Int64,
                    Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{GsOptionsF64}, Int32, Ptr{Cvoid}),
                cm, cf, device_chart(rep.g), length(Y), pointer(Y), P, stride, Sp, Cp, Op,
                device_options(opt), Int32(something(ws, 0)), stream().stream)
        end
    end
    rc == 0 || gs_error(rc)
    return nothing
end

# The status form: Y holds y₀ and receives x* = g(y*); Float64 or Float32.
function solve_batch!(Y::ROCVector{SVector{N, T}}, S::ROCVector{SolverStatus{T}},
        rep::HIPReparam, method::Newton = Newton(), p = nothing;
        order::Union{Nothing, AbstractVector{<:Integer}} = nothing,
        options::Union{Nothing, Options} = nothing,
        workgroupsize::Union{Nothing, Integer} = nothing, kwargs...) where {N, T <: Union{Float32, Float64}}
    cm = device_method(Adapt_real(method, T))
    served_reparam(rep, cm, T) || return invoke(solve_batch!,
        Tuple{AbstractVector{<:Union{SVector, Number}}, AbstractVector{SolverStatus{T}}, Reparametrized, Any, Any},
        Y, S, rep, method, p; order, options, workgroupsize, kwargs...)
    length(S) == length(Y) || throw(DimensionMismatch(
        "status vector length $(length(S)) does not match batch length $(length(Y))"))
    _validate_batch(Y, p, order)
    _validate_batch_problem(rep)
    ws = _validate_workgroupsize(workgroupsize)
    opt = resolve_options(T, options, kwargs)
    isempty(Y) && return S
    hip_launch_reparam!(Y, S, nothing, rep, cm, p, opt, order, ws)
    return S
end

# The counts form.
function solve_batch!(Y::ROCVector{SVector{N, T}}, rep::HIPReparam, method::Newton = Newton(), p = nothing;
        order::Union{Nothing, AbstractVector{<:Integer}} = nothing,
        options::Union{Nothing, Options} = nothing,
        workgroupsize::Union{Nothing, Integer} = nothing, kwargs...) where {N, T <: Union{Float32, Float64}}
    cm = device_method(Adapt_real(method, T))
    served_reparam(rep, cm, T) || return invoke(solve_batch!,
        Tuple{AbstractVector{<:Union{SVector, Number}}, Reparametrized, Any, Any},
        Y, rep, method, p; order, options, workgroupsize, kwargs...)
    _validate_batch(Y, p, order)
    _validate_batch_problem(rep)
    ws = _validate_workgroupsize(workgroupsize)
    opt = resolve_options(T, options, kwargs)
    isempty(Y) && return _tally_return_codes(ReturnCode[])
    codes = ROCArray{ReturnCode}(undef, length(Y))
    hip_launch_reparam!(Y, nothing, codes, rep, cm, p, opt, order, ws)
    return _tally_return_codes(Array(codes))
end

# ---- the StationarityProblem batch (stationarity.jl:642-698) ----------------------------------

"""
    HenonHeilesAction()

The two-step discrete action of the variational midpoint rule for Hénon–Heiles on the full state
`x = (q₋, q, q₊)`: `S(x, (h, λ)) = L_d(q₋, q) + L_d(q, q₊)`, `L_d(a, b) = |b − a|²/(2h) −
h V((a + b)/2)`, `V(q) = (q₁² + q₂²)/2 + λ (q₁² q₂ − q₂³/3)`. `StationarityProblem(HenonHeilesAction();
E = (3, 4), U = (5, 6))` gives the discrete Euler–Lagrange equations for `q₊`. The device uses the
closed forms of ∇S and ∇²S; on the host the reference differentiates this `S` itself.
"""
struct HenonHeilesAction end

function (::HenonHeilesAction)(x::SVector{6, T}, p) where {T}
    h, λ = p
    V(q1, q2) = (q1^2 + q2^2) / 2 + λ * (q1^2 * q2 - q2^3 / 3)
    Ld(a1, a2, b1, b2) = ((b1 - a1)^2 + (b2 - a2)^2) / (2h) - h * V((a1 + b1) / 2, (a2 + b2) / 2)
    return Ld(x[1], x[2], x[3], x[4]) + Ld(x[3], x[4], x[5], x[6])
end

struct GsStationarity
    n_eq::Int32
    b_kind::Int32
    e::NTuple{8, Int32}
    u::NTuple{8, Int32}
    b_shared::NTuple{8, Float64}
    b_data::Ptr{Cvoid}
    b_stride::Int64
end

_stat_indices(::Colon) = ()
_stat_indices(idx) = Tuple(Int32.(collect(idx)))

# The index sets and the datum as gs_stationarity, or `nothing` where the device form does not hold
# them (|E| other than 2 or 6, a callable datum): those run the reference's generic batch. A shared
# datum travels as doubles (the library uses it as T); a PerSystem datum holds SVector{m, T}.
function device_stationarity(prob::StationarityProblem, b_src, m::Int, ::Type{T} = Float64) where {T}
    m in (2, 6) || return nothing
    e, u = _stat_indices(prob.E), _stat_indices(prob.U)
    pad(t) = ntuple(k -> k <= length(t) ? t[k] : Int32(0), 8)
    kind, shared, data, stride = Int32(0), ntuple(_ -> 0.0, 8), C_NULL, Int64(0)
    if b_src === nothing
    elseif b_src isa Number
        kind, shared = Int32(1), ntuple(k -> k <= m ? Float64(b_src) : 0.0, 8)
    elseif b_src isa AbstractVector{<:Real}
        length(b_src) == m || throw(ArgumentError("datum b must have length $m, got $(length(b_src))"))
        v = Float64.(collect(b_src))
        kind, shared = Int32(1), ntuple(k -> k <= m ? v[k] : 0.0, 8)
    elseif b_src isa PerSystem && b_src.data isa ROCVector{<:SVector{M, T} where {M}}
        length(eltype(b_src.data)) == m || throw(ArgumentError(
            "datum b must have length $m, got $(length(eltype(b_src.data)))"))
        kind, data, stride = Int32(2), Ptr{Cvoid}(pointer(b_src.data)), Int64(sizeof(T) * m)
    else
        return nothing
    end
    return GsStationarity(Int32(length(e)), kind, pad(e), pad(u), shared, data, stride)
end

_action_records(p::Tuple{Real, Real}, ::Type{T}) where {T} = (ROCArray([T(p[1]), T(p[2])]), Int64(0))
_action_records(p::SVector{2, <:Real}, ::Type{T}) where {T} = _action_records((p[1], p[2]), T)
_action_records(p::PerSystem{<:ROCVector{SVector{2, T}}}, ::Type{T}) where {T} = (p.data, Int64(2 * sizeof(T)))

# The action's parameters the device reads: (h, λ) shared, or one SVector{2, T} per system in device
# memory, T the batch's precision. Anything else (a host vector, another record type) is no device
# launch: it goes to the generic method, which refuses or runs it as the reference does.
served_action_params(p::Tuple{Real, Real}, ::Type) = true
served_action_params(p::SVector{2, <:Real}, ::Type) = true
served_action_params(p::PerSystem{<:ROCVector{SVector{2, T}}}, ::Type{T}) where {T} = true
served_action_params(p, ::Type) = false

function hip_launch_stationarity!(X, S, codes, cs::GsStationarity, cm::GsMethod, p, opt::Options{R}, order,
        ws) where {R}
    P, stride = _action_records(p, R)
    cf = GsFamily(Int32(8), Int32(6), Int32(6), Int32(2), ntuple(_ -> 0.0, 8))
    order_dev = order === nothing ? nothing : ROCArray(Int64.(collect(order)))
    Sp = S === nothing ? C_NULL : pointer(S)
    Cp = codes === nothing ? C_NULL : pointer(codes)
    Op = order_dev === nothing ? C_NULL : pointer(order_dev)
    GC.@preserve P order_dev begin
        rc = if R === Float32
            ccall((:gs_solve_batch_stationarity_f32, libgsolve), Cint,
                (Ref{GsMethod}, Ref{GsFamily}, Ref{GsStationarity}, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Int64,
                    Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{GsOptionsF64}, Int32, Ptr{Cvoid}),
                cm, cf, cs, length(X), pointer(X), pointer(P), stride, Sp, Cp, Op,
                device_options(opt), Int32(something(ws, 0)), stream().stream)
        else
            ccall((:gs_solve_batch_stationarity_f64, libgsolve), Cint,
                (Ref{GsMethod}, Ref{GsFamily}, Ref{GsStationarity}, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Int64,
                    Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{GsOptionsF64}, Int32, Ptr{Cvoid}),
                cm, cf, cs, length(X), pointer(X), pointer(P), stride, Sp, Cp, Op,
                device_options(opt), Int32(something(ws, 0)), stream().stream)
        end
    end
    rc == 0 || gs_error(rc)
    return nothing
end

served_stationarity(cm) = cm !== nothing && cm.globalization in (0, 1, 3)

# The status form: X holds each system's full state (SVector{6, T}, T Float64 or Float32); p = (h, λ)
# shared or PerSystem of SVector{2, T}; the datum from the keyword `b` or the problem's own.
function solve_batch!(X::ROCVector{SVector{6, T}}, S::ROCVector{SolverStatus{T}},
        prob::StationarityProblem{HenonHeilesAction}, method::Newton = Newton(), p = nothing;
        order::Union{Nothing, AbstractVector{<:Integer}} = nothing,
        options::Union{Nothing, Options} = nothing, b = nothing,
        workgroupsize::Union{Nothing, Integer} = nothing, kwargs...) where {T <: Union{Float32, Float64}}
    b_src = b === nothing ? prob.b : b
    cm = device_method(Adapt_real(method, T))
    cs = device_stationarity(prob, b_src, prob.E isa Colon ? 6 : length(prob.E), T)
    (cs === nothing || !served_stationarity(cm) || !served_action_params(p, T)) && return invoke(solve_batch!,
        Tuple{AbstractArray, AbstractArray{SolverStatus{T}}, StationarityProblem, Any, Any},
        X, S, prob, method, p; order, options, b, workgroupsize, kwargs...)
    _validate_batch(X, p, order)
    b_src isa PerSystem && _validate_batch(X, b_src, nothing)
    length(S) == length(X) || throw(DimensionMismatch(
        "status vector S has length $(length(S)), expected $(length(X)) to match the batch X"))
    ws = _validate_workgroupsize(workgroupsize)
    opt = resolve_options(T, options, kwargs)
    isempty(X) && return S
    GC.@preserve b_src hip_launch_stationarity!(X, S, nothing, cs, cm, p, opt, order, ws)
    return S
end

# The counts form.
function solve_batch!(X::ROCVector{SVector{6, T}}, prob::StationarityProblem{HenonHeilesAction},
        method::Newton = Newton(), p = nothing;
        order::Union{Nothing, AbstractVector{<:Integer}} = nothing,
        options::Union{Nothing, Options} = nothing, b = nothing,
        workgroupsize::Union{Nothing, Integer} = nothing, kwargs...) where {T <: Union{Float32, Float64}}
    b_src = b === nothing ? prob.b : b
    cm = device_method(Adapt_real(method, T))
    cs = device_stationarity(prob, b_src, prob.E isa Colon ? 6 : length(prob.E), T)
    (cs === nothing || !served_stationarity(cm) || !served_action_params(p, T)) && return invoke(solve_batch!,
        Tuple{AbstractArray, StationarityProblem, Any, Any},
        X, prob, method, p; order, options, b, workgroupsize, kwargs...)
    _validate_batch(X, p, order)
    b_src isa PerSystem && _validate_batch(X, b_src, nothing)
    ws = _validate_workgroupsize(workgroupsize)
    opt = resolve_options(T, options, kwargs)
    isempty(X) && return _tally_return_codes(ReturnCode[])
    codes = ROCArray{ReturnCode}(undef, length(X))
    GC.@preserve b_src hip_launch_stationarity!(X, nothing, codes, cs, cm, p, opt, order, ws)
    return _tally_return_codes(Array(codes))
end

# ---- the least-squares batch (an R2 form the reference does not have) ------------------------

"""
    ExpSumFits()

Exponential-sum least-squares fits `y(t) ≈ Σ_{k≤K} a_k exp(−λ_k t)` at `t_j = 5j/(m−1)`, with
`x = (a_1..a_K, λ_1..λ_K)` in the first `2K` of the 8 entries of a fit's `SVector{8, Float64}`;
`LeastSquaresProblem(ExpSumFits())` is the problem of `gs_solve_expsum_lsq_f64`.
"""
struct ExpSumFits end

"""
    ExpSumData(K, m, y_offset, Y)

The per-fit data of an `ExpSumFits` batch on the device: `K::ROCVector{Int32}` (1..4),
`m::ROCVector{Int32}` (samples, ≥ 2K), `y_offset::ROCVector{Int64}` and the concatenated samples
`Y::ROCVector{Float64}`; fit i's samples are `Y[y_offset[i] + 1 : y_offset[i] + m[i]]`.
"""
struct ExpSumData{VK <: ROCVector{Int32}, VM <: ROCVector{Int32}, VO <: ROCVector{Int64}, VY <: ROCVector{Float64}}
    K::VK
    m::VM
    y_offset::VO
    Y::VY
end
Base.length(d::ExpSumData) = length(d.K)

# The status form: every fit's R1 solve (solver.jl:148-189, 277-378) with Newton + LevenbergMarquardt,
# in one call; the fits are bucketed by size on the device, so there is no `order` or workgroup size.
function solve_batch!(X::ROCVector{SVector{8, Float64}}, S::ROCVector{SolverStatus{Float64}},
        prob::LeastSquaresProblem{ExpSumFits}, method::Newton, data::ExpSumData;
        options::Union{Nothing, Options} = nothing, kwargs...)
    cm = device_lm_method(Adapt_real(method))
    length(data) == length(X) || throw(DimensionMismatch(
        "ExpSumData length $(length(data)) does not match batch length $(length(X))"))
    length(S) == length(X) || throw(DimensionMismatch(
        "status vector length $(length(S)) does not match batch length $(length(X))"))
    opt = resolve_options(Float64, options, kwargs)
    isempty(X) && return S
    co = device_options(opt)
    rc = ccall((:gs_solve_expsum_lsq_f64, libgsolve), Cint,
        (Ref{GsMethod}, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64,
            Ptr{Cvoid}, Ref{GsOptionsF64}, Ptr{Cvoid}),
        cm, length(X), pointer(X), pointer(data.K), pointer(data.m), pointer(data.y_offset),
        pointer(data.Y), length(data.Y), pointer(S), co, stream().stream)
    rc == 0 || gs_error(rc)
    return S
end

# The counts form.
function solve_batch!(X::ROCVector{SVector{8, Float64}}, prob::LeastSquaresProblem{ExpSumFits},
        method::Newton, data::ExpSumData; kwargs...)
    S = ROCArray{SolverStatus{Float64}}(undef, length(X))
    solve_batch!(X, S, prob, method, data; kwargs...)
    return _tally_return_codes([s.code for s in Array(S)])
end

# ---- the learned start (configuration 5; no counterpart in the reference) ---------------------

const MLP_IN, MLP_HID, MLP_OUT = 32, 128, 16
const MLP_NPARAM = 39312
const MLP_NPACKED = 38912

# gs_mlp of include/gsolve.h
struct GsMlp
    params::Ptr{Float32}
    packed::Ptr{UInt16}
    in_shift::NTuple{8, Float32}
    in_scale::NTuple{8, Float32}
    out_shift::NTuple{4, Float32}
    out_scale::NTuple{4, Float32}
end

"""
    LearnedStart(; seed = 0)

A multilayer perceptron (3 × 128 ReLU layers, BF16 operands on MFMA) that predicts the Newton start
of the `HenonHeilesMidpoint` step as a correction to the explicit midpoint step, trained on the GPU
by `train!` on solutions the method generates, and used by `solve_batch_learned!`. The parameters
(FP32), their packed BF16 copy, the Adam moments and the normalisation live on the device.
"""
mutable struct LearnedStart
    params::ROCVector{Float32}
    packed::ROCVector{UInt16}
    m::ROCVector{Float32}
    v::ROCVector{Float32}
    grad::ROCVector{Float32}
    loss::ROCVector{Float32}
    step::Int
    in_shift::NTuple{8, Float32}
    in_scale::NTuple{8, Float32}
    out_shift::NTuple{4, Float32}
    out_scale::NTuple{4, Float32}
end

# A small generator of its own (splitmix64; the package does not depend on Random): uniform draws
# in [0, 1), normal draws by Box–Muller, and Fisher–Yates permutations.
mutable struct SplitMix
    s::UInt64
end
function next!(g::SplitMix)
    g.s += 0x9e3779b97f4a7c15
    z = g.s
    z = (z ⊻ (z >> 30)) * 0xbf58476d1ce4e5b9
    z = (z ⊻ (z >> 27)) * 0x94d049bb133111eb
    return z ⊻ (z >> 31)
end
uniform!(g::SplitMix) = Float64(next!(g) >> 11) * 0x1.0p-53
randn!(g::SplitMix) = sqrt(-2 * log(1 - uniform!(g))) * cos(2π * uniform!(g))
function randperm!(g::SplitMix, n)
    v = collect(1:n)
    for i in n:-1:2
        j = Int(next!(g) % UInt64(i)) + 1
        v[i], v[j] = v[j], v[i]
    end
    return v
end

# He-normal hidden weights (fan-in 5 for the first layer, whose other inputs are padding), a small
# output layer, zero biases; each W row-major, inputs × outputs, in the order W1 b1 W2 b2 W3 b3 W4 b4.
function mlp_initial_parameters(seed)
    rng = SplitMix(UInt64(seed))
    P = zeros(Float32, MLP_NPARAM)
    w2 = MLP_IN * MLP_HID + MLP_HID
    w3 = w2 + MLP_HID * MLP_HID + MLP_HID
    w4 = w3 + MLP_HID * MLP_HID + MLP_HID
    for i in 0:4, j in 0:(MLP_HID - 1)
        P[i * MLP_HID + j + 1] = Float32(randn!(rng) * sqrt(2 / 5))
    end
    for off in (w2, w3), i in 0:(MLP_HID - 1), j in 0:(MLP_HID - 1)
        P[off + i * MLP_HID + j + 1] = Float32(randn!(rng) * sqrt(2 / MLP_HID))
    end
    for i in 0:(MLP_HID - 1), j in 0:3
        P[w4 + i * MLP_OUT + j + 1] = Float32(randn!(rng) * sqrt(1 / MLP_HID) * 0.1)
    end
    return P
end

function mlp_check(rc)
    rc == 0 || gs_error(rc)
    return nothing
end

function LearnedStart(; seed::Integer = 0)
    net = LearnedStart(ROCArray(mlp_initial_parameters(seed)), ROCArray(zeros(UInt16, MLP_NPACKED)),
        ROCArray(zeros(Float32, MLP_NPARAM)), ROCArray(zeros(Float32, MLP_NPARAM)),
        ROCArray(zeros(Float32, MLP_NPARAM)), ROCArray(zeros(Float32, 1)), 0,
        ntuple(_ -> 0.0f0, 8), ntuple(_ -> 1.0f0, 8), ntuple(_ -> 0.0f0, 4), ntuple(_ -> 1.0f0, 4))
    mlp_check(ccall((:gs_mlp_pack, libgsolve), Cint, (Ptr{Float32}, Ptr{UInt16}, Ptr{Cvoid}),
        pointer(net.params), pointer(net.packed), stream().stream))
    return net
end

gs_mlp(net::LearnedStart) = GsMlp(pointer(net.params), pointer(net.packed), net.in_shift, net.in_scale,
    net.out_shift, net.out_scale)

hh_field(x) = SVector(x[3], x[4], -x[1] - 2 * x[1] * x[2], (-x[2] - x[1] * x[1]) + x[2] * x[2])

"""
    fit_normalization!(net, Xk, X1, h)

The features' and targets' shifts and scales (means and standard deviations) from training pairs:
states `Xk` and the method's solutions `X1` of their midpoint steps (device vectors of
`SVector{4, Float64}`).
"""
function fit_normalization!(net::LearnedStart, Xk::ROCVector{SVector{4, Float64}},
        X1::ROCVector{SVector{4, Float64}}, h::Real)
    xk, x1 = Array(Xk), Array(X1)
    corr = map(xk, x1) do x, y
        xt = x + h * hh_field(x + (h / 2) * hh_field(x))
        (y - xt) / (h * h * h)
    end
    msd(v, j) = (μ = sum(e[j] for e in v) / length(v); σ = sqrt(sum((e[j] - μ)^2 for e in v) / length(v)); (μ, σ))
    s_in = [msd(xk, j) for j in 1:4]
    s_out = [msd(corr, j) for j in 1:4]
    net.in_shift = ntuple(j -> j <= 4 ? Float32(s_in[j][1]) : (j == 5 ? Float32(h) : 0.0f0), 8)
    net.in_scale = ntuple(j -> j <= 4 ? (s_in[j][2] > 0 ? Float32(1 / s_in[j][2]) : 1.0f0) : (j == 5 ? 1.0f0 : 0.0f0), 8)
    net.out_shift = ntuple(j -> Float32(s_out[j][1]), 4)
    net.out_scale = ntuple(j -> s_out[j][2] > 0 ? Float32(s_out[j][2]) : 1.0f0, 4)
    return net
end

"""
    train!(net, Xk, X1, h; epochs = 8, batch = 4096, lr = 3e-3, lr_final = 3e-5, seed = 0)

Minibatch Adam with a cosine learning-rate schedule over shuffled epochs, on the device's forward and
backward kernels (`gs_mlp_grad`, `gs_mlp_adam`); returns the mean loss of each epoch.
"""
function train!(net::LearnedStart, Xk::ROCVector{SVector{4, Float64}}, X1::ROCVector{SVector{4, Float64}},
        h::Real; epochs::Integer = 8, batch::Integer = 4096, lr::Real = 3e-3, lr_final::Real = 3e-5,
        seed::Integer = 0)
    n = length(Xk)
    batch % 32 == 0 || throw(ArgumentError("batch must be a multiple of 32"))
    steps = n ÷ batch
    steps ≥ 1 || throw(ArgumentError("fewer training pairs than one batch"))
    rng = SplitMix(UInt64(seed) + 0x5eed)
    total = epochs * steps
    history = Float64[]
    t = 0
    for _ in 1:epochs
        idx = ROCArray(Int64.(randperm!(rng, n) .- 1))
        acc = 0.0
        for k in 0:(steps - 1)
            GC.@preserve idx begin
                mlp_check(ccall((:gs_mlp_grad, libgsolve), Cint,
                    (Ref{GsMlp}, Ptr{Cvoid}, Ptr{Cvoid}, Float64, Ptr{Int64}, Int32, Ptr{Float32}, Ptr{Float32}, Ptr{Cvoid}),
                    gs_mlp(net), pointer(Xk), pointer(X1), Float64(h), pointer(idx) + 8 * k * batch, Int32(batch),
                    pointer(net.grad), pointer(net.loss), stream().stream))
            end
            acc += Float64(Array(net.loss)[1])
            rate = lr_final + (lr - lr_final) * (1 + cos(π * t / max(total - 1, 1))) / 2
            net.step += 1
            mlp_check(ccall((:gs_mlp_adam, libgsolve), Cint,
                (Ptr{Float32}, Ptr{UInt16}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Int64, Float32, Float32,
                    Float32, Float32, Ptr{Cvoid}),
                pointer(net.params), pointer(net.packed), pointer(net.grad), pointer(net.m), pointer(net.v),
                Int64(net.step), Float32(rate), 0.9f0, 0.999f0, 1.0f-8, stream().stream))
            t += 1
        end
        push!(history, acc / steps)
    end
    return history
end

"""
    solve_batch_learned!(X, S, prob::NonlinearProblem{HenonHeilesMidpoint}, method, p::BatchParams, net;
                         options = nothing, kwargs...)

`solve_batch!` from the learned start: for each system the start x̂₊ of the network, `method` from
x̂₊, and plain Newton from x_k for every system whose solve from x̂₊ does not return SUCCESS (that
solve's status is then the system's), in one fused kernel. `method` is Newton with
`LUFactorization()` or `LUFactorization(Float32)`, `Static`, with or without `Refinement`. Returns the
number of systems that fell back.
"""
function solve_batch_learned!(X::ROCVector{SVector{4, Float64}}, S::ROCVector{SolverStatus{Float64}},
        prob::NonlinearProblem{HenonHeilesMidpoint}, method::Newton, p::BatchParams, net::LearnedStart;
        options::Union{Nothing, Options} = nothing, kwargs...)
    cm = device_method(method)
    (cm === nothing || cm.globalization != 0 || cm.reuse != 0) && throw(ArgumentError(
        "the learned start takes Newton with LUFactorization, the Static globalization and no reuse"))
    Xk = p.per_system::ROCVector{SVector{4, Float64}}
    length(Xk) == length(X) || throw(DimensionMismatch(
        "BatchParams length $(length(Xk)) does not match batch length $(length(X))"))
    length(S) == length(X) || throw(DimensionMismatch(
        "status vector length $(length(S)) does not match batch length $(length(X))"))
    opt = resolve_options(Float64, options, kwargs)
    isempty(X) && return 0
    fallbacks = Ref{Int64}(0)
    rc = ccall((:gs_solve_learned_henon_heiles_f64, libgsolve), Cint,
        (Ref{GsMlp}, Ref{GsMethod}, Int64, Ptr{Cvoid}, Float64, Ptr{Cvoid}, Ptr{Cvoid}, Ref{GsOptionsF64},
            Ref{Int64}, Ptr{Cvoid}),
        gs_mlp(net), cm, length(X), pointer(Xk), Float64(p.shared), pointer(X), pointer(S), device_options(opt),
        fallbacks, stream().stream)
    rc == 0 || gs_error(rc)
    return Int(fallbacks[])
end

end # module
