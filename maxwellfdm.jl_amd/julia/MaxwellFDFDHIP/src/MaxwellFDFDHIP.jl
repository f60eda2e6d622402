# MaxwellFDFDHIP — MaxwellFDFD's frequency-domain hot path on an MI355X, called
# from Julia through libmfdfd's C ABI (include/mfdfd.h).
#
# `using MaxwellFDFDHIP` in place of `using MaxwellFDFD` keeps every setter and
# source function of MaxwellFDFD (re-exported unchanged) and replaces the
# exported assembly functions of src/model/model.jl:23-28 by device versions
# with the same signatures: create_paramops, create_curls, create_Mcs,
# create_A, create_b, create_linsys, h_from_e, e_from_h.  It adds the solve the
# reference leaves to `A \ b` (solve_linsys) and an rrule for the ε → field
# map (solve_field).  Only ModelFull (3D) models go to the device.
#
# Material assignment and subpixel smoothing (calc_matparams!, full.jl:16-70)
# stay MaxwellFDFD's own: create_paramops runs them on the CPU exactly as the
# reference does and assembles Pε, Pμ on the device from the arrays they fill.
#
# Ownership: every array passed to the library is Julia's and is preserved
# (GC.@preserve) for the duration of the call; the library keeps no Julia
# pointer and never calls back.  Results are fresh Julia arrays.
module MaxwellFDFDHIP

using LinearAlgebra
using SparseArrays
using ChainRulesCore
using MaxwellBase: EE, HH, FieldType
import MaxwellFDFD
using MaxwellFDFD: Model, ModelFull, set_ωpml!, set_boundft!, set_Npml!, set_kbloch!, create_e⁻ⁱᵏᴸ,
                   clear_objs!, clear_srcs!, add_srce!, add_srcm!, create_srcs

export set_ωpml!, set_boundft!, set_Npml!, set_kbloch!, create_e⁻ⁱᵏᴸ, clear_objs!, clear_srcs!, add_srce!, add_srcm!,
       create_srcs
export create_paramops, create_curls, create_linsys, create_A, create_b, h_from_e, e_from_h, create_Mcs
export MaxwellOperator, solve_linsys, solve_field

# ---------------------------------------------------------------- library

# The shared library: MAXWELLFDFD_HIP_LIB overrides the in-tree build.
const libmfdfd = get(ENV, "MAXWELLFDFD_HIP_LIB", joinpath(@__DIR__, "..", "..", "..", "lib", "libmfdfd.so"))

const MFD_OK = Cint(0)
const MFD_ERR_ARG = Cint(1)
const MFD_ERR_UNSUPPORTED = Cint(2)
const MFD_ERR_NOCONV = Cint(4)
const MFD_FP64 = Cint(0)
const MFD_MIXED = Cint(1)

struct MfdNoConvergence <: Exception
    msg::String
end

last_error() = unsafe_string(ccall((:mfd_last_error, libmfdfd), Cstring, ()))

# status codes -> Julia exceptions (ArgumentError where the reference throws one)
function check(rc::Cint)
    rc == MFD_OK && return nothing
    msg = last_error()
    rc == MFD_ERR_ARG && throw(ArgumentError(msg))
    rc == MFD_ERR_NOCONV && throw(MfdNoConvergence(msg))
    rc == MFD_ERR_UNSUPPORTED && error("unsupported: " * msg)
    error("libmfdfd error $rc: " * msg)
end

const CTX = Ref{Ptr{Cvoid}}(C_NULL)

"The process-wide library context on device 0 (one HIP stream)."
function context()
    if CTX[] == C_NULL
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:mfd_ctx_create, libmfdfd), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, r))
        CTX[] = r[]
    end
    return CTX[]
end

# ------------------------------------------------------------ C structures

struct MfdGrid                       # mfd_grid
    N::NTuple{3,Int64}
    lprim::NTuple{3,Ptr{Float64}}
    isbloch::NTuple{3,Int32}
    boundft::NTuple{3,Int32}
    npml::NTuple{2,NTuple{3,Int64}}
    omega_pml::ComplexF64
    kbloch::NTuple{3,Float64}
    pml_m::Float64
    pml_lnR::Float64
    order_cmpfirst::Int32
end

struct MfdCsc                        # mfd_csc: a SparseMatrixCSC{ComplexF64,Int64} in place
    m::Int64
    n::Int64
    nnz::Int64
    colptr::Ptr{Int64}
    rowval::Ptr{Int64}
    nzval::Ptr{ComplexF64}
end

mutable struct MfdSolveOpts          # mfd_solve_opts
    tol::Float64
    maxit::Int64
    precision::Int32
    method::Int32
    inner_tol::Float64
    check_every::Int32
    use_graph::Int32
end

mutable struct MfdSolveStats         # mfd_solve_stats
    iterations::Int64
    outer_cycles::Int64
    relres::Float64
    seconds::Float64
    converged::Int32
    iterations_fp32::Int64
end
MfdSolveStats() = MfdSolveStats(0, 0, 0.0, 0.0, 0, 0)

c_ft(ft::FieldType) = ft == EE ? Int32(0) : Int32(1)

csc_of(A::SparseMatrixCSC{ComplexF64,Int64}) =
    MfdCsc(size(A, 1), size(A, 2), nnz(A), pointer(A.colptr), pointer(A.rowval), pointer(A.nzval))

# a CSC the library can read in place (rows ascending within each column, as
# SparseMatrixCSC keeps them)
as_lib(A::AbstractMatrix) = convert(SparseMatrixCSC{ComplexF64,Int64}, A)

# The N + 1 primal planes of each axis (the vertices grid.l[1][w] plus the
# positive-end bound grid.bounds[2][w]), kept alive with the descriptor.
function grid_desc(mdl::ModelFull)
    g = mdl.grid
    lp = ntuple(w -> Float64[g.l[1][w]; g.bounds[2][w]], 3)
    N = ntuple(w -> Int64(g.N[w]), 3)
    desc = MfdGrid(N, ntuple(w -> pointer(lp[w]), 3), ntuple(w -> Int32(g.isbloch[w]), 3),
                   ntuple(w -> c_ft(mdl.boundft[w]), 3),
                   (ntuple(w -> Int64(mdl.Npml[1][w]), 3), ntuple(w -> Int64(mdl.Npml[2][w]), 3)),
                   ComplexF64(mdl.ωpml), ntuple(w -> Float64(mdl.kbloch[w]), 3),
                   4.0, -16.0, Int32(mdl.order_cmpfirst))
    return desc, lp
end

# two-phase CSC call: query nnz, then fill Julia-owned arrays wrapped without copying
function csc_two_phase(fnnz, ffill, n::Integer)
    nz = Ref{Int64}(0)
    check(fnnz(nz))
    colptr = Vector{Int64}(undef, n + 1)
    rowval = Vector{Int64}(undef, max(nz[], 1))
    nzval = Vector{ComplexF64}(undef, max(nz[], 1))
    check(ffill(colptr, rowval, nzval))
    resize!(rowval, nz[])
    resize!(nzval, nz[])
    return SparseMatrixCSC(n, n, colptr, rowval, nzval)
end

# ------------------------------------------------- model-based assembly (model.jl)

"create_curls (model.jl:160-175) on the device: (Cₑ, Cₘ)."
function create_curls(mdl::ModelFull)
    desc, lp = grid_desc(mdl)
    n = 3 * prod(desc.N)
    GC.@preserve lp begin
        C = ntuple(2) do which
            csc_two_phase(nz -> ccall((:mfd_curl_nnz, libmfdfd), Cint, (Ptr{Cvoid}, Ref{MfdGrid}, Cint, Ref{Int64}),
                                      context(), desc, which - 1, nz),
                          (cp, rv, v) -> ccall((:mfd_curl_csc, libmfdfd), Cint,
                                               (Ptr{Cvoid}, Ref{MfdGrid}, Cint, Ptr{Int64}, Ptr{Int64}, Ptr{ComplexF64}),
                                               context(), desc, which - 1, cp, rv, v), n)
        end
    end
    return C
end

"create_Mcs (model.jl:287-306) on the device: (Mcₑ, Mcₘ)."
function create_Mcs(mdl::ModelFull)
    desc, lp = grid_desc(mdl)
    n = 3 * prod(desc.N)
    GC.@preserve lp begin
        M = ntuple(2) do which
            csc_two_phase(nz -> ccall((:mfd_Mcs_nnz, libmfdfd), Cint, (Ptr{Cvoid}, Ref{MfdGrid}, Cint, Ref{Int64}),
                                      context(), desc, which - 1, nz),
                          (cp, rv, v) -> ccall((:mfd_Mcs_csc, libmfdfd), Cint,
                                               (Ptr{Cvoid}, Ref{MfdGrid}, Cint, Ptr{Int64}, Ptr{Int64}, Ptr{ComplexF64}),
                                               context(), desc, which - 1, cp, rv, v), n)
        end
    end
    return M
end

"""
create_paramops (model.jl:141-158): fills mdl.εarr/μarr with MaxwellFDFD's own
calc_matparams! (CPU, as the reference), then assembles (Pε, Pμ) on the device.
"""
function create_paramops(mdl::ModelFull)
    MaxwellFDFD.calc_matparams!(mdl)
    desc, lp = grid_desc(mdl)
    n = 3 * prod(desc.N)
    ε = convert(Array{ComplexF64,5}, mdl.εarr)
    μ = convert(Array{ComplexF64,5}, mdl.μarr)
    GC.@preserve lp ε μ begin
        P = ntuple(2) do which
            parr = which == 1 ? ε : μ
            csc_two_phase(nz -> ccall((:mfd_paramop_nnz, libmfdfd), Cint,
                                      (Ptr{Cvoid}, Ref{MfdGrid}, Cint, Ptr{ComplexF64}, Ref{Int64}),
                                      context(), desc, which - 1, parr, nz),
                          (cp, rv, v) -> ccall((:mfd_paramop_csc, libmfdfd), Cint,
                                               (Ptr{Cvoid}, Ref{MfdGrid}, Cint, Ptr{ComplexF64}, Ptr{Int64}, Ptr{Int64},
                                                Ptr{ComplexF64}), context(), desc, which - 1, parr, cp, rv, v), n)
        end
    end
    return P
end

# ------------------------------------- matrix-argument API (model.jl:209-284)

create_A(ft::FieldType, ω::Number, Ps::Tuple{AbstractMatrix,AbstractMatrix}, Cs::Tuple{AbstractMatrix,AbstractMatrix}) =
    create_A(ft, ω, Ps..., Cs...)

"""
create_A(ft, ω, Pε, Pμ, Cₑ, Cₘ) (model.jl:225-246), formed on the device:
A = Cₘ (Pμ \\ Cₑ) − ω² Pε (EE) or Cₑ (Pε \\ Cₘ) − ω² Pμ (HH) with the same
structure rules as SparseArrays (diagonal divisor, structural entries of the
product kept, numerical zeros of the subtraction dropped).
"""
function create_A(ft::FieldType, ω::Number, Pε::AbstractMatrix, Pμ::AbstractMatrix, Cₑ::AbstractMatrix,
                  Cₘ::AbstractMatrix)
    M = map(as_lib, (Pε, Pμ, Cₑ, Cₘ))
    n = size(M[1], 2)
    GC.@preserve M begin
        d = map(csc_of, M)
        return csc_two_phase(nz -> ccall((:mfd_A_from_csc_nnz, libmfdfd), Cint,
                                         (Ptr{Cvoid}, Cint, ComplexF64, Ref{MfdCsc}, Ref{MfdCsc}, Ref{MfdCsc},
                                          Ref{MfdCsc}, Ref{Int64}),
                                         context(), c_ft(ft), ComplexF64(ω), d[1], d[2], d[3], d[4], nz),
                             (cp, rv, v) -> ccall((:mfd_A_from_csc, libmfdfd), Cint,
                                                  (Ptr{Cvoid}, Cint, ComplexF64, Ref{MfdCsc}, Ref{MfdCsc}, Ref{MfdCsc},
                                                   Ref{MfdCsc}, Ptr{Int64}, Ptr{Int64}, Ptr{ComplexF64}),
                                                  context(), c_ft(ft), ComplexF64(ω), d[1], d[2], d[3], d[4], cp, rv, v), n)
    end
end

create_b(ft::FieldType, ω::Number, Ps::Tuple{AbstractMatrix,AbstractMatrix}, Cs::Tuple{AbstractMatrix,AbstractMatrix},
         js::Tuple{AbstractVector,AbstractVector}) = create_b(ft, ω, Ps..., Cs..., js...)

"create_b(ft, ω, Pε, Pμ, Cₑ, Cₘ, jₑ, jₘ) (model.jl:251-274) on the device."
function create_b(ft::FieldType, ω::Number, Pε::AbstractMatrix, Pμ::AbstractMatrix, Cₑ::AbstractMatrix,
                  Cₘ::AbstractMatrix, jₑ::AbstractVector, jₘ::AbstractVector)
    M = map(as_lib, (Pε, Pμ, Cₑ, Cₘ))
    je = convert(Vector{ComplexF64}, jₑ)
    jm = convert(Vector{ComplexF64}, jₘ)
    length(je) == length(jm) == size(M[1], 2) || throw(ArgumentError("jₑ and jₘ must have n entries"))
    b = Vector{ComplexF64}(undef, length(je))
    GC.@preserve M je jm b begin
        d = map(csc_of, M)
        check(ccall((:mfd_b_from_csc, libmfdfd), Cint,
                    (Ptr{Cvoid}, Cint, ComplexF64, Ref{MfdCsc}, Ref{MfdCsc}, Ref{MfdCsc}, Ref{MfdCsc},
                     Ptr{ComplexF64}, Ptr{ComplexF64}, Ptr{ComplexF64}),
                    context(), c_ft(ft), ComplexF64(ω), d[1], d[2], d[3], d[4], je, jm, b))
    end
    return b
end

create_linsys(ft::FieldType, ω::Number, Ps::Tuple{AbstractMatrix,AbstractMatrix},
              Cs::Tuple{AbstractMatrix,AbstractMatrix}, js::Tuple{AbstractVector,AbstractVector}) =
    create_linsys(ft, ω, Ps..., Cs..., js...)

"create_linsys (model.jl:209-220): (A, b), both formed on the device."
create_linsys(ft::FieldType, ω::Number, Pε::AbstractMatrix, Pμ::AbstractMatrix, Cₑ::AbstractMatrix,
              Cₘ::AbstractMatrix, jₑ::AbstractVector, jₘ::AbstractVector) =
    (create_A(ft, ω, Pε, Pμ, Cₑ, Cₘ), create_b(ft, ω, Pε, Pμ, Cₑ, Cₘ, jₑ, jₘ))

function field_csc(sym::Symbol, v::AbstractVector, ω::Number, P::AbstractMatrix, Cx::AbstractMatrix,
                   j::AbstractVector)
    Pl, Cl = as_lib(P), as_lib(Cx)
    x = convert(Vector{ComplexF64}, v)
    s = convert(Vector{ComplexF64}, j)
    length(x) == length(s) == size(Cl, 2) || throw(ArgumentError("field and source vectors must have n entries"))
    out = Vector{ComplexF64}(undef, length(x))
    GC.@preserve Pl Cl x s out begin
        dp, dc = csc_of(Pl), csc_of(Cl)
        rc = if sym === :h
            ccall((:mfd_h_from_e_csc, libmfdfd), Cint,
                  (Ptr{Cvoid}, Ptr{ComplexF64}, ComplexF64, Ref{MfdCsc}, Ref{MfdCsc}, Ptr{ComplexF64}, Ptr{ComplexF64}),
                  context(), x, ComplexF64(ω), dp, dc, s, out)
        else
            ccall((:mfd_e_from_h_csc, libmfdfd), Cint,
                  (Ptr{Cvoid}, Ptr{ComplexF64}, ComplexF64, Ref{MfdCsc}, Ref{MfdCsc}, Ptr{ComplexF64}, Ptr{ComplexF64}),
                  context(), x, ComplexF64(ω), dp, dc, s, out)
        end
        check(rc)
    end
    return out
end

"h_from_e (model.jl:276-279): (i/ω) (Pμ \\ (Cₑ e + jₘ)) on the device."
h_from_e(e::AbstractVector, ω::Number, Pμ::AbstractMatrix, Cₑ::AbstractMatrix, jₘ::AbstractVector) =
    field_csc(:h, e, ω, Pμ, Cₑ, jₘ)
h_from_e(e::AbstractVector, ω::Number, M::Tuple{AbstractMatrix,AbstractMatrix}, C::Tuple{AbstractMatrix,AbstractMatrix},
         j::Tuple{AbstractVector,AbstractVector}) = h_from_e(e, ω, M[2], C[1], j[2])

"e_from_h (model.jl:281-284): (−i/ω) (Pε \\ (Cₘ h − jₑ)) on the device."
e_from_h(h::AbstractVector, ω::Number, Pε::AbstractMatrix, Cₘ::AbstractMatrix, jₑ::AbstractVector) =
    field_csc(:e, h, ω, Pε, Cₘ, jₑ)
e_from_h(h::AbstractVector, ω::Number, M::Tuple{AbstractMatrix,AbstractMatrix}, C::Tuple{AbstractMatrix,AbstractMatrix},
         j::Tuple{AbstractVector,AbstractVector}) = e_from_h(h, ω, M[1], C[2], j[1])

# ------------------------------------------------------------------ solve

"""
Device operator: MaxwellOperator(A) from an assembled sparse A (create_A /
create_linsys), or MaxwellOperator(mdl, ω; ft) matrix-free from a model whose
εarr/μarr are filled (create_paramops or MaxwellFDFD.calc_matparams!).
"""
mutable struct MaxwellOperator
    h::Ptr{Cvoid}
    n::Int
    ω::ComplexF64
    function MaxwellOperator(h::Ptr{Cvoid}, n::Integer, ω::Number)
        op = new(h, n, ComplexF64(ω))
        finalizer(op) do o
            o.h == C_NULL || ccall((:mfd_op_destroy, libmfdfd), Cint, (Ptr{Cvoid},), o.h)
            o.h = C_NULL
        end
        return op
    end
end

function MaxwellOperator(A::AbstractMatrix)
    Al = as_lib(A)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve Al begin
        check(ccall((:mfd_op_create_csc, libmfdfd), Cint, (Ptr{Cvoid}, Ref{MfdCsc}, Ref{Ptr{Cvoid}}),
                    context(), csc_of(Al), r))
    end
    return MaxwellOperator(r[], size(Al, 1), 0)
end

function MaxwellOperator(mdl::ModelFull, ω::Number; ft::FieldType=EE,
                         εarr::AbstractArray=mdl.εarr, μarr::AbstractArray=mdl.μarr)
    desc, lp = grid_desc(mdl)
    ε = convert(Array{ComplexF64,5}, εarr)
    μ = convert(Array{ComplexF64,5}, μarr)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve lp ε μ begin
        check(ccall((:mfd_op_create, libmfdfd), Cint,
                    (Ptr{Cvoid}, Ref{MfdGrid}, Cint, ComplexF64, Ptr{ComplexF64}, Ptr{ComplexF64}, Ref{Ptr{Cvoid}}),
                    context(), desc, c_ft(ft), ComplexF64(ω), ε, μ, r))
    end
    return MaxwellOperator(r[], 3 * prod(desc.N), ω)
end

function solve_opts(tol, precision::Symbol, maxit)
    o = MfdSolveOpts(0.0, 0, 0, 0, 0.0, 0, 0)
    ccall((:mfd_solve_opts_default, libmfdfd), Cvoid, (Ref{MfdSolveOpts},), o)
    o.tol = tol
    o.precision = precision === :mixed ? MFD_MIXED : MFD_FP64
    maxit === nothing || (o.maxit = maxit)
    return o
end

"""
solve_linsys(A, b; tol=1e-8, precision=:fp64) — the user's `A \\ b` on the
device: iterative, to ‖b − A e‖/‖b‖ ≤ tol in FP64 (precision = :mixed runs
complex-FP32 inner cycles with FP64 refinement).  A is a MaxwellOperator or
an assembled sparse A; b one vector or n×k columns.  Returns (e, stats).
"""
function solve_linsys(op::MaxwellOperator, b::AbstractVecOrMat; tol::Real=1e-8, precision::Symbol=:fp64,
                      maxit=nothing, x0=nothing)
    bb = convert(Array{ComplexF64}, b)
    size(bb, 1) == op.n || throw(ArgumentError("b must have $(op.n) rows"))
    nrhs = size(bb, 2)
    e = x0 === nothing ? zeros(ComplexF64, size(bb)) : convert(Array{ComplexF64}, copy(x0))
    size(e) == size(bb) || throw(ArgumentError("x0 must have the size of b"))
    o = solve_opts(tol, precision, maxit)
    st = MfdSolveStats()
    GC.@preserve bb e begin
        check(ccall((:mfd_solve, libmfdfd), Cint,
                    (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{ComplexF64}, Int64, Ref{MfdSolveOpts}, Ref{MfdSolveStats}),
                    op.h, bb, e, nrhs, o, st))
    end
    return e, st
end
solve_linsys(A::AbstractMatrix, b::AbstractVecOrMat; kw...) = solve_linsys(MaxwellOperator(A), b; kw...)

function adjoint_solve(op::MaxwellOperator, g::AbstractVector; tol::Real=1e-8)
    gg = convert(Vector{ComplexF64}, g)
    λ = zeros(ComplexF64, length(gg))
    o = solve_opts(tol, :fp64, nothing)
    st = MfdSolveStats()
    GC.@preserve gg λ begin
        check(ccall((:mfd_solve_adjoint, libmfdfd), Cint,
                    (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{ComplexF64}, Ref{MfdSolveOpts}, Ref{MfdSolveStats}),
                    op.h, gg, λ, o, st))
    end
    return λ
end

function adjoint_grad(op::MaxwellOperator, e::AbstractVector, λ::AbstractVector)
    ee = convert(Vector{ComplexF64}, e)
    ll = convert(Vector{ComplexF64}, λ)
    grad = zeros(Float64, length(ee))
    GC.@preserve ee ll grad begin
        check(ccall((:mfd_adjoint_grad, libmfdfd), Cint,
                    (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{ComplexF64}, Ptr{Float64}), op.h, ee, ll, grad))
    end
    return grad
end

# εarr of mdl with its diagonal replaced, per E DOF, by the real vector ε
# (component-first or component-last as mdl.order_cmpfirst says)
function εarr_with_diag(mdl::ModelFull, ε::AbstractVector{<:Real})
    εarr = convert(Array{ComplexF64,5}, copy(mdl.εarr))
    Nx, Ny, Nz = size(εarr)[1:3]
    M = Nx * Ny * Nz
    length(ε) == 3M || throw(ArgumentError("ε must have one entry per E DOF (3·$M)"))
    for c in 1:3, k in 1:Nz, j in 1:Ny, i in 1:Nx
        p = i + Nx * ((j - 1) + Ny * (k - 1))        # 1-based cell index
        dof = mdl.order_cmpfirst ? c + 3 * (p - 1) : p + M * (c - 1)
        εarr[i, j, k, c, c] = ε[dof]
    end
    return εarr
end

"""
solve_field(ε, mdl, ω, b; tol=1e-10) — the E field of A(ε) e = b, where ε is
the real diagonal permittivity per E DOF (the rest of the model's εarr and μarr
is kept).  Differentiable: its rrule solves the adjoint Aᵀλ = (∂f/∂e)ᵀ on the
device (Aᵀ = S A S⁻¹) and returns ∂f/∂ε = 2 Re(ω² λ ∘ e).
"""
function solve_field(ε::AbstractVector{<:Real}, mdl::ModelFull, ω::Number, b::AbstractVector; tol::Real=1e-10)
    op = MaxwellOperator(mdl, ω; εarr=εarr_with_diag(mdl, ε))
    e, _ = solve_linsys(op, b; tol=tol)
    return vec(e)
end

function ChainRulesCore.rrule(::typeof(solve_field), ε::AbstractVector{<:Real}, mdl::ModelFull, ω::Number,
                              b::AbstractVector; tol::Real=1e-10)
    op = MaxwellOperator(mdl, ω; εarr=εarr_with_diag(mdl, ε))
    e, _ = solve_linsys(op, b; tol=tol)
    e = vec(e)
    function solve_field_pullback(ē)
        ēv = unthunk(ē)
        # ChainRules' cotangent of a complex output of a real loss f is
        # ∂f/∂Re(e) + i ∂f/∂Im(e); the holomorphic gradient (∂f/∂e)ᵀ the
        # adjoint solve takes is its conjugate over 2
        g = conj.(convert(Vector{ComplexF64}, ēv)) ./ 2
        λ = adjoint_solve(op, g; tol=tol)
        return NoTangent(), adjoint_grad(op, e, λ), NoTangent(), NoTangent(), NoTangent()
    end
    return e, solve_field_pullback
end

end # module MaxwellFDFDHIP
