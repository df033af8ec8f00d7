# EqcycleHIP.jl -- Julia bindings of libeqcycle_hip.so for the eqcycle time loop.
#
# Drop-in replacements, same names / positional arguments / error behaviour as
# the reference routines (paths relative to the eqcycle repository):
#   FBC!        src/NRsearch_serial.jl:6-45
#   NRsearch!   src/NRsearch_serial.jl:47-98
#   IDS!, IDS2! src/otherFunctions_serial.jl:5-59
#   slrFunc!    src/otherFunctions_serial.jl:62-83
#   dtevol!     src/dtevol.jl:5-44
#   KOp / KniOp mul!            Ksparse'*u, nKsparse'*u, kni'*x  (src/main.jl:246,267,311,256)
#   AmgPreconditioner ldiv!     aspreconditioner(ruge_stuben(kni)) (src/main.jl:200-201)
#   cg!(x, ::KniOp, b; Pl, tol) IterativeSolvers.cg! (src/main.jl:256), whole solve on the GPU
#   main(P)                     src/main.jl:67-476, device-resident
#
# Usage in run.jl, after `P = setParameters(24e3, 20)`:
#   include("julia/EqcycleHIP.jl"); using .EqcycleHIP
#   EqcycleHIP.Context(P; W, iglob, ...) # mesh, material W, params_*; becomes the current context
#   O = EqcycleHIP.main(P)               # the reference's main(P), unchanged call site
# FBC!, NRsearch!, IDS!, IDS2!, slrFunc!, dtevol! and main(P) take exactly the
# reference's arguments and run on the current context (use_context!(ctx) to
# switch); FBC! and slrFunc! read the fault parameters from their P argument.
# Every routine also has a method with a leading ctx::Context.
module EqcycleHIP

using LinearAlgebra
using SparseArrays

const LIB = get(ENV, "EQCYCLE_HIP_LIB", joinpath(@__DIR__, "..", "eqcycle_amd", "libeqcycle_hip.so"))

const EQ_OK = 0
const EQ_NR_NOCONV = 3
const EQ_NR_NAN = 4
const EQ_PCG_NOCONV = 5
const EQ_MASK_ALL, EQ_MASK_INTERIOR, EQ_MASK_NEG, EQ_MASK_FAULT_ROWS = 0, 1, 2, 3
const EQ_PREC_FP64, EQ_PREC_FP32, EQ_PREC_MIXED, EQ_PREC_F32PC = 0, 1, 2, 3
const EQ_PC_NONE, EQ_PC_JACOBI, EQ_PC_AMG = 0, 1, 2
const EQ_SMOOTH_JACOBI, EQ_SMOOTH_CHEBYSHEV, EQ_SMOOTH_CHEBYSHEV_TILES = 0, 1, 2

struct MeshDesc
    ngll::Int64
    nelx::Int64
    nely::Int64
    dxe::Float64
    dye::Float64
    iglob::Ptr{Int64}
    xgll::Ptr{Float64}
    wgll::Ptr{Float64}
    H::Ptr{Float64}
end

struct ParamsDesc
    FltNglob::Int64; nglob::Int64; IDstate::Int64
    Vpl::Float64; Vthres::Float64; Vevne::Float64; dt0::Float64
    fo::Ptr{Float64}; Vo::Ptr{Float64}; xLf::Ptr{Float64}; M::Ptr{Float64}
    BcLC::Ptr{Float64}; BcTC::Ptr{Float64}; FltB::Ptr{Float64}; FltZ::Ptr{Float64}; FltX::Ptr{Float64}
    cca::Ptr{Float64}; ccb::Ptr{Float64}; Seff::Ptr{Float64}; tauo::Ptr{Float64}; XiLf::Ptr{Float64}
    iFlt::Ptr{Int64}; iBcL::Ptr{Int64}; n_iBcL::Int64; iBcT::Ptr{Int64}; n_iBcT::Int64
    FltIglobBC::Ptr{Int64}; n_FltIglobBC::Int64; out_seis::Ptr{Int64}; n_out_seis::Int64
end

struct FaultParams                     # eq_fault_params: the params_farray fields FBC! / slrFunc! read
    fo::Ptr{Float64}; Vo::Ptr{Float64}; xLf::Ptr{Float64}; cca::Ptr{Float64}
    ccb::Ptr{Float64}; Seff::Ptr{Float64}; tauo::Ptr{Float64}; FltZ::Ptr{Float64}
end
FaultParams(P) = FaultParams(pointer(P.fo), pointer(P.Vo), pointer(P.xLf), pointer(P.cca), pointer(P.ccb),
                             pointer(P.Seff), pointer(P.tauo), pointer(P.FltZ))

mutable struct Context
    h::Ptr{Cvoid}
    nglob::Int
    nI::Int
    keep::Vector{Any}          # arrays whose pointers the library may hold during set-up
end

# the context the reference-signature methods run on (the last one created)
const CURRENT = Ref{Union{Context,Nothing}}(nothing)
use_context!(ctx::Context) = (CURRENT[] = ctx)
function current()
    CURRENT[] === nothing && error("EqcycleHIP: create a Context first")
    return CURRENT[]
end

lasterr(ctx::Context) = unsafe_string(ccall((:eq_last_error, LIB), Cstring, (Ptr{Cvoid},), ctx.h))
function check(ctx::Context, rc)
    rc == EQ_OK && return rc
    error("libeqcycle_hip: ", lasterr(ctx))
end

"""Context(P; W, iglob, dxe, dye, xgll, wgll, H, amg=true): one mesh on one GPU.

`P` is the 5-tuple of setParameters; the mesh arrays and W are the locals of
setParameters (par.jl:99-152) that the reference does not return -- pass them
explicitly."""
function Context(P; W, iglob, NelX, NelY, dxe, dye, xgll, wgll, H, amg = true)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    ig = convert(Array{Int64}, iglob)
    Hc = convert(Matrix{Float64}, H)
    GC.@preserve ig xgll wgll Hc begin
        d = MeshDesc(5, NelX, NelY, dxe, dye, pointer(ig), pointer(xgll), pointer(wgll), pointer(Hc))
        rc = ccall((:eq_create, LIB), Cint, (Ref{MeshDesc}, Ref{Ptr{Cvoid}}), d, h)
        rc == EQ_OK || error("eq_create failed ($rc)")
    end
    ctx = Context(h[], P[1].nglob, length(P[4].FltNI), Any[])
    finalizer(c -> ccall((:eq_destroy, LIB), Cvoid, (Ptr{Cvoid},), c.h), ctx)
    GC.@preserve W check(ctx, ccall((:eq_set_material, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), ctx.h, W))
    set_params!(ctx, P)
    amg && check(ctx, ccall((:eq_amg_setup, LIB), Cint,
                           (Ptr{Cvoid}, Cdouble, Cint, Int64, Cint, Cint, Cint, Cint),
                           ctx.h, 0.25, 10, 2000, 1, EQ_SMOOTH_CHEBYSHEV_TILES, 5, 1))
    return use_context!(ctx)
end

function set_params!(ctx::Context, P)
    f, ia = P[3], P[4]
    a = (f.fo, f.Vo, f.xLf, f.M, f.BcLC, f.BcTC, f.FltB, f.FltZ, f.FltX, f.cca, f.ccb, f.Seff, f.tauo, f.XiLf,
         ia.iFlt, ia.iBcL, ia.iBcT, ia.FltIglobBC, ia.out_seis)
    GC.@preserve a begin
        d = ParamsDesc(P[1].FltNglob, P[1].nglob, P[1].IDstate, P[2].Vpl, P[2].Vthres, P[2].Vevne, P[2].dt0,
                       pointer(f.fo), pointer(f.Vo), pointer(f.xLf), pointer(f.M), pointer(f.BcLC),
                       pointer(f.BcTC), pointer(f.FltB), pointer(f.FltZ), pointer(f.FltX), pointer(f.cca),
                       pointer(f.ccb), pointer(f.Seff), pointer(f.tauo), pointer(f.XiLf),
                       pointer(ia.iFlt), pointer(ia.iBcL), length(ia.iBcL), pointer(ia.iBcT), length(ia.iBcT),
                       pointer(ia.FltIglobBC), length(ia.FltIglobBC), pointer(ia.out_seis), length(ia.out_seis))
        check(ctx, ccall((:eq_set_params, LIB), Cint, (Ptr{Cvoid}, Ref{ParamsDesc}), ctx.h, d))
    end
end

# ---- operators: the AbstractMatrix protocol cg! needs (test_01.jl:9-27) ------
struct EqCsr
    n::Int64
    m::Int64
    rowptr::Ptr{Int64}
    col::Ptr{Int64}
    val::Ptr{Float64}
end

# Keep the reference's own hierarchy: ml = ruge_stuben(kni) (src/main.jl:200)
# goes to the device level by level.  A symmetric level operator's CSC arrays
# are its CSR arrays; the CSC arrays of R = P' are the CSR arrays of P.
function upload_amg!(ctx::Context, ml; smoother = EQ_SMOOTH_JACOBI, degree = 1, omega = nothing)
    As = SparseMatrixCSC[lvl.A for lvl in ml.levels]
    push!(As, ml.final_A)
    keep = Any[]
    function desc(M::SparseMatrixCSC, n, m)
        rp = Int64.(M.colptr) .- 1
        cl = Int64.(M.rowval) .- 1
        v = Vector{Float64}(M.nzval)
        push!(keep, rp, cl, v)
        EqCsr(n, m, pointer(rp), pointer(cl), pointer(v))
    end
    Ad = [desc(A, size(A, 1), size(A, 2)) for A in As]
    Pd = [desc(lvl.R, size(lvl.R, 2), size(lvl.R, 1)) for lvl in ml.levels]
    om = omega === nothing ? Ptr{Float64}(C_NULL) : pointer(omega)
    GC.@preserve keep omega check(ctx, ccall((:eq_amg_upload, LIB), Cint,
                                             (Ptr{Cvoid}, Cint, Ptr{EqCsr}, Ptr{EqCsr}, Ptr{Float64}, Cint, Cint),
                                             ctx.h, length(Ad), Ad, Pd, om, smoother, degree))
    return AmgPreconditioner(ctx)
end

struct KOp <: AbstractMatrix{Float64}   # Ksparse' (mask EQ_MASK_ALL) or nKsparse' (EQ_MASK_NEG)
    ctx::Context
    mask::Int
end
Base.size(A::KOp) = (A.ctx.nglob, A.ctx.nglob)
Base.eltype(::KOp) = Float64
Base.getindex(::KOp, i::Int, j::Int) = error("matrix-free operator")
function LinearAlgebra.mul!(y::Vector{Float64}, A::KOp, x::Vector{Float64})
    check(A.ctx, ccall((:eq_ku_apply, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Cint),
                       A.ctx.h, x, y, A.mask, EQ_PREC_FP64))
    y
end
Base.:*(A::KOp, x::Vector{Float64}) = mul!(similar(x), A, x)

struct KniOp <: AbstractMatrix{Float64}  # kni' = -K[FltNI, FltNI]
    ctx::Context
end
Base.size(A::KniOp) = (A.ctx.nI, A.ctx.nI)
Base.eltype(::KniOp) = Float64
Base.getindex(::KniOp, i::Int, j::Int) = error("matrix-free operator")
function LinearAlgebra.mul!(y::Vector{Float64}, A::KniOp, x::Vector{Float64})
    check(A.ctx, ccall((:eq_kni_apply, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint),
                       A.ctx.h, x, y, EQ_PREC_FP64))
    y
end

struct AmgPreconditioner
    ctx::Context
end
function LinearAlgebra.ldiv!(x::Vector{Float64}, p::AmgPreconditioner, b::Vector{Float64})
    check(p.ctx, ccall((:eq_amg_ldiv, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint),
                       p.ctx.h, b, x, EQ_PREC_FP64))
    x
end

"""cg!(x, kni, b; Pl, tol, maxiter): the whole solve on the GPU, IterativeSolvers semantics
(||b - kni x|| <= tol ||b||).  A non-converged solve returns x, silently, as cg! does."""
function cg!(x::Vector{Float64}, A::KniOp, b::Vector{Float64}; Pl = nothing, tol = sqrt(eps()),
             maxiter::Int = size(A, 2), prec::Int = EQ_PREC_FP64)
    it = Ref{Int64}(0); rr = Ref{Float64}(0)
    rc = ccall((:eq_cg, LIB), Cint,
               (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cdouble, Int64, Cint, Cint, Ref{Int64}, Ref{Float64}),
               A.ctx.h, x, b, tol, maxiter, Pl === nothing ? EQ_PC_NONE : EQ_PC_AMG, prec, it, rr)
    rc == EQ_PCG_NOCONV || check(A.ctx, rc)
    x
end

# ---- fault routines (same positional signatures as the reference) -------------
function FBC!(ctx::Context, IDstate, P, NFBC, FltNglob, psi1, Vf1, tau1, psi2, Vf2, tau2, psi, Vf, FltVfree, dt)
    fail = Ref{Int64}(0)
    rc = GC.@preserve P ccall((:eq_fbc, LIB), Cint,
               (Ptr{Cvoid}, Int64, Ref{FaultParams}, Int64, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cdouble,
                Ref{Int64}),
               ctx.h, IDstate, FaultParams(P), NFBC, FltNglob, psi1, Vf1, tau1, psi2, Vf2, tau2, psi, Vf,
               FltVfree, dt, fail)
    if rc == EQ_NR_NAN
        println("Fault Location = ", fail[])
        @error("NR SEARCH FAILED!")
        return
    end
    rc == EQ_NR_NOCONV || check(ctx, rc)
    return psi1, Vf1, tau1, psi2, Vf2, tau2
end

function NRsearch!(ctx::Context, fo, Vo, cca, ccb, Seff, tau, tauo, psi, FltZ, FltVfree)
    Vf = Ref{Float64}(0); t = Ref{Float64}(0); k = Ref{Int32}(0)
    rc = ccall((:eq_nrsearch, LIB), Cint,
               (Ptr{Cvoid}, Int64, Ref{Float64}, Ref{Float64}, Ref{Float64}, Ref{Float64}, Ref{Float64},
                Ref{Float64}, Ref{Float64}, Ref{Float64}, Ref{Float64}, Ref{Float64}, Ref{Float64},
                Ref{Float64}, Ref{Int32}),
               ctx.h, 1, fo, Vo, cca, ccb, Seff, Float64(tau), tauo, psi, FltZ, FltVfree, Vf, t, k)
    rc == EQ_NR_NOCONV && println("k = ", -k[])
    rc in (EQ_OK, EQ_NR_NOCONV) || check(ctx, rc)
    return Vf[], t[]
end

function IDS!(ctx::Context, xLf, Vo, psi, dt, Vf, cnd, IDstate)
    out = Ref{Float64}(0)
    check(ctx, ccall((:eq_ids, LIB), Cint, (Ptr{Cvoid}, Int64, Ref{Float64}, Ref{Float64}, Ref{Float64}, Cdouble,
                                            Ref{Float64}, Cdouble, Int64, Ref{Float64}),
                     ctx.h, 1, xLf, Vo, psi, dt, Vf, cnd, IDstate, out))
    out[]
end

function IDS2!(ctx::Context, xLf, Vo, psi, psi1, dt, Vf, Vf1, IDstate)
    out = Ref{Float64}(0)
    check(ctx, ccall((:eq_ids2, LIB), Cint, (Ptr{Cvoid}, Int64, Ref{Float64}, Ref{Float64}, Ref{Float64},
                                             Ref{Float64}, Cdouble, Ref{Float64}, Ref{Float64}, Int64, Ref{Float64}),
                     ctx.h, 1, xLf, Vo, psi, psi1, dt, Vf, Vf1, IDstate, out))
    out[]
end

function slrFunc!(ctx::Context, P, NFBC, FltNglob, psi, psi1, Vf, Vf1, IDstate, tau1, dt)
    GC.@preserve P check(ctx, ccall((:eq_slr, LIB), Cint,
                                    (Ptr{Cvoid}, Ref{FaultParams}, Int64, Int64, Ptr{Float64}, Ptr{Float64},
                                     Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Float64}, Cdouble),
                                    ctx.h, FaultParams(P), NFBC, FltNglob, psi, psi1, Vf, Vf1, IDstate, tau1, dt))
    return psi1, Vf1
end

function dtevol!(ctx::Context, dt, dtmin, XiLf, FaultNglob, NFBC, Vf, isolver)
    out = Ref{Float64}(0)
    check(ctx, ccall((:eq_dtevol, LIB), Cint, (Ptr{Cvoid}, Cdouble, Cdouble, Ptr{Float64}, Int64, Int64,
                                               Ptr{Float64}, Int64, Ref{Float64}),
                     ctx.h, dt, dtmin, XiLf, FaultNglob, NFBC, Vf, isolver, out))
    out[]
end

# The reference's own signatures, on the current context: call sites of
# src/main.jl and src/NRsearch_serial.jl need no edit.
FBC!(IDstate, P, NFBC, FltNglob, psi1, Vf1, tau1, psi2, Vf2, tau2, psi, Vf, FltVfree, dt) =
    FBC!(current(), IDstate, P, NFBC, FltNglob, psi1, Vf1, tau1, psi2, Vf2, tau2, psi, Vf, FltVfree, dt)
NRsearch!(fo, Vo, cca, ccb, Seff, tau, tauo, psi, FltZ, FltVfree) =
    NRsearch!(current(), fo, Vo, cca, ccb, Seff, tau, tauo, psi, FltZ, FltVfree)
IDS!(xLf, Vo, psi, dt, Vf, cnd, IDstate = 2) = IDS!(current(), xLf, Vo, psi, dt, Vf, cnd, IDstate)
IDS2!(xLf, Vo, psi, psi1, dt, Vf, Vf1, IDstate = 2) = IDS2!(current(), xLf, Vo, psi, psi1, dt, Vf, Vf1, IDstate)
slrFunc!(P, NFBC, FltNglob, psi, psi1, Vf, Vf1, IDstate, tau1, dt) =
    slrFunc!(current(), P, NFBC, FltNglob, psi, psi1, Vf, Vf1, IDstate, tau1, dt)
dtevol!(dt, dtmin, XiLf, FaultNglob, NFBC, Vf, isolver) =
    dtevol!(current(), dt, dtmin, XiLf, FaultNglob, NFBC, Vf, isolver)

# The dynamic branch of main.jl's loop body (src/main.jl:297-344) in one call:
# d, v, a (nglob) and psi (FltNglob) are advanced in place; tau and Vf receive
# the new fault traction and the pre-step slip rate.  Returns nothing after
# the reference's NR SEARCH FAILED! report, as FBC! does.
function newmark_step!(ctx::Context, d, v, a, psi, tau, Vf, FltVfree, dt)
    rc = ccall((:eq_newmark_step, LIB), Cint,
               (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                Ptr{Float64}, Cdouble),
               ctx.h, d, v, a, psi, tau, Vf, FltVfree, dt)
    if rc == EQ_NR_NAN
        println(lasterr(ctx))
        @error("NR SEARCH FAILED!")
        return
    end
    rc == EQ_NR_NOCONV || check(ctx, rc)
    return d, v, a, psi, tau
end

# ---- device-resident loop: main(P) (src/main.jl:67-476) ---------------------
const EQ_GUESS_PREVIOUS, EQ_GUESS_LEARNED, EQ_GUESS_GALERKIN, EQ_GUESS_LINEAR = 0, 1, 2, 3

struct SolverOpts                      # eq_solver_opts, field for field
    tol::Float64
    maxiter::Int64
    pc::Cint
    prec::Cint
    guess::Cint
    record_snapshots::Cint
    graphs::Cint
    cheb_single::Cint
    sync_pass2::Cint
    galerkin_full::Cint
    basis_fp16::Cint
    extrapolate::Cint
    galerkin_split::Cint
end

"""set_solver!(ctx; tol = 1e-6, ...): the cg! settings of the loop's quasi-static
solves (src/main.jl:256: tol = 1e-6, previous-step start)."""
function set_solver!(ctx::Context; tol = 1e-6, maxiter = 0, pc = EQ_PC_AMG, prec = EQ_PREC_FP64,
                     guess = EQ_GUESS_PREVIOUS, record_snapshots = true, graphs = false)
    # cheb_single, sync_pass2, galerkin_full, basis_fp16, extrapolate, galerkin_split at the library's defaults
    o = SolverOpts(tol, maxiter, pc, prec, guess, record_snapshots, graphs, 0, 0, 0, 0, 0, 0)
    check(ctx, ccall((:eq_set_solver, LIB), Cint, (Ptr{Cvoid}, Ref{SolverOpts}), ctx.h, o))
end

"""state_init!(ctx, v_init; psi_init = nothing): the initial state of
src/main.jl:99-190 (v = v_init, then v .-= Vpl/2; d = a = 0; psi = psi_init,
or the formula of src/main.jl:125 when nothing)."""
function state_init!(ctx::Context, v_init; psi_init = nothing)
    psi = psi_init === nothing ? nothing : convert(Vector{Float64}, psi_init)
    GC.@preserve psi check(ctx, ccall((:eq_state_init, LIB), Cint, (Ptr{Cvoid}, Cdouble, Ptr{Float64}),
                                      ctx.h, v_init, psi === nothing ? Ptr{Float64}(C_NULL) : pointer(psi)))
end

struct StepInfo                        # eq_step_info, field for field
    t::Float64
    dt_used::Float64
    dt_next::Float64
    Vfmax::Float64
    it::Int64
    isolver_used::Int64
    isolver_next::Int64
    pcg_iters::NTuple{2,Int64}
    nr_noconv::Int64
end

"""step!(ctx): one pass of the loop body (src/main.jl:221-451) without the
output bookkeeping; nothing after the reference's NR SEARCH FAILED! report."""
function step!(ctx::Context)
    info = Ref{StepInfo}()
    rc = ccall((:eq_step, LIB), Cint, (Ptr{Cvoid}, Ref{StepInfo}), ctx.h, info)
    if rc == EQ_NR_NAN
        println(lasterr(ctx))
        @error("NR SEARCH FAILED!")
        return
    end
    check(ctx, rc)
    return info[]
end

mutable struct Results                 # eq_results, field for field
    cap_steps::Int64; cap_events::Int64; cap_is::Int64; cap_seis::Int64; cap_rec::Int64
    time_::Ptr{Float64}; Vfmax::Ptr{Float64}; isolver::Ptr{Int64}; pcg_iters::Ptr{Int64}
    tStart::Ptr{Float64}; tEnd::Ptr{Float64}; hypo::Ptr{Float64}
    taubefore::Ptr{Float64}; tauafter::Ptr{Float64}; delfafter::Ptr{Float64}
    is_slip::Ptr{Float64}; is_slipvel::Ptr{Float64}; is_stress::Ptr{Float64}
    seismic_slip::Ptr{Float64}; seismic_slipvel::Ptr{Float64}; seismic_stress::Ptr{Float64}
    dSeis::Ptr{Float64}; vSeis::Ptr{Float64}; aSeis::Ptr{Float64}
    n_steps::Int64; n_start::Int64; n_end::Int64; n_is::Int64; n_seis::Int64; n_rec::Int64
    n_qs::Int64; n_dyn::Int64; total_pcg_iters::Int64; nr_noconv::Int64
    wall_qs::Float64; wall_dyn::Float64
end

"""main(ctx, P; v_init, psi_init = nothing, total_time = P[2].Total_time, ...):
the whole loop of src/main.jl:67-476 on the device, from the initial state
(state_init!) to total_time.  Returns the fields of struct results
(src/main.jl:479-498) under the reference's names, fault matrices as
[FltNglob x n] and receiver traces as [n x receivers].  After NR SEARCH
FAILED! the results up to that step are returned."""
function main(ctx::Context, P; v_init, psi_init = nothing, total_time = P[2].Total_time,
              cap_steps = 200_000, cap_events = 400, cap_is = 4_000, cap_seis = 30_000, cap_rec = 190_000,
              prec = EQ_PREC_FP64, guess = EQ_GUESS_PREVIOUS)
    F = P[1].FltNglob
    nr = length(P[4].out_seis)
    set_solver!(ctx; prec = prec, guess = guess)
    state_init!(ctx, v_init; psi_init = psi_init)
    time_ = zeros(cap_steps); Vfmax = zeros(cap_steps); iso = zeros(Int64, cap_steps)
    iters = zeros(Int64, 2 * cap_steps)
    tStart = zeros(cap_events); tEnd = zeros(cap_events); hypo = zeros(cap_events)
    taub = zeros(F, cap_events); taua = zeros(F, cap_events); delf = zeros(F, cap_events)
    is_slip = zeros(F, cap_is); is_slipvel = zeros(F, cap_is); is_stress = zeros(F, cap_is)
    s_slip = zeros(F, cap_seis); s_slipvel = zeros(F, cap_seis); s_stress = zeros(F, cap_seis)
    dS = zeros(cap_rec, nr); vS = zeros(cap_rec, nr); aS = zeros(cap_rec, nr)
    bufs = (time_, Vfmax, iso, iters, tStart, tEnd, hypo, taub, taua, delf, is_slip, is_slipvel, is_stress,
            s_slip, s_slipvel, s_stress, dS, vS, aS)
    R = GC.@preserve bufs begin
        r = Results(cap_steps, cap_events, cap_is, cap_seis, cap_rec,
                    pointer(time_), pointer(Vfmax), pointer(iso), pointer(iters),
                    pointer(tStart), pointer(tEnd), pointer(hypo), pointer(taub), pointer(taua), pointer(delf),
                    pointer(is_slip), pointer(is_slipvel), pointer(is_stress),
                    pointer(s_slip), pointer(s_slipvel), pointer(s_stress),
                    pointer(dS), pointer(vS), pointer(aS), 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0.0, 0.0)
        rc = ccall((:eq_run, LIB), Cint, (Ptr{Cvoid}, Cdouble, Int64, Ref{Results}), ctx.h, total_time, 0, r)
        if rc == EQ_NR_NAN
            println(lasterr(ctx))
            @error("NR SEARCH FAILED!")
        else
            check(ctx, rc)
        end
        r
    end
    ns = min(R.n_steps, cap_steps); ne = min(R.n_end, cap_events); nst = min(R.n_start, cap_events)
    nis = min(R.n_is, cap_is); nse = min(R.n_seis, cap_seis); nrc = min(R.n_rec, cap_rec)
    return (seismic_stress = s_stress[:, 1:nse], seismic_slipvel = s_slipvel[:, 1:nse],
            seismic_slip = s_slip[:, 1:nse], index_eq = Int64[], is_stress = is_stress[:, 1:nis],
            is_slipvel = is_slipvel[:, 1:nis], is_slip = is_slip[:, 1:nis], dSeis = dS[1:nrc, :],
            vSeis = vS[1:nrc, :], aSeis = aS[1:nrc, :], tStart = tStart[1:nst], tEnd = tEnd[1:ne],
            taubefore = taub[:, 1:nst], tauafter = taua[:, 1:ne], delfafter = delf[:, 1:ne],
            hypo = hypo[1:nst], time_ = time_[1:ns], Vfmax = Vfmax[1:ns])
end

"""main(P): src/main.jl:67-476 with the reference's call signature, on the
current context.  The initial velocity is the reference's v .= 0.5e-3
(src/main.jl:101-102) and psi its formula (:125)."""
main(P) = main(current(), P; v_init = 0.5e-3)

end # module
