"""
    VortexPanelHIP

Julia binding of libvortexhip (the MI355X Galerkin hot path) for VortexPanel.jl: `assemble`, `factorize_galerkin` and
`solve_galerkin` with the reference's signatures (src/assembly/assemble.jl:377-447, src/assembly/solve.jl:51-99),
calling the C ABI of include/vortexhip.h through `ccall`. Indices cross the boundary 0-based; host arrays are
allocated here and kept alive with `GC.@preserve` for the duration of each call; the library owns every device buffer
behind its opaque handles.

Load it next to VortexPanel:

    using VortexPanel
    include(joinpath(VORTEXHIP_ROOT, "vortexpanel.jl_amd", "julia", "VortexPanelHIP.jl"))
    A, b, stats = VortexPanelHIP.assemble(g, dm, els, σ, V∞; opts=AssemblyOptions(), wake=wake)
    F = VortexPanelHIP.factorize_galerkin(dm, A; W=wake_rows(dm, wake), precision=:mixed)
    sol = VortexPanelHIP.solve_galerkin(F, b)

The library path comes from `ENV["VORTEXHIP_LIB"]`, else the in-tree build `vortexpanel.jl_amd/lib/libvortexhip.so`.
"""
module VortexPanelHIP

using LinearAlgebra
using SparseArrays
using StaticArrays
import VortexPanel
import VortexPanel: AssemblyOptions, DOFMapping, RigidWake, SurfaceGeometry, bordered_rows, ExplicitSources, MixedMesh

const libvortexhip = get(ENV, "VORTEXHIP_LIB", normpath(joinpath(@__DIR__, "..", "lib", "libvortexhip.so")))

const VP_OK = Cint(0)
const VP_E_ARGUMENT = Cint(-1)
const VP_E_DIMENSION = Cint(-2)
const VP_E_DEVICE = Cint(-3)
const VP_E_SINGULAR = Cint(-4)
const VP_E_CONVERGENCE = Cint(-5)
const TIER_NAMES = (:self, :edge, :vertex, :contact, :near, :far)   # src/assembly/assemble.jl:260

# ---- C structs (field order and types of include/vortexhip.h) ---------------------------------------------------

struct VPAssemblyOptions      # vp_assembly_options
    test::Cint
    far_ratio::Cdouble
    far_points::Cint
    near_points::Cint
    near_ratio::Cdouble
    near_depth::Cint
    touch_levels::Cint
    touch_ratio::Cdouble
    touch_order::Cint
    corner_angle::Cdouble
    contact_rel::Cdouble
    kernel_far::Cdouble
    precision::Cint
    quad_touch_levels::Cint
end

struct VPSurface              # vp_surface
    nn::Cint
    x::Ptr{Cdouble}
    nt::Cint
    tris::Ptr{Cint}
    corner_slot::Ptr{Cint}
    nslot::Cint
    ntouch::Cint
    touch::Ptr{Cint}
end

struct VPWake                 # vp_wake
    nw::Cint
    x::Ptr{Cdouble}
    ntri::Cint
    tris::Ptr{Cint}
    nfil::Cint
    filaments::Ptr{Cint}
end

struct VPTreeOptions          # vp_tree_options
    theta::Cdouble
    degree::Cint
    precision::Cint
end

struct VPMixedMesh            # vp_mixed_mesh
    nn::Cint
    x::Ptr{Cdouble}
    nquad::Cint
    quad_nodes::Ptr{Cint}
    quad_slots::Ptr{Cint}
    ntri::Cint
    tri_nodes::Ptr{Cint}
    tri_slots::Ptr{Cint}
    nslot::Cint
    ntouch::Cint
    touch::Ptr{Cint}
end

struct VPConstraints          # vp_constraints
    nslot::Cint
    nq::Cint
    nw::Cint
    nb::Cint
    C_rowptr::Ptr{Cint}
    C_col::Ptr{Cint}
    C_val::Ptr{Cdouble}
    W_rowptr::Ptr{Cint}
    W_col::Ptr{Cint}
    W_val::Ptr{Cdouble}
    e::Ptr{Cdouble}
    c::Ptr{Cdouble}
end

# ---- context and errors -----------------------------------------------------------------------------------------

mutable struct Context
    ptr::Ptr{Cvoid}
    device::Int
end

const DEFAULT = Ref{Union{Nothing,Context}}(nothing)

"""The library context on `device` (one HIP stream; calls on a context are serialised by the caller)."""
function Context(device::Integer=0)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:vp_create, libvortexhip), Cint, (Cint, Ref{Ptr{Cvoid}}), device, out)
    rc == VP_OK || error("vortexhip: no usable HIP device $device (status $rc)")
    ctx = Context(out[], device)
    finalizer(c -> c.ptr == C_NULL || ccall((:vp_destroy, libvortexhip), Cvoid, (Ptr{Cvoid},), c.ptr), ctx)
    return ctx
end

default_context() = something(DEFAULT[], (DEFAULT[] = Context(0)))

last_error(ctx::Context) = unsafe_string(ccall((:vp_last_error, libvortexhip), Cstring, (Ptr{Cvoid},), ctx.ptr))

"""FP64 iterative refinement of a solve did not meet its stopping test (no counterpart: LAPACK's solve has none)."""
struct ConvergenceError <: Exception
    msg::String
end

"""Rethrow a status as the exception the reference raises for the same condition."""
function check(ctx::Context, rc::Cint)
    rc == VP_OK && return nothing
    msg = last_error(ctx)
    rc == VP_E_ARGUMENT && throw(ArgumentError(msg))          # check_options, assemble.jl:43-46
    rc == VP_E_DIMENSION && throw(DimensionMismatch(msg))     # reduce_system, src/dofmap/system.jl:28
    rc == VP_E_SINGULAR && throw(SingularException(0))         # getrf info > 0
    rc == VP_E_CONVERGENCE && throw(ConvergenceError(msg))
    error("vortexhip: $msg")
end

# ---- flat arrays (0-based) --------------------------------------------------------------------------------------

flat3(v::AbstractVector{<:SVector{3}}, ::Type{T}) where {T} = T[p[k] for p in v for k in 1:3]

test_code(s::Symbol) = s === :hat ? 0 : s === :onering ? 1 : s === :dual ? 2 :
    throw(ArgumentError("test must be :hat, :onering or :dual, not $s"))

function c_options(o::AssemblyOptions, precision::Symbol)
    precision in (:fp64, :mixed) || throw(ArgumentError("precision must be :fp64 or :mixed, not $precision"))
    return VPAssemblyOptions(test_code(o.test), o.far_ratio, o.far_points, o.near_points, o.near_ratio, o.near_depth,
        o.touch_levels, o.touch_ratio, o.touch_order, o.corner_angle, o.contact_rel, o.kernel_far,
        precision === :mixed ? 1 : 0, o.quad_touch_levels)
end

struct SurfaceArrays
    x::Vector{Cdouble}
    tris::Vector{Cint}
    corner_slot::Vector{Cint}
    touch::Vector{Cint}
    nslot::Int
end

function SurfaceArrays(g::SurfaceGeometry, dm::DOFMapping)
    x = flat3(g.x, Cdouble)
    tris = Cint[t[k] - 1 for t in g.tris for k in 1:3]
    cs = Cint[c[k] - 1 for c in dm.corner_slot for k in 1:3]
    touch = Cint[v - 1 for p in g.touch_pairs for v in (p[1], p[2])]
    return SurfaceArrays(x, tris, cs, touch, dm.nslot)
end

c_surface(s::SurfaceArrays) = VPSurface(length(s.x) ÷ 3, pointer(s.x), length(s.tris) ÷ 3, pointer(s.tris),
    pointer(s.corner_slot), s.nslot, length(s.touch) ÷ 2, isempty(s.touch) ? C_NULL : pointer(s.touch))

struct WakeArrays
    x::Vector{Cdouble}
    tris::Vector{Cint}
    filaments::Vector{Cint}
end

WakeArrays(w::RigidWake) = WakeArrays(flat3(w.x, Cdouble), Cint[t[k] - 1 for t in w.tris for k in 1:3],
    Cint[v - 1 for f in w.filaments for v in (f[1], f[2])])

c_wake(w::WakeArrays) = VPWake(length(w.x) ÷ 3, pointer(w.x), length(w.tris) ÷ 3, pointer(w.tris),
    length(w.filaments) ÷ 2, isempty(w.filaments) ? C_NULL : pointer(w.filaments))

# ---- assemble (src/assembly/assemble.jl:377-447) ------------------------------------------------------------------

"""
    assemble(g, dm, els, σ, V∞; opts=AssemblyOptions(), external=nothing, wake=nothing, precision=:fp64, ctx)

The reference's `assemble` on the GPU: returns `(A, b, stats)` with A nslot × (nslot + nwake) (row = test slot,
column = source slot then wake node), b and `stats = (pairs, points, time)` with `pairs` and `points` a
`NamedTuple{TIER_NAMES}` of counts (assemble.jl:445), so `stats.pairs.contact` and `sum(stats.pairs)` work as there. `els` is accepted for
signature compatibility; the library builds its element records from `g.x`/`g.tris` (the same construction as
`build_elements`). `external` (an arbitrary closure) cannot cross the C boundary: pass `nothing`.
"""
function assemble(g::SurfaceGeometry, dm::DOFMapping, els, σ::AbstractVector, V∞::SVector{3};
                  opts::AssemblyOptions=AssemblyOptions(), external=nothing, wake::Union{Nothing,RigidWake}=nothing,
                  precision::Symbol=:fp64, ctx::Context=default_context())
    external === nothing || throw(ArgumentError("external velocity closures are not supported by the GPU path"))
    length(σ) == length(g.tris) || throw(DimensionMismatch("σ has $(length(σ)) entries for $(length(g.tris)) triangles"))
    o = c_options(opts, precision)
    s = SurfaceArrays(g, dm)
    w = wake === nothing ? nothing : WakeArrays(wake)
    nw = wake === nothing ? 0 : length(wake.x)
    A = Matrix{Cdouble}(undef, dm.nslot, dm.nslot + nw)     # column-major, exactly the library's layout
    b = Vector{Cdouble}(undef, dm.nslot)
    sig = Vector{Cdouble}(σ)
    vinf = Cdouble[V∞[1], V∞[2], V∞[3]]
    pairs = zeros(Clonglong, 6)
    points = zeros(Clonglong, 6)
    t = @elapsed GC.@preserve s w A b sig vinf pairs points begin
        sc = Ref(c_surface(s))
        wref = Ref(w === nothing ? VPWake(0, C_NULL, 0, C_NULL, 0, C_NULL) : c_wake(w))
        wc = w === nothing ? Ptr{VPWake}(C_NULL) : Base.unsafe_convert(Ptr{VPWake}, wref)
        rc = GC.@preserve wref ccall((:vp_assemble, libvortexhip), Cint,
            (Ptr{Cvoid}, Ref{VPSurface}, Ptr{VPWake}, Ptr{Cdouble}, Ptr{Cdouble}, Ref{VPAssemblyOptions},
             Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Clonglong}, Ptr{Clonglong}),
            ctx.ptr, sc, wc, sig, vinf, Ref(o), A, b, pairs, points)
        check(ctx, rc)
    end
    stats = (pairs=NamedTuple{TIER_NAMES}(Tuple(Int.(pairs))), points=NamedTuple{TIER_NAMES}(Tuple(Int.(points))), time=t)
    return A, b, stats
end

"""
    assemble(g, dm, mesh::MixedMesh, V∞; opts=AssemblyOptions(), external=nothing, wake=nothing, ctx)

The reference's mixed-mesh `assemble` (src/assembly/mixed.jl:444-505) on the GPU: the quads of `mesh` kept whole as
isoparametric quads (inner integrals by quad_field's tensor Gauss and polar rules, test integrals on the parameter
square's tiers with `opts.quad_touch_levels`), every other face a flat triangle; sources -V∞·n. Returns `(A, b,
stats)` as the triangle method, FP64. The element touch pairs are formed from `g.touch_pairs` through
`mesh.element_of` as the reference does (mixed.jl:453-457).
"""
function assemble(g::SurfaceGeometry, dm::DOFMapping, mesh::MixedMesh, V∞::SVector{3};
                  opts::AssemblyOptions=AssemblyOptions(), external=nothing, wake::Union{Nothing,RigidWake}=nothing,
                  ctx::Context=default_context())
    external === nothing || throw(ArgumentError("external velocity closures are not supported by the GPU path"))
    o = c_options(opts, :fp64)
    x = flat3(g.x, Cdouble)
    qn = Cint[v - 1 for q in mesh.quad_nodes for v in q]
    qs = Cint[v - 1 for q in mesh.quad_slots for v in q]
    tn = Cint[v - 1 for t in mesh.tri_nodes for v in t]
    ts = Cint[v - 1 for t in mesh.tri_slots for v in t]
    tp = Cint[]
    for (a, b) in g.touch_pairs
        ea, eb = mesh.element_of[a], mesh.element_of[b]
        ea == eb || append!(tp, (ea - 1, eb - 1))
    end
    w = wake === nothing ? nothing : WakeArrays(wake)
    nw = wake === nothing ? 0 : length(wake.x)
    A = Matrix{Cdouble}(undef, dm.nslot, dm.nslot + nw)
    b = Vector{Cdouble}(undef, dm.nslot)
    vinf = Cdouble[V∞[1], V∞[2], V∞[3]]
    pairs = zeros(Clonglong, 6)
    points = zeros(Clonglong, 6)
    t = @elapsed GC.@preserve x qn qs tn ts tp w A b vinf pairs points begin
        m = Ref(VPMixedMesh(length(g.x), pointer(x), length(mesh.quads), pointer(qn), pointer(qs), length(mesh.tris),
                            pointer(tn), pointer(ts), dm.nslot, length(tp) ÷ 2, pointer(tp)))
        wref = Ref(w === nothing ? VPWake(0, C_NULL, 0, C_NULL, 0, C_NULL) : c_wake(w))
        wc = w === nothing ? Ptr{VPWake}(C_NULL) : Base.unsafe_convert(Ptr{VPWake}, wref)
        rc = GC.@preserve wref ccall((:vp_assemble_mixed, libvortexhip), Cint,
            (Ptr{Cvoid}, Ref{VPMixedMesh}, Ptr{VPWake}, Ptr{Cdouble}, Ref{VPAssemblyOptions}, Ptr{Cdouble},
             Ptr{Cdouble}, Ptr{Clonglong}, Ptr{Clonglong}, Ptr{UInt8}),
            ctx.ptr, m, wc, vinf, Ref(o), A, b, pairs, points, C_NULL)
        check(ctx, rc)
    end
    stats = (pairs=NamedTuple{TIER_NAMES}(Tuple(Int.(pairs))), points=NamedTuple{TIER_NAMES}(Tuple(Int.(points))), time=t)
    return A, b, stats
end

# ---- bordered solve (src/assembly/solve.jl:15-99) -----------------------------------------------------------------

"""
The factored bordered system on the device, with the fields `solve_galerkin(F, b)` reads in the reference's
`GalerkinFactorization` (solve.jl:65-72): `se`, `C`, `nq`, `nb`, `cond`.
"""
mutable struct GalerkinFactorizationHIP{T,M}
    ptr::Ptr{Cvoid}
    ctx::Context
    se::Vector{T}
    C::M
    nq::Int
    nb::Int
    cond::T
end

csr(S::SparseMatrixCSC) = (St = sparse(transpose(S)); (Cint.(St.colptr .- 1), Cint.(St.rowval .- 1), Vector{Cdouble}(St.nzval)))

"""
    factorize_galerkin(dm, A; W=nothing, precision=:fp64, ctx)

Forms M = [CᵀA[C;W] se·e; sc·cᵀ 0] on the device and factors it (`precision = :fp64`: FP64 LU on FP64 MFMA;
`:mixed`: FP32 LU on FP32 MFMA); solves refine in FP64 against the FP64 M.
"""
function factorize_galerkin(dm::DOFMapping, A::AbstractMatrix; W=nothing, precision::Symbol=:fp64,
                            ctx::Context=default_context())
    precision in (:fp64, :mixed) || throw(ArgumentError("precision must be :fp64 or :mixed, not $precision"))
    nw = W === nothing ? 0 : size(W, 1)
    size(A) == (dm.nslot, dm.nslot + nw) || throw(DimensionMismatch("A is $(size(A)), expected ($(dm.nslot), $(dm.nslot + nw))"))
    Ad = Matrix{Cdouble}(A)
    return GC.@preserve Ad factorize_pointer(dm, pointer(Ad), dm.nslot, false, W, precision, ctx)
end

# vp_factorize of A at `Aptr` (column-major, leading dimension lda; device memory when on_device, as a plan's A)
function factorize_pointer(dm::DOFMapping, Aptr::Ptr{Cdouble}, lda::Integer, on_device::Bool, W, precision::Symbol,
                           ctx::Context)
    nw = W === nothing ? 0 : size(W, 1)
    rows = bordered_rows(dm)
    nb = length(rows)
    Cr, Cc, Cv = csr(dm.C)
    Wr, Wc, Wv = W === nothing ? (Cint[0], Cint[], Cdouble[]) : csr(sparse(W))
    e = Cdouble[r.e[i] for r in rows for i in 1:dm.nq]
    c = Cdouble[r.c[i] for r in rows for i in 1:dm.nq]
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve Cr Cc Cv Wr Wc Wv e c begin
        k = VPConstraints(dm.nslot, dm.nq, nw, nb, pointer(Cr), pointer(Cc), pointer(Cv),
            nw == 0 ? C_NULL : pointer(Wr), nw == 0 ? C_NULL : pointer(Wc), nw == 0 ? C_NULL : pointer(Wv),
            pointer(e), pointer(c))
        rc = ccall((:vp_factorize, libvortexhip), Cint,
            (Ptr{Cvoid}, Ptr{Cdouble}, Clonglong, Cint, Ref{VPConstraints}, Cint, Ref{Ptr{Cvoid}}),
            ctx.ptr, Aptr, lda, on_device ? 1 : 0, Ref(k), precision === :mixed ? 1 : 0, out)
        check(ctx, rc)
    end
    n = Ref{Cint}(0)
    se = zeros(Cdouble, nb)
    ccall((:vp_factor_info, libvortexhip), Cint, (Ptr{Cvoid}, Ref{Cint}, Ptr{Cdouble}), out[], n, se)
    cnd = Ref{Cdouble}(0)
    check(ctx, ccall((:vp_factor_cond, libvortexhip), Cint, (Ptr{Cvoid}, Ref{Cdouble}), out[], cnd))
    F = GalerkinFactorizationHIP(out[], ctx, se, dm.C, dm.nq, nb, cnd[])
    finalizer(f -> f.ptr == C_NULL || ccall((:vp_factor_destroy, libvortexhip), Cvoid, (Ptr{Cvoid},), f.ptr), F)
    return F
end

"""
    solve_galerkin(F::GalerkinFactorizationHIP, b)

`solve_galerkin(F, b)` of solve.jl:95-99: `(q, μ, λ, cond)`. `b` may also be a matrix whose columns are right hand
sides (an angle-of-attack sweep solved as one batch); q, μ, λ then have one column per right hand side.
"""
solve_galerkin(F::GalerkinFactorizationHIP, b::AbstractVecOrMat) = Base.structdiff(solve_residual(F, b), NamedTuple{(:residual,)})

"""The solve of `solve_galerkin(F, b)` plus the library's relative residual |M x - rhs|₂ / |rhs|₂ (`residual`)."""
function solve_residual(F::GalerkinFactorizationHIP, b::AbstractVecOrMat)
    B = reshape(Matrix{Cdouble}(reshape(b, size(b, 1), :)), size(b, 1), :)
    nrhs = size(B, 2)
    n = F.nq + F.nb
    R = vcat(Matrix{Cdouble}(F.C' * B), zeros(Cdouble, F.nb, nrhs))
    Rr = Matrix{Cdouble}(transpose(R))          # row-major n × nrhs for the library
    Xr = similar(Rr)
    iters = Ref{Cint}(0)
    resid = Ref{Cdouble}(0)
    GC.@preserve Rr Xr begin
        rc = ccall((:vp_factor_solve, libvortexhip), Cint,
            (Ptr{Cvoid}, Ptr{Cdouble}, Cint, Ptr{Cdouble}, Ref{Cint}, Ref{Cdouble}), F.ptr, Rr, nrhs, Xr, iters, resid)
        check(F.ctx, rc)
    end
    X = Matrix(transpose(Xr))
    q = X[1:F.nq, :]
    λ = F.se .* X[F.nq+1:n, :]
    b isa AbstractVector && return (q=vec(q), μ=F.C * vec(q), λ=vec(λ), cond=F.cond, residual=resid[])
    return (q=q, μ=F.C * q, λ=λ, cond=F.cond, residual=resid[])
end

"""
    solve_galerkin(dm, A, b; W=nothing, precision=:fp64)

The one-shot form of solve.jl:51-59: `(q, μ, λ, cond, residual)`.
"""
function solve_galerkin(dm::DOFMapping, A::AbstractMatrix, b::AbstractVector; W=nothing, precision::Symbol=:fp64,
                        ctx::Context=default_context())
    F = factorize_galerkin(dm, A; W=W, precision=precision, ctx=ctx)
    return solve_residual(F, b)
end

"""
    consistency(dm, A, b; W=nothing)

solve.jl:111-115 on the device: `(rows, columns, rhs)`, the largest absolute row and column sums of A [C; W] over its
largest absolute entry, and |Σb| / Σ|b|.
"""
function consistency(dm::DOFMapping, A::AbstractMatrix, b::AbstractVector; W=nothing, ctx::Context=default_context())
    nw = W === nothing ? 0 : size(W, 1)
    size(A) == (dm.nslot, dm.nslot + nw) || throw(DimensionMismatch("A is $(size(A)), expected ($(dm.nslot), $(dm.nslot + nw))"))
    length(b) == dm.nslot || throw(DimensionMismatch("b has $(length(b)) entries, expected $(dm.nslot)"))
    Cr, Cc, Cv = csr(dm.C)
    Wr, Wc, Wv = W === nothing ? (Cint[0], Cint[], Cdouble[]) : csr(sparse(W))
    Ad, bd = Matrix{Cdouble}(A), Vector{Cdouble}(b)
    out = zeros(Cdouble, 3)
    GC.@preserve Cr Cc Cv Wr Wc Wv Ad bd out begin
        k = VPConstraints(dm.nslot, dm.nq, nw, 0, pointer(Cr), pointer(Cc), pointer(Cv),
            nw == 0 ? C_NULL : pointer(Wr), nw == 0 ? C_NULL : pointer(Wc), nw == 0 ? C_NULL : pointer(Wv),
            C_NULL, C_NULL)
        check(ctx, ccall((:vp_consistency, libvortexhip), Cint,
            (Ptr{Cvoid}, Ptr{Cdouble}, Clonglong, Cint, Ptr{Cdouble}, Ref{VPConstraints}, Ptr{Cdouble}),
            ctx.ptr, Ad, dm.nslot, 0, bd, Ref(k), out))
    end
    return (rows=out[1], columns=out[2], rhs=out[3])
end

# ---- N-body velocity sums (src/post/Post.jl:32-112, src/freewake/step.jl:176-187) --------------------------------

"""
    surface_velocity(g, dm, σ, μ, V∞; wake=nothing, μ_wake=nothing, kernel_far=10, precision=:fp64, ctx)

Post.surface_velocity on the GPU: the velocity at every centroid (`Vector{SVector{3}}`). `μ_wake` is one value per wake
node (`wake_values`).
"""
function surface_velocity(g::SurfaceGeometry, dm::DOFMapping, σ::AbstractVector, μ::AbstractVector, V∞::SVector{3};
                          wake::Union{Nothing,RigidWake}=nothing, μ_wake=nothing, kernel_far::Real=10.0,
                          precision::Symbol=:fp64, ctx::Context=default_context())
    s = SurfaceArrays(g, dm)
    w = wake === nothing ? nothing : WakeArrays(wake)
    sig, m = Vector{Cdouble}(σ), Vector{Cdouble}(μ)
    mw = wake === nothing ? Cdouble[0] : Vector{Cdouble}(μ_wake)
    vinf = Cdouble[V∞[1], V∞[2], V∞[3]]
    u = zeros(Cdouble, 3, length(g.tris))
    GC.@preserve s w sig m mw vinf u begin
        wref = Ref(w === nothing ? VPWake(0, C_NULL, 0, C_NULL, 0, C_NULL) : c_wake(w))
        wp = w === nothing ? Ptr{VPWake}(C_NULL) : Base.unsafe_convert(Ptr{VPWake}, wref)
        rc = GC.@preserve wref ccall((:vp_surface_velocity, libvortexhip), Cint,
            (Ptr{Cvoid}, Ref{VPSurface}, Ptr{VPWake}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble,
             Cint, Ptr{Cdouble}),
            ctx.ptr, Ref(c_surface(s)), wp, sig, m, mw, vinf, kernel_far, precision === :mixed ? 1 : 0, u)
        check(ctx, rc)
    end
    return [SVector(u[1, i], u[2, i], u[3, i]) for i in axes(u, 2)]
end

"""
    body_velocity(g, σ, γ, xs; kernel_far=10, precision=:fp64, ctx)

Post.body_velocity on the GPU (γ: the elements' sheets, `sheet_strengths`).
"""
function body_velocity(g::SurfaceGeometry, σ::AbstractVector, γ::AbstractVector, xs::AbstractVector{<:SVector{3}};
                       kernel_far::Real=10.0, precision::Symbol=:fp64, ctx::Context=default_context())
    x = flat3(g.x, Cdouble)
    tris = Cint[t[k] - 1 for t in g.tris for k in 1:3]
    sig, gam, X = Vector{Cdouble}(σ), flat3(γ, Cdouble), flat3(xs, Cdouble)
    u = zeros(Cdouble, 3, length(xs))
    GC.@preserve x tris sig gam X u begin
        sf = VPSurface(length(g.x), pointer(x), length(g.tris), pointer(tris), C_NULL, 0, 0, C_NULL)
        check(ctx, ccall((:vp_body_velocity, libvortexhip), Cint,
            (Ptr{Cvoid}, Ref{VPSurface}, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Ptr{Cdouble}, Cdouble, Cint, Ptr{Cdouble}),
            ctx.ptr, Ref(sf), sig, gam, length(xs), X, kernel_far, precision === :mixed ? 1 : 0, u))
    end
    return [SVector(u[1, i], u[2, i], u[3, i]) for i in axes(u, 2)]
end

"""
    blob_velocities(blobs, X; precision=:fp64, tree=nothing, ctx)

FreeWake.blob_velocities on the GPU: the velocity of Rosenhead–Moore particles (`Blob`: x, α, δ) at the points X.
With `tree = (θ, degree)` the sums are approximated by the treecode (`vp_blob_velocities_tree`, in the same
precision; no counterpart in the reference): target clusters well separated from a source cluster interpolate its field from the
(degree+1)^3 Chebyshev points of their box, every other pair is summed directly, each blob with its own core.
"""
function blob_velocities(blobs::AbstractVector, X::AbstractVector{<:SVector{3}}; precision::Symbol=:fp64,
                         tree::Union{Nothing,Tuple{Real,Integer}}=nothing, ctx::Context=default_context())
    bx = flat3([b.x for b in blobs], Cdouble)
    ba = flat3([b.α for b in blobs], Cdouble)
    bd = Cdouble[b.δ for b in blobs]
    Xf = flat3(X, Cdouble)
    u = zeros(Cdouble, 3, length(X))
    GC.@preserve bx ba bd Xf u begin
        if tree === nothing
            check(ctx, ccall((:vp_blob_velocities, libvortexhip), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Ptr{Cdouble}, Cint, Ptr{Cdouble}),
                ctx.ptr, length(blobs), bx, ba, bd, length(X), Xf, precision === :mixed ? 1 : 0, u))
        else
            check(ctx, ccall((:vp_blob_velocities_tree, libvortexhip), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Ptr{Cdouble}, Ref{VPTreeOptions},
                 Ptr{Cdouble}, Ptr{Cvoid}),
                ctx.ptr, length(blobs), bx, ba, bd, length(X), Xf,
                Ref(VPTreeOptions(tree[1], tree[2], precision === :mixed ? 1 : 0)), u, C_NULL))
        end
    end
    return [SVector(u[1, i], u[2, i], u[3, i]) for i in axes(u, 2)]
end

# ---- the free wake's velocity sums as one HIP graph -----------------------------------------------------------------

"""
    WakeField(g; kernel_far, precision)

The free-wake step's velocity sums (`wake_velocities`, src/freewake/step.jl:166-173) as one HIP graph on the device:
the body's sources and sheets and the wake's blobs at the targets.  The body's elements are set up once; every call
rewrites the graph's node arguments, so the particle count may change from step to step.
"""
mutable struct WakeField
    ptr::Ptr{Cvoid}
    ctx::Context
    nt::Int
end

WakeField(g::SurfaceGeometry; kernel_far::Real=10.0, precision::Symbol=:fp64, ctx::Context=default_context()) =
    WakeField(flat3(g.x, Cdouble), Cint[t[k] - 1 for t in g.tris for k in 1:3]; kernel_far, precision, ctx)

# from the body's elements (as `wake_velocities` passes `sim.els`): each triangle with its own corners
WakeField(els::AbstractVector; kernel_far::Real=10.0, precision::Symbol=:fp64, ctx::Context=default_context()) =
    WakeField(Cdouble[e.x[k][d] for e in els for k in 1:3 for d in 1:3], Cint[0:3length(els)-1;]; kernel_far, precision,
              ctx)

function WakeField(x::Vector{Cdouble}, tris::Vector{Cint}; kernel_far::Real=10.0, precision::Symbol=:fp64,
                   ctx::Context=default_context())
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve x tris begin
        sf = VPSurface(length(x) ÷ 3, pointer(x), length(tris) ÷ 3, pointer(tris), C_NULL, 0, 0, C_NULL)
        check(ctx, ccall((:vp_wake_field_create, libvortexhip), Cint,
            (Ptr{Cvoid}, Ref{VPSurface}, Cdouble, Cint, Ref{Ptr{Cvoid}}),
            ctx.ptr, Ref(sf), kernel_far, precision === :mixed ? 1 : 0, out))
    end
    wf = WakeField(out[], ctx, length(tris) ÷ 3)
    finalizer(w -> w.ptr == C_NULL || ccall((:vp_wake_field_destroy, libvortexhip), Cvoid, (Ptr{Cvoid},), w.ptr), wf)
    return wf
end

"""
    (wf::WakeField)(σ, γ, X, blobs; base = nothing)

base (V∞, or nothing) + the body's kernels (σ per triangle, γ the sheet strength per triangle) + the blobs at X.
"""
function (wf::WakeField)(σ::AbstractVector, γ::AbstractVector, X::AbstractVector{<:SVector{3}}, blobs::AbstractVector;
                         base=nothing)
    sig = Vector{Cdouble}(σ)
    gam = flat3(γ, Cdouble)
    Xf = flat3(X, Cdouble)
    bx = flat3([b.x for b in blobs], Cdouble)
    ba = flat3([b.α for b in blobs], Cdouble)
    bd = Cdouble[b.δ for b in blobs]
    bv = base === nothing ? C_NULL : Cdouble[base[1], base[2], base[3]]
    u = zeros(Cdouble, 3, length(X))
    GC.@preserve sig gam Xf bx ba bd bv u begin
        check(wf.ctx, ccall((:vp_wake_field_eval, libvortexhip), Cint,
            (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Ptr{Cdouble}, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble},
             Ptr{Cdouble}, Ptr{Cdouble}),
            wf.ptr, sig, gam, length(X), Xf, length(blobs), bx, ba, bd, bv, u))
    end
    return [SVector(u[1, i], u[2, i], u[3, i]) for i in axes(u, 2)]
end

# ---- explicit_rhs (src/assembly/explicit.jl:31-86) ------------------------------------------------------------------

"""
    ExplicitRHS(els, dm; opts=AssemblyOptions(), ctx)

A `vp_explicit` handle on one body: `explicit_rhs` of known sheets, filaments and particles with the body's plans kept
on the device between calls (the plan of the known elements is moved in place while their counts stay, as a free
wake's rows move every step).  The body is given by its elements (`FlatTriangle`s, as the reference passes `sim.els`):
explicit_rhs pairs them only with the known elements, so they cross as triangles with their own corners.
"""
mutable struct ExplicitRHS
    ptr::Ptr{Cvoid}
    ctx::Context
    nslot::Int
end

function ExplicitRHS(els::AbstractVector, dm::DOFMapping; opts::AssemblyOptions=AssemblyOptions(),
                     ctx::Context=default_context())
    nt = length(els)
    length(dm.corner_slot) == nt || throw(DimensionMismatch("dm has $(length(dm.corner_slot)) triangles, els $nt"))
    x = Cdouble[e.x[k][d] for e in els for k in 1:3 for d in 1:3]
    tris = Cint[0:3nt-1;]
    cs = Cint[c[k] - 1 for c in dm.corner_slot for k in 1:3]
    o = c_options(opts, :fp64)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve x tris cs begin
        sf = VPSurface(3nt, pointer(x), nt, pointer(tris), pointer(cs), dm.nslot, 0, C_NULL)
        check(ctx, ccall((:vp_explicit_create, libvortexhip), Cint,
            (Ptr{Cvoid}, Ref{VPSurface}, Ref{VPAssemblyOptions}, Ref{Ptr{Cvoid}}), ctx.ptr, Ref(sf), Ref(o), out))
    end
    h = ExplicitRHS(out[], ctx, dm.nslot)
    finalizer(h -> h.ptr == C_NULL || ccall((:vp_explicit_destroy, libvortexhip), Cvoid, (Ptr{Cvoid},), h.ptr), h)
    return h
end

function (h::ExplicitRHS)(src::ExplicitSources)
    sx = Cdouble[t.x[k][d] for t in src.sheets for k in 1:3 for d in 1:3]
    sm = Cdouble[m[k] for m in src.sheet_μ for k in 1:3]
    fx = Cdouble[p[d] for s in src.segments for p in (s.a, s.b) for d in 1:3]
    fg = Cdouble[g[k] for g in src.segment_Γ for k in 1:2]
    bx = flat3([b.x for b in src.blobs], Cdouble)
    ba = flat3([b.α for b in src.blobs], Cdouble)
    bd = Cdouble[b.δ for b in src.blobs]
    length(sm) == length(sx) ÷ 3 || throw(DimensionMismatch("one sheet_μ per sheet"))
    length(fg) == length(fx) ÷ 3 || throw(DimensionMismatch("one segment_Γ per segment"))
    b = zeros(Cdouble, h.nslot)
    GC.@preserve sx sm fx fg bx ba bd b begin
        check(h.ctx, ccall((:vp_explicit_rhs, libvortexhip), Cint,
            (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Ptr{Cdouble},
             Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
            h.ptr, length(src.sheets), sx, sm, length(src.segments), fx, fg, length(src.blobs), bx, ba, bd, b))
    end
    return b
end

# one handle per body element vector, so the reference's step! (which calls explicit_rhs(sim.els, dm, src; opts)
# every step) reuses its plans
const EXPLICIT_HANDLES = WeakKeyDict{Any,Tuple{DOFMapping,AssemblyOptions,ExplicitRHS}}()

"""
    explicit_rhs(els, dm, src::ExplicitSources; opts=AssemblyOptions(), ctx)

The reference's `explicit_rhs` (src/assembly/explicit.jl:31) on the GPU: -∫ψ_i n·u_src dA for every slot.  The handle
of `els` is kept between calls (while `dm` and `opts` stay the same), so a free-wake loop pays for its plans once.
"""
function explicit_rhs(els::AbstractVector, dm::DOFMapping, src::ExplicitSources; opts::AssemblyOptions=AssemblyOptions(),
                      ctx::Context=default_context())
    e = get(EXPLICIT_HANDLES, els, nothing)
    if e === nothing || e[1] !== dm || e[2] != opts || e[3].ctx !== ctx
        e = (dm, opts, ExplicitRHS(els, dm; opts=opts, ctx=ctx))
        EXPLICIT_HANDLES[els] = e
    end
    return e[3](src)
end

const WAKE_FIELDS = WeakKeyDict{Any,Tuple{Float64,Symbol,WakeField}}()

"""
    wake_field(els; kernel_far=10.0, precision=:fp64, ctx)

The `WakeField` of a body's element vector, made on the first call and kept while `els` lives: the handle the
reference's `wake_velocities` (src/freewake/step.jl:166-173) uses every step,
`wake_field(sim.els; kernel_far)(sim.σ, sheet_strengths(sim.els, sim.dm, sol.μ), X, blobs; base=sim.V∞)`.
"""
function wake_field(els::AbstractVector; kernel_far::Real=10.0, precision::Symbol=:fp64, ctx::Context=default_context())
    e = get(WAKE_FIELDS, els, nothing)
    if e === nothing || e[1] != kernel_far || e[2] !== precision || e[3].ctx !== ctx
        e = (Float64(kernel_far), precision, WakeField(els; kernel_far, precision, ctx))
        WAKE_FIELDS[els] = e
    end
    return e[3]
end

# ---- the adjoint's device pieces -----------------------------------------------------------------------------------

"""
    solve_transposed(F, r)

Mᵀ x = r with the factorization (the adjoint system of a Galerkin objective; no counterpart in the reference).
"""
function solve_transposed(F::GalerkinFactorizationHIP, r::AbstractVector)
    R = Vector{Cdouble}(r)
    X = similar(R)
    iters, resid = Ref{Cint}(0), Ref{Cdouble}(0)
    GC.@preserve R X check(F.ctx, ccall((:vp_factor_solve_transposed, libvortexhip), Cint,
        (Ptr{Cvoid}, Ptr{Cdouble}, Cint, Ptr{Cdouble}, Ref{Cint}, Ref{Cdouble}), F.ptr, R, 1, X, iters, resid))
    return X
end

# ---- galerkin_doublets and its reverse rule ------------------------------------------------------------------------

"""
    DoubletState(g, dm, wake, V∞, dV∞; opts, precision, ctx)

The forward state of `galerkin_doublets`: a device plan of `g` and `wake` assembled at V∞ with σ = -V∞·n
(`vp_plan_create`, `vp_plan_assemble`), the bordered system factored on the device from the plan's own A
(`vp_plan_device_A`, `vp_factorize` with `A_on_device`), and the solve: the unknowns q, the slot doublets μ = C q
and the wake nodes' doublets W q. The pullback keeps the plan (its rules and A) and the factors.
"""
mutable struct DoubletState
    ctx::Context
    plan::Ptr{Cvoid}
    F::Any
    dm::DOFMapping
    wake::RigidWake
    nn::Int
    V::Vector{Cdouble}
    dV::Vector{Cdouble}
    q::Vector{Float64}
    μ::Vector{Float64}
    μ_wake::Vector{Float64}
    tris::Vector{SVector{3,Int}}
    x::Vector{SVector{3,Float64}}
end

function DoubletState(g::SurfaceGeometry, dm::DOFMapping, wake::RigidWake, V∞::SVector{3}, dV∞::SVector{3};
                      opts::AssemblyOptions=AssemblyOptions(), precision::Symbol=:fp64, ctx::Context=default_context())
    o = c_options(opts, precision)
    s = SurfaceArrays(g, dm)
    w = WakeArrays(wake)
    plan = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve s w begin
        check(ctx, ccall((:vp_plan_create, libvortexhip), Cint,
            (Ptr{Cvoid}, Ref{VPSurface}, Ref{VPWake}, Ref{VPAssemblyOptions}, Ref{Ptr{Cvoid}}),
            ctx.ptr, Ref(c_surface(s)), Ref(c_wake(w)), Ref(o), plan))
    end
    σ = Cdouble[-dot(V∞, normalize(cross(g.x[t[2]] - g.x[t[1]], g.x[t[3]] - g.x[t[1]]))) for t in g.tris]
    V = Cdouble[V∞[1], V∞[2], V∞[3]]
    pairs, points = zeros(Clonglong, 6), zeros(Clonglong, 6)
    check(ctx, ccall((:vp_plan_assemble, libvortexhip), Cint,
        (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Clonglong}, Ptr{Clonglong}), plan[], σ, V, pairs, points))
    b = zeros(Cdouble, dm.nslot)
    check(ctx, ccall((:vp_plan_read, libvortexhip), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}), plan[], C_NULL, b))
    Ad, bd, lda = Ref{Ptr{Cdouble}}(C_NULL), Ref{Ptr{Cdouble}}(C_NULL), Ref{Clonglong}(0)
    check(ctx, ccall((:vp_plan_device_A, libvortexhip), Cint,
        (Ptr{Cvoid}, Ref{Ptr{Cdouble}}, Ref{Ptr{Cdouble}}, Ref{Clonglong}), plan[], Ad, bd, lda))
    W = VortexPanel.wake_rows(dm, wake.te)
    F = factorize_pointer(dm, Ad[], lda[], true, W, precision, ctx)
    sol = solve_galerkin(F, b)
    st = DoubletState(ctx, plan[], F, dm, wake, length(g.x), V, Cdouble[dV∞[1], dV∞[2], dV∞[3]], sol.q, sol.μ,
                      Vector(W * sol.q), copy(g.tris), copy(g.x))
    finalizer(x -> x.plan == C_NULL || ccall((:vp_plan_destroy, libvortexhip), Cvoid, (Ptr{Cvoid},), x.plan), st)
    return st
end

"""
    galerkin_doublets(g, dm, wake, V∞; opts, precision, ctx)

The slot doublets μ of `assemble` + `solve_galerkin` for geometry g with its rigid wake at freestream V∞ (σ = -V∞·n):
the primal of the reverse rule below, which differentiates them with respect to the node coordinates `g.x` and to
V∞ along a direction `dV∞` (for d/dα, dV∞ = dV∞/dα).
"""
galerkin_doublets(g::SurfaceGeometry, dm::DOFMapping, wake::RigidWake, V∞::SVector{3}; kw...) =
    DoubletState(g, dm, wake, V∞, zero(V∞); kw...).μ

"""
    doublets_pullback(st, μ̄) -> (x̄, V̄)

The vector-Jacobian product of μ(x, V∞) with a cotangent μ̄ (nslot): x̄ (a vector of SVector{3} over the nodes) and
V̄, the derivative along `st.dV` (per unit of the direction's parameter). The adjoint system Mᵀψ = [Cᵀμ̄; 0] is solved
with the forward factors (`vp_factor_solve_transposed`); λ = C ψ_q weights the backward assembly (`vp_plan_adjoint`:
d/dx [λᵀ(b - A[μ; μ_wake])] with the plan's rules, and λᵀ db along dV∞); the wake nodes' gradients go to their
trailing edge nodes (the rigid wake moves with them); the zero-mean border rows add -ψ_b ∂(c̃_k·q)/∂x through the
slot areas. The same steps as `vphip.adjoint.WingAdjoint.pullback_mu`.
"""
function doublets_pullback(st::DoubletState, μ̄::AbstractVector)
    dm, F = st.dm, st.F
    n = F.nq + F.nb
    R = zeros(Cdouble, n)
    R[1:F.nq] = dm.C' * Vector{Float64}(μ̄)
    ψ = solve_transposed(F, R)
    λ = Vector{Cdouble}(dm.C * ψ[1:F.nq])
    μfull = vcat(Vector{Cdouble}(st.μ), Vector{Cdouble}(st.μ_wake))
    nw = length(st.wake.x)
    G = zeros(Cdouble, 3 * (st.nn + nw))
    λb = zeros(Cdouble, 1)
    check(st.ctx, ccall((:vp_plan_adjoint, libvortexhip), Cint,
        (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
        st.plan, 1, λ, μfull, st.V, st.dV, G, λb))
    x̄ = [SVector(G[3i-2], G[3i-1], G[3i]) for i in 1:st.nn]
    for k in 1:nw   # a wake node follows its trailing edge node
        j = dm.wake_te.node[st.wake.te[k]]
        x̄[j] += SVector(G[3(st.nn+k)-2], G[3(st.nn+k)-1], G[3(st.nn+k)])
    end
    # the border rows c̃_k = c_k scale/|c_k| (solve.jl:30-35), c_k = Cᵀ(slot_area on body k): only d(c_k·q)/dx at
    # fixed q is left (c̃_k·q = 0), through the slot areas (a third of each corner's triangle), d area/d a = n × (c - b)/2
    ψb = ψ[F.nq+1:n]
    for (k, r) in enumerate(bordered_rows(dm))
        on = dm.slot_body .== r.body
        scale = F.se[k] * sqrt(length(dm.body_q[r.body]))
        coef = -ψb[k] * scale / norm(r.c)
        for (t, tri) in enumerate(st.tris)
            a, b, c = st.x[tri[1]], st.x[tri[2]], st.x[tri[3]]
            nrm = normalize(cross(b - a, c - a))
            cs = dm.corner_slot[t]
            mt = sum(on[cs[j]] ? st.μ[cs[j]] : 0.0 for j in 1:3) / 3
            f = coef * mt / 2
            x̄[tri[1]] += f * cross(nrm, c - b)
            x̄[tri[2]] += f * cross(nrm, a - c)
            x̄[tri[3]] += f * cross(nrm, b - a)
        end
    end
    return x̄, λb[1]
end

# The reverse rule for ChainRulesCore, defined when the caller has loaded it:
#
#     using ChainRulesCore
#     function ChainRulesCore.rrule(::typeof(VortexPanelHIP.galerkin_doublets), g, dm, wake, V∞; dV∞, kw...)
#         st = VortexPanelHIP.DoubletState(g, dm, wake, V∞, dV∞; kw...)
#         function pullback(μ̄)
#             x̄, v̄ = VortexPanelHIP.doublets_pullback(st, unthunk(μ̄))
#             return NoTangent(), Tangent{typeof(g)}(; x=x̄), NoTangent(), NoTangent(), v̄ * dV∞
#         end
#         return st.μ, pullback
#     end
#
# (V̄ = v̄ dV∞ is the cotangent's projection on the direction dV∞: with dV∞ = dV∞/dα the derivative in α is v̄.)

end # module
