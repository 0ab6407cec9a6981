# TestParticleAMD.jl -- thin ccall layer from TestParticle.jl to libtpamd.so.
#
# A package next to TestParticle.jl (Project.toml beside src/): it adds a
# method of `solve` that dispatches on the device backend `MI355X()`,
# mirroring the KernelAbstractions entry point
#     solve(prob, ::Boris{false}, backend::KA.Backend, ensemblealg; dt, ...)
# (src/boris/boris_kernel.jl:312-382), and extends it to MultistepBoris and the
# adaptive solvers; and methods that route the reference's own
# `solve(prob, Boris(); dt, ...)` to the GPU when p carries device fields.
# Fields are described with the device menu (DeviceField) rather than Julia
# closures, since a GPU cannot call host code.  `trajectory_loss` and
# `loss_gradient` wrap the adjoint; the ChainRulesCore extension
# (ext/TestParticleAMDChainRulesCoreExt.jl) gives `trajectory_loss` its rrule.
#
# The build container has no Julia, so this file is not run there; the
# struct mirrors and the ccall signatures below are checked against
# include/tpamd.h by the Python tests (tests/test_julia_binding.py), and the
# C ABI itself is exercised through ctypes.
module TestParticleAMD

using TestParticle
using TestParticle: TraceProblem, Boris, MultistepBoris, build_solution
using SciMLBase: EnsembleSolution, ReturnCode, LinearInterpolation
using StaticArrays

const libtpamd = joinpath(@__DIR__, "..", "..", "_build", "libtpamd.so")

struct MI355X
    device::Int
end
MI355X() = MI355X(0)

# mirrors tpamd_solve_desc (include/tpamd.h)
Base.@kwdef mutable struct SolveDesc
    alg::Int32 = 0
    adaptive::Int32 = 0
    order_N::Int32 = 2
    nsub::Int32 = 1
    form::Int32 = 0           # 0 reference (atan, sincos: multistep_boris.jl:94-120), 1 rational
    precision::Int32 = 0
    save_start::Int32 = 1
    save_end::Int32 = 1
    save_everystep::Int32 = 1
    bnd_flags::Int32 = 0
    safety::Float64 = 0.0
    q2m::Float64 = 0.0
    t0::Float64 = 0.0
    t1::Float64 = 0.0
    dt::Float64 = 0.0
    maxiters::Int64 = 1_000_000
    savestepinterval::Int64 = 1
    max_steps::Int64 = 0
    nout_cap::Int64 = 0
    r_in::Float64 = 0.0
    r_out::Float64 = Inf
    box_lo::NTuple{3, Float64} = (-Inf, -Inf, -Inf)
    box_hi::NTuple{3, Float64} = (Inf, Inf, Inf)
    dt_pp::Ptr{Float64} = C_NULL
    nsteps_pp::Int64 = 0
    block_size::Int32 = 0
    flags::Int32 = 0
    mass::Float64 = 0.0
    save_fields::Int32 = 0
    save_work::Int32 = 0
    particles_per_lane::Int32 = 0
    reserved::Int32 = 0
end

# mirrors tpamd_stats (include/tpamd.h)
mutable struct Stats
    kernel_ms::Float64
    total_ms::Float64
    particle_steps::Int64
    nt::Int64
    nout::Int64
    max_len::Int64
    aux_ms::Float64
end
Stats() = Stats(0, 0, 0, 0, 0, 0, 0)

const _ctx = Ref{Ptr{Cvoid}}(C_NULL)

function context(dev::Int)
    if _ctx[] == C_NULL
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:tpamd_init, libtpamd), Int32, (Int32, Ptr{Ptr{Cvoid}}), dev, h)
        rc == 0 || error(unsafe_string(ccall((:tpamd_last_error, libtpamd), Cstring, (Ptr{Cvoid},), C_NULL)))
        _ctx[] = h[]
    end
    return _ctx[]
end

# tpamd_status: 1-4 are the reference's ArgumentErrors (boris.jl:349-356,
# types.jl:70-75); 9 (TPAMD_ERR_NONFINITE) a loss or gradient that is not finite
function check(ctx, rc)
    rc == 0 && return
    msg = unsafe_string(ccall((:tpamd_last_error, libtpamd), Cstring, (Ptr{Cvoid},), ctx))
    rc in (1, 2, 3, 4) && throw(ArgumentError(msg))
    rc == 9 && throw(DomainError(rc, msg))
    error(msg)
end

# fixed step or adaptive, from the algorithm's type (src/types.jl:51-107): a
# fixed-step MultistepBoris{N, false}(; n) still carries safety = 0.1
is_adaptive(::Boris{A}) where {A} = A
is_adaptive(::MultistepBoris{N, A}) where {N, A} = A
multistep_order(::MultistepBoris{N}) where {N} = N

# --- device field menu -------------------------------------------------------
abstract type DeviceField end
struct ZeroDeviceField <: DeviceField end
struct UniformDeviceField <: DeviceField
    v::SVector{3, Float64}
end
struct DipoleDeviceField <: DeviceField
    b0::Float64
    radius::Float64
end
"Gridded field: A is Julia's 3 x nx x ny x nz array on uniform ranges x, y, z."
struct GridDeviceField{T} <: DeviceField
    A::Array{T, 4}
    x::AbstractRange
    y::AbstractRange
    z::AbstractRange
    extrap::Int32     # 0 FillExtrap(NaN), 1 ClampExtrap(), 2 periodic (exclusive)
    storage::Int32    # 0 Float64, 1 Float32
end
struct UxBDeviceField <: DeviceField
    u::SVector{3, Float64}
    B::DeviceField
end
"""
Gridded field on a general mesh: `build_interpolator(gridtype, A, g1, g2, g3, order, bc)`
for RectilinearGrid / non-uniform CartesianGrid axes, order 3 (uniform axes),
StructuredGrid (`coord = 1`, axes r, θ, ϕ) and 2D / 1D fields (an axis of
one node is ignored). `extrap` per axis: 0 fill NaN, 1 clamp, 2 exclusive
periodic, 3 inclusive periodic.
"""
struct MeshDeviceField{T} <: DeviceField
    A::Array{T, 4}
    axes::NTuple{3, Vector{Float64}}
    coord::Int32
    order::Int32
    extrap::NTuple{3, Int32}
    storage::Int32
end
"""
Time-dependent gridded field, the device form of
`LazyTimeInterpolator(times, loader)` (src/utility/interpolation.jl:337-430):
`loader(i)` returns the `GridDeviceField` of snapshot i on one shared grid.
Every snapshot is loaded once and kept resident on the device.
"""
struct TimeGridDeviceField <: DeviceField
    times::Vector{Float64}
    loader::Function
end

handle(ctx, ::ZeroDeviceField, cache) = _mk(ctx) do h
    ccall((:tpamd_field_zero, libtpamd), Int32, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}), ctx, h)
end
handle(ctx, f::UniformDeviceField, cache) = _mk(ctx) do h
    v = Vector(f.v)
    ccall((:tpamd_field_uniform, libtpamd), Int32, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Ptr{Cvoid}}), ctx, v, h)
end
handle(ctx, f::DipoleDeviceField, cache) = _mk(ctx) do h
    ccall((:tpamd_field_dipole, libtpamd), Int32, (Ptr{Cvoid}, Float64, Float64, Ptr{Ptr{Cvoid}}),
          ctx, f.b0, f.radius, h)
end
function handle(ctx, f::GridDeviceField{T}, cache) where {T}
    get!(cache, f) do
        _mk(ctx) do h
            origin = [first(f.x), first(f.y), first(f.z)]
            spacing = [step(f.x), step(f.y), step(f.z)]
            GC.@preserve f ccall((:tpamd_field_grid, libtpamd), Int32,
                (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Int64, Int64, Int64, Ptr{Float64}, Ptr{Float64},
                 Int32, Int32, Ptr{Ptr{Cvoid}}),
                ctx, f.A, T === Float64 ? 0 : 1, size(f.A, 2), size(f.A, 3), size(f.A, 4),
                origin, spacing, f.extrap, f.storage, h)
        end
    end
end
function handle(ctx, f::MeshDeviceField{T}, cache) where {T}
    get!(cache, f) do
        _mk(ctx) do h
            n = Int64[size(f.A, 2), size(f.A, 3), size(f.A, 4)]
            axp = [pointer(a) for a in f.axes]
            ex = Int32[f.extrap...]
            GC.@preserve f ccall((:tpamd_field_mesh, libtpamd), Int32,
                (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{Int64}, Ptr{Ptr{Float64}}, Int32, Int32,
                 Ptr{Int32}, Int32, Ptr{Ptr{Cvoid}}),
                ctx, f.A, T === Float64 ? 0 : 1, n, axp, f.coord, f.order, ex, f.storage, h)
        end
    end
end
function handle(ctx, f::TimeGridDeviceField, cache)
    get!(cache, f) do
        snaps = [f.loader(i)::GridDeviceField for i in eachindex(f.times)]
        g = snaps[1]
        T = eltype(g.A)
        all(s -> size(s.A) == size(g.A) && eltype(s.A) === T && s.x == g.x && s.y == g.y &&
                 s.z == g.z && s.extrap == g.extrap && s.storage == g.storage, snaps) ||
            throw(ArgumentError("every snapshot must share the grid of snapshot 1"))
        _mk(ctx) do h
            origin = [first(g.x), first(g.y), first(g.z)]
            spacing = [step(g.x), step(g.y), step(g.z)]
            ptrs = [pointer(s.A) for s in snaps]
            GC.@preserve snaps ccall((:tpamd_field_time_grid, libtpamd), Int32,
                (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Int32, Int64, Int64, Int64, Int32, Ptr{Float64},
                 Ptr{Float64}, Ptr{Float64}, Int32, Int32, Ptr{Ptr{Cvoid}}),
                ctx, ptrs, T === Float64 ? 0 : 1, size(g.A, 2), size(g.A, 3), size(g.A, 4),
                length(f.times), f.times, origin, spacing, g.extrap, g.storage, h)
        end
    end
end
function handle(ctx, f::UxBDeviceField, cache)
    hB = handle(ctx, f.B, cache)
    _mk(ctx) do h
        u = Vector(f.u)
        ccall((:tpamd_field_uxb, libtpamd), Int32, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Cvoid}, Ptr{Ptr{Cvoid}}),
              ctx, u, hB, h)
    end
end
function _mk(f, ctx)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ctx, f(h))
    return h[]
end

# --- solve --------------------------------------------------------------------
"""
    solve(prob::TraceProblem, alg, backend::MI355X, [ensemblealg]; dt, trajectories=1,
          savestepinterval=1, isoutside=nothing, save_start=true, save_end=true,
          save_everystep=true, save_fields=false, save_work=false, maxiters=1_000_000,
          precision=:fp64, form=:reference)

Device counterpart of `solve(prob, Boris(), backend; ...)`
(src/boris/boris_kernel.jl:312-382) for Boris, MultistepBoris and the adaptive
variants.  `prob.p = (q2m, m, E::DeviceField, B::DeviceField, F)`.
"""
function TestParticle.solve(prob::TraceProblem, alg, backend::MI355X, ensemblealg = nothing;
        dt = 0.0, trajectories::Int = 1, savestepinterval::Int = 1, isoutside = nothing,
        save_start::Bool = true, save_end::Bool = true, save_everystep::Bool = true,
        save_fields::Bool = false, save_work::Bool = false,
        maxiters::Int = 1_000_000, precision::Symbol = :fp64, form::Symbol = :reference,
        nout_cap::Int = 0)
    ctx = context(backend.device)
    q2m, m, E, B, _ = prob.p
    T = Float64
    # initial states: prob_func on the host, as the reference does (boris.jl:501-503)
    xv0 = Matrix{Float64}(undef, 6, trajectories)
    for i in 1:trajectories
        xv0[:, i] .= prob.prob_func(prob, (sim_id = i, repeat = false)).u0
    end
    d = SolveDesc(q2m = q2m, t0 = prob.tspan[1], t1 = prob.tspan[2], dt = dt,
                  savestepinterval = savestepinterval, maxiters = maxiters,
                  save_start = save_start, save_end = save_end, save_everystep = save_everystep,
                  precision = precision === :fp32c ? 1 : 0, mass = m,
                  save_fields = save_fields, save_work = save_work)
    # record width, as _prepare_saved_data builds it (boris.jl:405-440)
    W = 6 + (save_fields ? 6 : 0) + (save_work ? 4 : 0)
    if alg isa MultistepBoris
        d.alg = 1; d.order_N = multistep_order(alg); d.nsub = alg.n
        d.form = form === :rational ? 1 : 0
    end
    d.adaptive = is_adaptive(alg) ? 1 : 0
    d.safety = is_adaptive(alg) ? alg.safety : 0.0
    isoutside === nothing || apply_boundary!(d, isoutside)
    nt, nout = Ref{Int64}(0), Ref{Int64}(0)
    if d.adaptive == 0
        check(C_NULL, ccall((:tpamd_plan, libtpamd), Int32,
            (Ptr{Cvoid}, Ref{SolveDesc}, Ref{Int64}, Ref{Int64}), C_NULL, d, nt, nout))
        d.nout_cap = nout[]
    else
        d.nout_cap = nout_cap > 0 ? nout_cap : (save_everystep ? 1024 : 2)
    end
    cache = IdDict()
    hE, hB = handle(ctx, E, cache), handle(ctx, B, cache)
    cap = d.nout_cap
    u = Array{Float64}(undef, W, cap, trajectories)
    t = Array{Float64}(undef, cap, trajectories)
    len = Vector{Int32}(undef, trajectories)
    ret = Vector{Int32}(undef, trajectories)
    st = Ref(Stats())
    GC.@preserve xv0 u t len ret begin
        rc = ccall((:tpamd_solve, libtpamd), Int32,
            (Ptr{Cvoid}, Ref{SolveDesc}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Int64,
             Ptr{Float64}, Ptr{Float64}, Ptr{Int32}, Ptr{Int32}, Ref{Stats}),
            ctx, d, hE, hB, xv0, trajectories, u, t, len, ret, st)
        check(ctx, rc)
    end
    sols = map(1:trajectories) do i
        n = len[i]
        ts = t[1:n, i]
        us = [SVector{W, T}(u[:, k, i]) for k in 1:n]
        rcode = ret[i] == 0 ? ReturnCode.Success : ReturnCode.Terminated
        build_solution(prob, :boris, ts, us; interp = LinearInterpolation(ts, us), retcode = rcode)
    end
    return EnsembleSolution(sols, st[].kernel_ms * 1e-3, true)
end

# isoutside menu: (:nan, :sphere_in => R, :sphere_out => R, :box => (lo, hi))
# --- preferred route: device fields in p -----------------------------------
# When p = (q2m, m, E, B, F) carries device fields, the reference's own call
#     TestParticle.solve(prob, Boris(); dt, savestepinterval)
# lands on the GPU.  These methods mirror the reference's solve signatures
# (src/boris/boris.jl:88-106, src/boris/adaptive_boris.jl:21-33) with the
# same algorithm and ensemble types, and add a constraint on p that names
# this module's field types: more specific than the reference's methods in
# every argument, so no ambiguity and no type piracy.
const DeviceParams = Tuple{Any, Any, DeviceField, DeviceField, Any}
const FixedAlg = Union{Boris{false}, MultistepBoris{N, false} where {N}}
const AdaptiveAlg = Union{Boris{true}, MultistepBoris{N, true} where {N}}

function TestParticle.solve(prob::TraceProblem{U, T, I, P}, alg::FixedAlg,
        ensemblealg::TestParticle.BasicEnsembleAlgorithm = TestParticle.EnsembleSerial();
        device::Int = 0, kw...) where {U, T, I, P <: DeviceParams}
    return TestParticle.solve(prob, alg, MI355X(device), ensemblealg; kw...)
end
function TestParticle.solve(prob::TraceProblem{U, T, I, P}, alg::AdaptiveAlg,
        ensemblealg::TestParticle.BasicEnsembleAlgorithm = TestParticle.EnsembleSerial();
        device::Int = 0, kw...) where {U, T, I, P <: DeviceParams}
    return TestParticle.solve(prob, alg, MI355X(device), ensemblealg; kw...)
end

"""
    prepare(x, y, z, E, B; device = MI355X(), order = 1, bc = 1, species, q, m, storage = Float64)
    prepare(E, B; device = MI355X(), species, q, m)

`TestParticle.prepare` for the device path: gridded arrays become
`GridDeviceField`s (bc: 0 fill NaN, 1 clamp, 2 periodic), `DeviceField`s pass
through, and the result routes `TestParticle.solve(prob, Boris(); ...)` to
the GPU (src/prepare.jl:105-170).
"""
function prepare(x::AbstractRange, y::AbstractRange, z::AbstractRange, E, B;
        device::MI355X = MI355X(), bc::Integer = 0, storage::Type = Float64, kw...)
    mk(f) = f isa DeviceField ? f :
        f === nothing ? ZeroDeviceField() :
        GridDeviceField(Array(f), x, y, z, Int32(bc), storage === Float32 ? Int32(1) : Int32(0))
    return prepare(mk(E), mk(B); device, kw...)
end
function prepare(E::DeviceField, B::DeviceField; device::MI355X = MI355X(),
        species = TestParticle.Proton, q = nothing, m = nothing)
    qq = q === nothing ? species.q : q
    mm = m === nothing ? species.m : m
    return (qq / mm, mm, E, B, TestParticle.ZeroField())
end

# --- gradients -----------------------------------------------------------------
"""
    loss_gradient(A, xv0; x, y, z, bc = 2, q2m, tspan, dt, targets, precision = :fp64,
                  ckpt_interval = 64, device = 0) -> (loss, dL_dA, dL_dxv0)

Loss L = mean_i |x_i(T) - x*_i|^2 of the final positions of a fixed-step
standard Boris solve with E = 0 and B the gridded field A (3 x nx x ny x nz
on the uniform ranges x, y, z; `bc` 0 fill NaN, 1 clamp, 2 periodic; `cells`
adds the cell-replicated device copy), and its gradients with respect to every value
of A and every initial state (xv0 and targets are 6 x N and 3 x N).  One
forward pass with FP64 checkpoints every `ckpt_interval` steps, then a reverse
sweep (tpamd_adjoint).  The reference differentiates the same solve with
ForwardDiff (test/test_derivatives.jl).  `precision = :fp32c` runs the steps
in float over float storage.  A loss or gradient that is not finite throws
a DomainError.
"""
function loss_gradient(A::Array{T, 4}, xv0::AbstractMatrix, targets::AbstractMatrix;
        x::AbstractRange, y::AbstractRange, z::AbstractRange, bc::Integer = 2, q2m::Real,
        tspan, dt::Real, precision::Symbol = :fp64, ckpt_interval::Integer = 64,
        cells::Bool = true, device::Integer = 0) where {T <: Union{Float32, Float64}}
    size(A, 1) == 3 || throw(ArgumentError("A must be 3 x nx x ny x nz"))
    size(xv0, 1) == 6 && size(targets) == (3, size(xv0, 2)) ||
        throw(ArgumentError("xv0 must be 6 x N and targets 3 x N"))
    ctx = context(device)
    storage = precision === :fp32c ? Int32(1) : Int32(0)
    hB = handle(ctx, GridDeviceField(A, x, y, z, Int32(bc), storage), IdDict())
    # the cell-replicated copy: one cache line per gather (same values)
    cells && check(ctx, ccall((:tpamd_field_grid_cells, libtpamd), Int32,
                              (Ptr{Cvoid}, Ptr{Cvoid}, Int32), ctx, hB, Int32(1)))
    N = size(xv0, 2)
    d = SolveDesc(q2m = q2m, t0 = tspan[1], t1 = tspan[2], dt = dt, save_start = 0,
                  save_everystep = 0, precision = precision === :fp32c ? 1 : 0, nout_cap = 1)
    x0 = Matrix{Float64}(xv0)
    tg = Matrix{Float64}(targets)
    loss = Ref(0.0)
    gA = Array{Float64}(undef, size(A))           # node layout [k][j][i][c] = 3 x nx x ny x nz
    gx = Matrix{Float64}(undef, 6, N)             # [N][6] = 6 x N
    st = Ref(Stats())
    try
        GC.@preserve x0 tg gA gx begin
            rc = ccall((:tpamd_adjoint, libtpamd), Int32,
                (Ptr{Cvoid}, Ref{SolveDesc}, Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Int32,
                 Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{Stats}),
                ctx, d, hB, x0, N, tg, Int32(ckpt_interval), loss, gA, gx, st)
            check(ctx, rc)
        end
    finally
        ccall((:tpamd_field_free, libtpamd), Int32, (Ptr{Cvoid}, Ptr{Cvoid}), ctx, hB)
    end
    return loss[], gA, gx
end

"""
    trajectory_loss(A, xv0, targets; kw...) -> loss

The loss of `loss_gradient` alone.  With ChainRulesCore loaded, its rrule
returns the adjoint's gradients for A and xv0 from the same call.
"""
trajectory_loss(A, xv0, targets; kw...) = loss_gradient(A, xv0, targets; kw...)[1]

function apply_boundary!(d::SolveDesc, spec)
    for s in (spec isa Tuple ? spec : (spec,))
        if s === :nan
            d.bnd_flags |= 1
        elseif s isa Pair && s.first === :sphere_in
            d.bnd_flags |= 2; d.r_in = s.second
        elseif s isa Pair && s.first === :sphere_out
            d.bnd_flags |= 4; d.r_out = s.second
        elseif s isa Pair && s.first === :box
            d.bnd_flags |= 8; d.box_lo = Tuple(s.second[1]); d.box_hi = Tuple(s.second[2])
        else
            throw(ArgumentError("isoutside must come from the device menu"))
        end
    end
end

end # module
