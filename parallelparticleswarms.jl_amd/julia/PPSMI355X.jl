# PPSMI355X.jl — Julia binding of libpps for ParallelParticleSwarms.jl.
#
# Loaded as a package extension of ParallelParticleSwarms (v1.4.0), it adds the
# backend type `MI355XBackend` and the SciMLBase methods the reference defines
# for its own backends:
#   init(prob, ParallelSyncPSOKernel|ParallelPSOKernel{<:MI355XBackend})
#       src/init.jl:56-115
#   solve!(cache, alg)      src/solve.jl:11-59
#   reinit!(cache)          src/init.jl:117-141
#   solve(prob, alg)        src/solve.jl:61-67
#   init / solve! / reinit! / solve for HybridPSO{<:MI355XBackend,<:SimpleLBFGS}
#                           src/init.jl:143-168, src/hybrid.jl:100-158
# Every call crosses into libpps through `ccall`; the C declarations are in
# include/pps.h.  User code changes only the backend:
#     solve(prob, ParallelSyncPSOKernel(1000; backend = MI355XBackend()), maxiters = 500)
# The objective must be one of the markers below (an arbitrary Julia closure
# cannot run inside a precompiled HIP kernel); anything else throws.
module PPSMI355X

using SciMLBase, StaticArrays, QuasiMonteCarlo
import ParallelParticleSwarms
import ParallelParticleSwarms: ParallelPSOKernel, ParallelSyncPSOKernel, HybridPSO,
    SimpleLBFGS
import SimpleOptimization
import LineSearch: StrongWolfeLineSearch
import SciMLBase: init, solve!, reinit!, solve, OptimizationProblem, OptimizationStats

const libpps = get(ENV, "PPS_LIBPPS", joinpath(@__DIR__, "..", "lib", "libpps.so"))

# ---- C ABI mirrors (include/pps.h) -----------------------------------------
const PPS_OK = Int32(0)
const PPS_OBJ = Dict(:rosenbrock => 1, :rastrigin => 2, :sphere => 3, :beale => 4,
    :booth => 5, :himmelblau => 6, :quadratic => 7, :boundary_quadratic => 8,
    :rosen2 => 9)
const PPS_MODE_SYNC, PPS_MODE_ASYNC, PPS_MODE_QLOCK = Int32(0), Int32(1), Int32(2)
const PPS_GBEST_PBEST, PPS_GBEST_CURRENT = Int32(0), Int32(1)

struct CProblem            # pps_problem
    objective::Int32
    d::Int32
    lb::Ptr{Float32}
    ub::Ptr{Float32}
    params::Ptr{Float32}
    n_params::Int32
    shift::Ptr{Float32}
    obs::Ptr{Float32}
    n_obs::Int32
    cons::Ptr{Float32}
    n_cons::Int32
    u0::Ptr{Float32}
end

struct CSwarmParams        # pps_swarm_params
    n::Int64
    mode::Int32
    gbest_from::Int32
    clamp::Int32
    launch::Int32
    seed::UInt64
    c1::Float32
    c2::Float32
    qlock_min_improvers::Int32
    grid_per_cu::Int32
    sampler::Int32
    particles_per_workgroup::Int32   # 0 = autotuned or the objective's own (256; 64 for ODE objectives)
end

struct CLbfgsParams        # pps_lbfgs_params
    maxiters::Int32
    history::Int32
    abstol::Float32
    armijo_c::Float32
    armijo_rho::Float32
    max_backtracks::Int32
    linesearch::Int32      # 0 Armijo, 1 Strong-Wolfe
    wolfe_c2::Float32
    alpha_init::Float32
    alpha_max::Float32
    ls_maxiters::Int32
    zoom_maxiters::Int32
    reltol::Float32
end

function check(rc)
    rc == PPS_OK && return nothing
    msg = unsafe_string(ccall((:pps_last_error, libpps), Cstring, ()))
    error("libpps error $rc: $msg")
end

# ---- backend and objective markers -----------------------------------------
struct MI355XBackend
    device::Int
end
MI355XBackend() = MI355XBackend(0)

"Marker objective: `OptimizationProblem(PPSObjective(:rosenbrock), u0, p; lb, ub)`."
struct PPSObjective{S}
    shift::Union{Nothing, Vector{Float32}}
end
PPSObjective(s::Symbol; shift = nothing) = PPSObjective{s}(shift === nothing ? nothing :
                                                           Float32.(collect(shift)))
(f::PPSObjective{S})(x, p) where {S} = eval_f64(f, x, p)

objective_code(::PPSObjective{S}) where {S} = Int32(PPS_OBJ[S])
objective_code(f) = throw(ArgumentError(
    "the MI355X backend runs a closed objective set; wrap it as PPSObjective(:name), got $(typeof(f))"))

const CTX = Ref{Ptr{Cvoid}}(C_NULL)
const ABI_VERSION = Int32(3)   # include/pps.h PPS_ABI_VERSION mirrored by CProblem & co.
function context(b::MI355XBackend)
    if CTX[] == C_NULL
        v = ccall((:pps_abi_version, libpps), Cint, ())
        v == ABI_VERSION || error("$libpps has ABI version $v, this binding mirrors $ABI_VERSION")
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:pps_ctx_create, libpps), Cint, (Cint, Ptr{Ptr{Cvoid}}), b.device, h))
        CTX[] = h[]
    end
    return CTX[]
end

"""
    PPSQuadraticConstraints(linear, quadratic, c)

Constraints g_k(x) = sum_j linear[k,j] x_j + sum_j quadratic[k,j] x_j^2 + c[k] <= 0
(k <= 8) for `OptimizationFunction(f; cons = PPSQuadraticConstraints(...))`.  Callable
as `cons(x, p)` (an SVector of the g_k), so the same problem runs on the reference's
own backends.  test/constraints.jl's `conss` is
`PPSQuadraticConstraints([-1 2; 1 -2; 0 0], [0 0; 0 0; 0.25 1], [-1, 1, -1])`.
"""
struct PPSQuadraticConstraints{M <: AbstractMatrix, V <: AbstractVector}
    linear::M
    quadratic::M
    c::V
end
function (g::PPSQuadraticConstraints)(x, p)
    m = length(g.c)
    return SVector{m}(ntuple(k -> begin
        acc = zero(eltype(x))
        for j in eachindex(x)
            g.linear[k, j] != 0 && (acc += eltype(x)(g.linear[k, j]) * x[j])
        end
        for j in eachindex(x)
            g.quadratic[k, j] != 0 && (acc += eltype(x)(g.quadratic[k, j]) * (x[j] * x[j]))
        end
        acc + eltype(x)(g.c[k])
    end, m))
end
cons_rows(g::PPSQuadraticConstraints, d) = Float32.(vec(permutedims(
    hcat(g.linear, g.quadratic, reshape(g.c, :, 1)))))   # row-major [a | q | c] per constraint

# Keeps the Julia arrays that a CProblem points into alive.
struct ProblemArrays
    lb::Vector{Float32}
    ub::Vector{Float32}
    p::Vector{Float32}
    shift::Union{Nothing, Vector{Float32}}
    cons::Vector{Float32}
    u0::Vector{Float32}
end

function cproblem(prob::OptimizationProblem)
    f = prob.f.f
    d = length(prob.u0)
    (prob.lb === nothing) == (prob.ub === nothing) ||
        throw(ArgumentError("give both lb and ub, or neither"))
    g = prob.f.cons
    g === nothing || g isa PPSQuadraticConstraints || throw(ArgumentError(
        "the MI355X backend evaluates constraints given as PPSQuadraticConstraints"))
    # lcons/ucons are not read: the reference's penalty uses g(x) <= 0 (src/utils.jl:359-361)
    # no bounds: the library fills -Inf/+Inf and starts from the unbounded
    # sampler around u0 (check_init_bounds + src/init.jl:70-74)
    a = ProblemArrays(prob.lb === nothing ? Float32[] : Float32.(collect(prob.lb)),
        prob.ub === nothing ? Float32[] : Float32.(collect(prob.ub)),
        prob.p === nothing || prob.p isa SciMLBase.NullParameters ? Float32[] :
        Float32.(collect(prob.p)), f isa PPSObjective ? f.shift : nothing,
        g === nothing ? Float32[] : cons_rows(g, d), Float32.(collect(prob.u0)))
    ptr(v) = isempty(v) ? Ptr{Float32}(C_NULL) : pointer(v)
    c = CProblem(objective_code(f), Int32(d), ptr(a.lb), ptr(a.ub), ptr(a.p),
        Int32(length(a.p)), a.shift === nothing ? Ptr{Float32}(C_NULL) : pointer(a.shift),
        Ptr{Float32}(C_NULL), Int32(0), ptr(a.cons), Int32(g === nothing ? 0 : length(g.c)),
        pointer(a.u0))
    return c, a
end

function eval_f64(f::PPSObjective, x, p)
    d = length(x)
    prob = (; f = (; f = f, cons = nothing), u0 = x, p = p,
        lb = fill(-Inf32, d), ub = fill(Inf32, d))
    a = ProblemArrays(Float32[], Float32[], p === nothing ? Float32[] : Float32.(collect(p)),
        f.shift, Float32[], Float32[])
    c = CProblem(objective_code(f), Int32(d), C_NULL, C_NULL,
        isempty(a.p) ? Ptr{Float32}(C_NULL) : pointer(a.p), Int32(length(a.p)),
        a.shift === nothing ? Ptr{Float32}(C_NULL) : pointer(a.shift), C_NULL, Int32(0),
        C_NULL, Int32(0), C_NULL)
    xs = Float64.(collect(x))
    out = Ref{Float64}(0.0)
    GC.@preserve a xs begin
        check(ccall((:pps_eval_f64, libpps), Cint,
            (Ref{CProblem}, Int64, Ptr{Float64}, Ref{Float64}), c, 1, xs, out))
    end
    return out[]
end

# ---- cache (the role of the reference's PSOCache, src/init.jl:1-7) ---------
mutable struct MI355XCache{P, A, S}
    prob::P
    alg::A
    handle::Ptr{Cvoid}
    gbest_pos::Vector{Float32}   # the reference's cache.gbest
    gbest_cost::Float32
    sampler::S                   # the reference's cache.sampler
end

"""
    PPSLatinHypercube()

Latin hypercube initial swarm drawn on the device (one particle per stratum of
every coordinate; a fresh design on every reinit!).  QuasiMonteCarlo samplers
(`LatinHypercubeSample()`, `SobolSample()`, ...) are sampled on the host exactly
as the reference does (src/init.jl:15-36) and handed to the library.
"""
struct PPSLatinHypercube end

const PPS_SAMPLER_UNIFORM, PPS_SAMPLER_UNBOUNDED, PPS_SAMPLER_LATIN_HYPERCUBE =
    Int32(0), Int32(1), Int32(2)
sampler_code(::ParallelParticleSwarms.GPUUniformSampler) = PPS_SAMPLER_UNIFORM
sampler_code(::ParallelParticleSwarms.GPUUnboundedSampler) = PPS_SAMPLER_UNBOUNDED
sampler_code(::PPSLatinHypercube) = PPS_SAMPLER_LATIN_HYPERCUBE
sampler_code(::QuasiMonteCarlo.SamplingAlgorithm) = PPS_SAMPLER_UNIFORM  # host-sampled

host_sampled(prob, sampler) = sampler isa QuasiMonteCarlo.SamplingAlgorithm &&
    prob.lb !== nothing && !(all(isinf, prob.lb) && all(isinf, prob.ub))

# QuasiMonteCarlo.sample(n, lb, ub, sampler) is d x n; its memory is the
# particle-major n x d layout pps_swarm_init_positions takes.
host_positions(prob, n, sampler) =
    Float32.(QuasiMonteCarlo.sample(n, collect(prob.lb), collect(prob.ub), sampler))

function free!(c::MI355XCache)
    if c.handle != C_NULL
        ccall((:pps_swarm_free, libpps), Cint, (Ptr{Cvoid},), c.handle)
        c.handle = C_NULL
    end
end

mode_of(::ParallelSyncPSOKernel) = PPS_MODE_SYNC
mode_of(a::ParallelPSOKernel) = a.global_update ? PPS_MODE_QLOCK : PPS_MODE_ASYNC

function read_gbest(h, d)
    pos = zeros(Float32, d)
    cost = Ref{Float32}(0)
    idx = Ref{Int64}(0)
    check(ccall((:pps_swarm_get_gbest, libpps), Cint,
        (Ptr{Cvoid}, Ptr{Float32}, Ref{Float32}, Ref{Int64}), h, pos, cost, idx))
    return pos, cost[], idx[]
end

# The kernels apply the reference's default penalty schedules θ_default,
# γ_default and h = sqrt (src/utils.jl:320-372, src/algorithms.jl:192-203); they
# only matter for constrained problems, where custom ones are refused.
function check_schedules(prob, alg)
    prob.f.cons === nothing && return nothing
    (alg.θ === ParallelParticleSwarms.θ_default && alg.γ === ParallelParticleSwarms.γ_default &&
     alg.h === sqrt) || throw(ArgumentError(
        "the MI355X backend applies the default penalty schedules (θ_default, γ_default, " *
        "h = sqrt); custom θ, γ or h with constraints are not supported"))
    return nothing
end

function init(prob::OptimizationProblem,
        alg::Union{ParallelSyncPSOKernel{<:MI355XBackend}, ParallelPSOKernel{<:MI355XBackend}},
        args...; seed::Integer = rand(UInt64),
        sampler = ParallelParticleSwarms.GPUUniformSampler(), kwargs...)
    check_schedules(prob, alg)
    c, arrays = cproblem(prob)
    prm = CSwarmParams(alg.num_particles, mode_of(alg), PPS_GBEST_PBEST, Int32(0), Int32(0),
        UInt64(seed), 1.4962f0, 1.4962f0, Int32(2), Int32(0), sampler_code(sampler), Int32(0))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve arrays begin
        if host_sampled(prob, sampler)
            x0 = host_positions(prob, alg.num_particles, sampler)
            check(ccall((:pps_swarm_init_positions, libpps), Cint,
                (Ptr{Cvoid}, Ref{CProblem}, Ref{CSwarmParams}, Ptr{Float32}, Ptr{Ptr{Cvoid}}),
                context(alg.backend), c, prm, x0, h))
        else
            check(ccall((:pps_swarm_init, libpps), Cint,
                (Ptr{Cvoid}, Ref{CProblem}, Ref{CSwarmParams}, Ptr{Ptr{Cvoid}}),
                context(alg.backend), c, prm, h))
        end
    end
    pos, cost, _ = read_gbest(h[], length(prob.u0))
    cache = MI355XCache(prob, alg, h[], pos, cost, sampler)
    finalizer(free!, cache)
    return cache
end

function solve!(cache::MI355XCache, args...; maxiters = 100, w = 0.7298f0, wdamp = 1.0f0,
        kwargs...)
    h, d = cache.handle, length(cache.prob.u0)
    if cache.alg isa ParallelSyncPSOKernel
        # the reference's sync solve! starts from cache.gbest every call
        check(ccall((:pps_swarm_set_gbest, libpps), Cint,
            (Ptr{Cvoid}, Ptr{Float32}, Float32), h, cache.gbest_pos, cache.gbest_cost))
    end
    t0 = time()
    check(ccall((:pps_swarm_run, libpps), Cint, (Ptr{Cvoid}, Int32, Float32, Float32),
        h, Int32(maxiters), Float32(w), Float32(wdamp)))
    t1 = time()
    pos, cost, _ = read_gbest(h, d)
    if !(cache.alg isa ParallelSyncPSOKernel)
        cache.gbest_pos, cache.gbest_cost = pos, cost
    end
    n = cache.alg.num_particles
    aos = zeros(Float32, d * n)
    check(ccall((:pps_swarm_get_positions, libpps), Cint, (Ptr{Cvoid}, Ptr{Float32}), h, aos))
    u0 = cache.prob.u0
    original = [typeof(u0)(view(aos, ((i - 1) * d + 1):(i * d))) for i in 1:n]
    u = typeof(u0)(pos)
    return SciMLBase.build_solution(
        SciMLBase.DefaultOptimizationCache(cache.prob.f, cache.prob.p), cache.alg,
        u, cache.prob.f(u, cache.prob.p); original = original,
        stats = OptimizationStats(; time = t1 - t0))
end

function reinit!(cache::MI355XCache)
    if host_sampled(cache.prob, cache.sampler)
        x0 = host_positions(cache.prob, cache.alg.num_particles, cache.sampler)
        check(ccall((:pps_swarm_reinit_positions, libpps), Cint, (Ptr{Cvoid}, Ptr{Float32}),
            cache.handle, x0))
    else
        check(ccall((:pps_swarm_reinit, libpps), Cint, (Ptr{Cvoid},), cache.handle))
    end
    cache.gbest_pos, cache.gbest_cost, _ = read_gbest(cache.handle, length(cache.prob.u0))
    return nothing
end

function solve(prob::OptimizationProblem,
        alg::Union{ParallelSyncPSOKernel{<:MI355XBackend}, ParallelPSOKernel{<:MI355XBackend}},
        args...; maxiters = 100, kwargs...)
    return solve!(init(prob, alg; kwargs...); maxiters, kwargs...)
end

# parameter_estim_ode! (src/ode_pso.jl:54-170) for the Lotka-Volterra model,
# solved inside the fused PSO kernel.  `data` is 2 x n_obs (sample k at
# t = k * dt * saveat_steps from u0, or k * saveat when `saveat` is given);
# ode_alg :rk4 or :tsit5 (the reference's default GPUTsit5), adaptive = true for
# its Val(true) path (dt = initial step).  `saveat` (an interval, :tsit5 only)
# decouples the steps from the samples, which are read from Tsit5's
# interpolant as GPUTsit5 fills saveat (the reference's benchmark: dt = 0.1,
# saveat = 0.01, benchmarks/Lotka_Volterra/lotka_volterra.jl:124-127).
# `model` :lotka_volterra (rates α, β, γ, δ) or :fitzhugh_nagumo (a, b, τ⁻¹, l;
# benchmarks/Fitzhugh_Nagumo/ode_estimation.jl:6-12).
# Returns (gbest position, FP32 loss).
function parameter_estim_lv(data::AbstractMatrix, lb, ub; backend = MI355XBackend(),
        num_particles = 2^14, maxiters = 300, u0 = (1.0f0, 1.0f0), dt = 0.01f0,
        saveat_steps = 10, saveat = nothing, ode_alg = :tsit5, adaptive = false,
        abstol = 1.0f-6, reltol = 1.0f-3, max_steps = 10000, w = 0.7298f0, wdamp = 1.0f0,
        model = :lotka_volterra, seed::Integer = rand(UInt64))
    (ode_alg === :rk4 && adaptive) && throw(ArgumentError("adaptive stepping uses :tsit5"))
    (ode_alg === :rk4 && saveat !== nothing) &&
        throw(ArgumentError("saveat between steps uses :tsit5's interpolant"))
    base = model === :lotka_volterra ? Int32(16) : model === :fitzhugh_nagumo ? Int32(21) :
           throw(ArgumentError("model is :lotka_volterra or :fitzhugh_nagumo"))
    if saveat !== nothing
        code = base + (adaptive ? Int32(4) : Int32(3))
        par = Float32[u0[1], u0[2], saveat, dt, abstol, reltol, max_steps]
    else
        code = base + (ode_alg === :rk4 ? Int32(0) : adaptive ? Int32(2) : Int32(1))
        par = adaptive ?
              Float32[u0[1], u0[2], Float32(dt) * saveat_steps, dt, abstol, reltol, max_steps] :
              Float32[u0[1], u0[2], dt, saveat_steps]
    end
    obs = Float32.(vec(data))                      # [u1(t0), u2(t0), u1(t1), ...]
    l, u = Float32.(collect(lb)), Float32.(collect(ub))
    z = zeros(Float32, 4)
    c = CProblem(code, Int32(4), pointer(l), pointer(u), pointer(par), Int32(length(par)),
        Ptr{Float32}(C_NULL), pointer(obs), Int32(size(data, 2)), Ptr{Float32}(C_NULL),
        Int32(0), pointer(z))
    prm = CSwarmParams(num_particles, PPS_MODE_SYNC, PPS_GBEST_PBEST, Int32(1), Int32(0),
        UInt64(seed), 1.4962f0, 1.4962f0, Int32(2), Int32(0), PPS_SAMPLER_UNIFORM, Int32(0))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve l u par obs z begin
        check(ccall((:pps_swarm_init, libpps), Cint,
            (Ptr{Cvoid}, Ref{CProblem}, Ref{CSwarmParams}, Ptr{Ptr{Cvoid}}),
            context(backend), c, prm, h))
    end
    try
        check(ccall((:pps_swarm_run, libpps), Cint, (Ptr{Cvoid}, Int32, Float32, Float32),
            h[], Int32(maxiters), Float32(w), Float32(wdamp)))
        pos, cost, _ = read_gbest(h[], 4)
        return pos, cost
    finally
        ccall((:pps_swarm_free, libpps), Cint, (Ptr{Cvoid},), h[])
    end
end

parameter_estim_fhn(data::AbstractMatrix, lb, ub; kwargs...) =
    parameter_estim_lv(data, lb, ub; model = :fitzhugh_nagumo, kwargs...)

# ---- HybridPSO{MI355XBackend, SimpleLBFGS} ----------------------------------
# init builds the PSO stage's cache (src/init.jl:143-159), reinit! renews it
# (src/init.jl:160-168), solve! runs the PSO stage, then projected L-BFGS from
# every particle's best position on the device, and keeps the local argmin when
# it beats the PSO result (src/hybrid.jl:100-158).

struct MI355XHybridCache{C, A}
    pso_cache::C
    alg::A
end

# Other properties read through to the PSO cache, as the reference's
# HybridPSOCache does (src/init.jl:170-176).
Base.getproperty(c::MI355XHybridCache, name::Symbol) =
    name in (:pso_cache, :alg) ? getfield(c, name) : getproperty(getfield(c, :pso_cache), name)

function init(prob::OptimizationProblem, alg::HybridPSO{<:MI355XBackend, <:SimpleLBFGS},
        args...; kwargs...)
    return MI355XHybridCache(init(prob, alg.pso, args...; kwargs...), alg)
end

function reinit!(cache::MI355XHybridCache)
    reinit!(cache.pso_cache)
    return nothing
end

"""
    PPSArmijo(; c = 1f-4, rho = 0.5f0, max_backtracks = 20)

Armijo backtracking for the local phase (the HybridPSO benchmark
configuration's line search; the reference itself accepts only
`StrongWolfeLineSearch`, whose c1, c2, α_init, α_max, maxiters and
zoom_maxiters are forwarded to the kernel as the reference forwards them,
src/hybrid.jl:131-136).
"""
struct PPSArmijo
    c::Float32
    rho::Float32
    max_backtracks::Int32
end
PPSArmijo(; c = 1.0f-4, rho = 0.5f0, max_backtracks = 20) = PPSArmijo(c, rho, max_backtracks)

history_size(::Val{N}) where {N} = Int32(N)
history_size(n::Integer) = Int32(n)

function lbfgs_params(local_maxiters, history, abstol, reltol, ls::StrongWolfeLineSearch)
    return CLbfgsParams(Int32(local_maxiters), history,
        abstol === nothing ? 0.0f0 : Float32(abstol), Float32(ls.c1), 0.5f0, Int32(20),
        Int32(1), Float32(ls.c2), Float32(ls.α_init), Float32(ls.α_max), Int32(ls.maxiters),
        Int32(ls.zoom_maxiters), reltol === nothing ? 0.0f0 : Float32(reltol))
end
function lbfgs_params(local_maxiters, history, abstol, reltol, ls::PPSArmijo)
    return CLbfgsParams(Int32(local_maxiters), history,
        abstol === nothing ? 0.0f0 : Float32(abstol), ls.c, ls.rho, ls.max_backtracks,
        Int32(0), 0.9f0, 1.0f0, 65536.0f0, Int32(20), Int32(30),
        reltol === nothing ? 0.0f0 : Float32(reltol))
end
lbfgs_params(local_maxiters, history, abstol, reltol, ls) = throw(ArgumentError(
    "HybridPSO with SimpleLBFGS on MI355XBackend requires a StrongWolfeLineSearch " *
    "(or PPSArmijo), got $(typeof(ls))"))

function solve!(cache::MI355XHybridCache, args...; maxiters = 100, local_maxiters = 50,
        abstol = nothing, reltol = nothing,
        linesearch = cache.alg.local_opt.linesearch, w = 0.7298f0, wdamp = 1.0f0, kwargs...)
    opt = cache.alg
    lp = lbfgs_params(local_maxiters,
        history_size(SimpleOptimization.__get_threshold(opt.local_opt)), abstol, reltol,
        linesearch)
    sol_pso = solve!(cache.pso_cache; maxiters, w, wdamp)
    prob = cache.pso_cache.prob
    d = length(prob.u0)
    bx = zeros(Float32, d)
    bf = Ref{Float32}(0)
    bi = Ref{Int64}(0)
    t0 = time()
    check(ccall((:pps_swarm_lbfgs, libpps), Cint,
        (Ptr{Cvoid}, Ref{CLbfgsParams}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32},
            Ref{Float32}, Ref{Int64}),
        cache.pso_cache.handle, lp, C_NULL, C_NULL, bx, bf, bi))
    best_u, best_obj = sol_pso.u, sol_pso.objective
    if isfinite(bf[])   # non-finite local results are +Inf (src/lbfgs.jl:20)
        local_u = typeof(prob.u0)(bx)
        local_obj = prob.f(local_u, prob.p)
        if local_obj < best_obj
            best_u, best_obj = local_u, local_obj
        end
    end
    return SciMLBase.build_solution(
        SciMLBase.DefaultOptimizationCache(prob.f, prob.p), opt, best_u, best_obj;
        stats = OptimizationStats(; time = (time() - t0) + sol_pso.stats.time))
end

function solve(prob::OptimizationProblem, alg::HybridPSO{<:MI355XBackend, <:SimpleLBFGS},
        args...; maxiters = 100, kwargs...)
    return solve!(init(prob, alg; kwargs...); maxiters, kwargs...)
end

export MI355XBackend, PPSObjective, PPSQuadraticConstraints, PPSLatinHypercube, PPSArmijo,
    parameter_estim_lv, parameter_estim_fhn

end # module
