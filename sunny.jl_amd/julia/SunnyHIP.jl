# SunnyHIP: device-backed `step!` for Sunny.jl systems on MI355X.
#
# The binding a Sunny maintainer would add: a wrapper type built from a
# `System` (the precedent is `EntangledSystem`, src/EntangledUnits/
# TypesAndAliasing.jl:282-286, which forwards `step!`), one `ccall` per
# entry point of include/sunny_hip.h, and `step!` methods that keep
# `Langevin` / `ImplicitMidpoint` unchanged (src/Integrators.jl:56-135).
# Julia is not available in the build container, so this file is the
# documented binding rather than a tested one; the same entry points are
# exercised through Python ctypes in tests/.
module SunnyHIP

import Sunny
using Sunny: System, Langevin, ImplicitMidpoint

const lib = joinpath(@__DIR__, "..", "sunnyhip", "libsunnyhip.so")

struct SnhSpec
    na::Int32; N::Int32; precision::Int32
    kappa::Ptr{Float64}; g::Ptr{Float64}; field::Ptr{Float64}
    stevens::Ptr{Float64}; kmax::Ptr{Int32}; Lambda::Ptr{Float64}
    nbonds::Int32; bond::Ptr{Int32}; bilin::Ptr{Float64}; biquad::Ptr{Float64}
    has_biquad::Ptr{Int32}; scalar::Ptr{Float64}
    ewald::Int32; mu0_muB2::Float64; demag::NTuple{9,Float64}; latvecs::NTuple{9,Float64}
    positions::Ptr{Float64}; nreplicas::Int32; dims::Ptr{Int32}
    layout::Int32   # 0 = SNH_LAYOUT_AUTO
    gen_start::Ptr{Int32}; gen_A::Ptr{Float64}; gen_B::Ptr{Float64}   # SU(N) general couplings
end

mutable struct DeviceSystem{N}
    sys::System{N}
    handle::Ptr{Cvoid}
    step::UInt64
    seed::UInt64
end

check(rc, h) = rc == 0 || (msg = unsafe_string(ccall((:snh_last_error, lib), Cstring, (Ptr{Cvoid},), h));
                            rc == 3 ? error(msg) : error("libsunnyhip: " * msg))

# Coupling tables in the layout of include/sunny_hip.h, built from the
# homogeneous interactions of `sys` (src/System/Types.jl:46-74, 106-140).
function coupling_tables(sys::System{N}) where N
    ints = sys.interactions_union
    na = length(ints)
    bonds = Int32[]; bilin = Float64[]; biquad = Float64[]; hasbq = Int32[]; scalar = Float64[]
    gen_start = Int32[0]; gen_A = Float64[]; gen_B = Float64[]
    rowmajor_c(M) = vec(permutedims(cat(real(permutedims(M)), imag(permutedims(M)); dims=3), (3, 1, 2)))
    for i in 1:na, pc in ints[i].pair
        (; bond, isculled) = pc
        append!(bonds, Int32[bond.i - 1, bond.j - 1, bond.n..., isculled])
        J = pc.bilin isa Number ? pc.bilin * Sunny.Mat3(Sunny.I) : pc.bilin
        append!(bilin, vec(permutedims(J)))                 # row-major
        B = pc.biquad isa Number ? pc.biquad * Sunny.diagm(Sunny.scalar_biquad_metric) : pc.biquad
        append!(biquad, vec(permutedims(Matrix(B))))
        push!(hasbq, iszero(pc.biquad) ? 0 : 1)
        push!(scalar, pc.scalar)
        for (A, B) in pc.general.data   # TensorDecomposition terms (Types.jl:23-27)
            append!(gen_A, rowmajor_c(Matrix(A))); append!(gen_B, rowmajor_c(Matrix(B)))
        end
        push!(gen_start, gen_start[end] + length(pc.general.data))
    end
    kappa = Float64[sys.κs[1, 1, 1, i] for i in 1:na]
    g = reduce(vcat, [vec(permutedims(sys.gs[1, 1, 1, i])) for i in 1:na])
    field = reduce(vcat, [Vector(sys.extfield[1, 1, 1, i]) for i in 1:na])
    stevens = zeros(28 * na); kmax = zeros(Int32, na)
    Lambda = zeros(2 * max(N, 1)^2 * na)
    for i in 1:na
        os = ints[i].onsite
        if N == 0
            stevens[28(i-1)+1:28i] = vcat(os.c0, os.c2, os.c4, os.c6); kmax[i] = os.kmax
        else
            M = Matrix(os)
            Lambda[2N^2*(i-1)+1:2N^2*i] = rowmajor_c(M)
        end
    end
    dims = Int32[sys.dims...]
    pos = reduce(vcat, sys.crystal.positions)
    return (; na, bonds, bilin, biquad, hasbq, scalar, kappa, g, field, stevens, kmax, Lambda, dims, pos,
            gen_start, gen_A, gen_B)
end

function to_device(sys::System{N}; precision=64, seed=rand(UInt64), autotune=false) where N
    t = coupling_tables(sys)
    ew = sys.ewald
    GC.@preserve t begin
        spec = SnhSpec(t.na, N, precision, pointer(t.kappa), pointer(t.g), pointer(t.field), pointer(t.stevens),
                       pointer(t.kmax), pointer(t.Lambda), length(t.scalar), pointer(t.bonds), pointer(t.bilin),
                       pointer(t.biquad), pointer(t.hasbq), pointer(t.scalar), isnothing(ew) ? 0 : 1,
                       isnothing(ew) ? 0.0 : ew.μ0_μB², isnothing(ew) ? ntuple(_ -> 0.0, 9) : Tuple(vec(permutedims(ew.demag))),
                       Tuple(vec(permutedims(sys.crystal.latvecs))), pointer(t.pos), 1, pointer(t.dims), 0,
                       isempty(t.gen_A) ? C_NULL : pointer(t.gen_start), pointer(t.gen_A), pointer(t.gen_B))
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:snh_create, lib), Cint, (Ref{SnhSpec}, Ref{Ptr{Cvoid}}), spec, h), C_NULL)
    end
    d = DeviceSystem{N}(sys, h[], 0, seed)
    finalizer(x -> ccall((:snh_destroy, lib), Cint, (Ptr{Cvoid},), x.handle), d)
    upload!(d)
    # measured choice of the tile plans (no reference counterpart; results
    # do not depend on it)
    autotune && check(ccall((:snh_autotune, lib), Cint, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}),
                            d.handle, 8, 3, C_NULL), d.handle)
    return d
end

# Zero-copy: Array{Vec3,4} is 3 contiguous Float64 per site in column-major
# (L1, L2, L3, na) order, exactly the C-ABI layout.
function upload!(d::DeviceSystem{N}) where N
    if N == 0
        check(ccall((:snh_set_dipoles, lib), Cint, (Ptr{Cvoid}, Ptr{Float64}), d.handle, d.sys.dipoles), d.handle)
    else
        check(ccall((:snh_set_coherents, lib), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}), d.handle, d.sys.coherents), d.handle)
    end
end

function download!(d::DeviceSystem{N}) where N
    check(ccall((:snh_get_dipoles, lib), Cint, (Ptr{Cvoid}, Ptr{Float64}), d.handle, d.sys.dipoles), d.handle)
    N > 0 && check(ccall((:snh_get_coherents, lib), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}), d.handle, d.sys.coherents), d.handle)
    return d.sys
end

function Sunny.step!(d::DeviceSystem, integrator::Langevin)
    Sunny.check_timestep_available(integrator)
    kT = Ref(integrator.kT)
    check(ccall((:snh_step_langevin, lib), Cint,
                (Ptr{Cvoid}, Float64, Float64, Ptr{Float64}, Ptr{Float64}, UInt64, UInt64, Int32),
                d.handle, integrator.dt, integrator.damping, kT, C_NULL, d.seed, d.step, 1), d.handle)
    d.step += 1
    return
end

function Sunny.step!(d::DeviceSystem, integrator::ImplicitMidpoint; max_iters=100)
    Sunny.check_timestep_available(integrator)
    iters = Ref{Int32}(0)
    check(ccall((:snh_step_midpoint, lib), Cint,
                (Ptr{Cvoid}, Float64, Float64, Float64, Int32, Int32, Ptr{Int32}),
                d.handle, integrator.dt, integrator.damping, integrator.tol, max_iters, 1, iters), d.handle)
    return
end

# ---------------------------------------------------------------------------
# Reverse rule for a deterministic ImplicitMidpoint trajectory (no such rule
# exists in Sunny; src/Fitting.jl:123-152 is the pattern).  `trajectory`
# runs nsteps of step! and returns the final dipoles; its pullback maps the
# cotangent of the final dipoles to cotangents of the labeled parameters
# (as a SystemTangent over `labels`, src/System/Types.jl:143-146) and of the
# initial dipoles.

using ChainRulesCore

function trajectory(d::DeviceSystem{0}, integrator::ImplicitMidpoint, nsteps::Int; checkpoint_every=50, max_iters=100)
    iszero(integrator.damping) || error("the midpoint adjoint needs damping = 0")
    check(ccall((:snh_midpoint_record, lib), Cint, (Ptr{Cvoid}, Float64, Float64, Int32, Int32, Int32),
                d.handle, integrator.dt, integrator.tol, max_iters, nsteps, checkpoint_every), d.handle)
    return copy(download!(d).dipoles)
end

# Unit couplings of parameter `label`: the system with only that parameter set to 1.
function param_tables(sys::System{0}, label::Symbol)
    unit = Sunny.clone_system(sys)
    for p in unit.params
        p.val = (p.label == label) ? 1.0 : 0.0
    end
    Sunny.repopulate_couplings_from_params!(unit)
    return coupling_tables(unit)
end

function ChainRulesCore.rrule(::typeof(trajectory), d::DeviceSystem{0}, integrator::ImplicitMidpoint, nsteps::Int;
                              labels::Vector{Symbol}, kws...)
    S_final = trajectory(d, integrator, nsteps; kws...)
    function trajectory_pullback(dS)
        dLdS = collect(reinterpret(Float64, unthunk(dS)))
        tabs = [param_tables(d.sys, l) for l in labels]
        # snh_spec per parameter on the system's bond list (layout of to_device)
        specs = map(tabs) do t
            SnhSpec(t.na, 0, 64, pointer(t.kappa), pointer(t.g), pointer(t.field), pointer(t.stevens), pointer(t.kmax),
                    pointer(t.Lambda), length(t.scalar), pointer(t.bonds), pointer(t.bilin), pointer(t.biquad), pointer(t.hasbq),
                    pointer(t.scalar), 0, 0.0, ntuple(_ -> 0.0, 9), ntuple(_ -> 0.0, 9), pointer(t.pos), 1,
                    pointer(t.dims), 0, C_NULL, C_NULL, C_NULL)
        end
        grad = zeros(length(labels))
        dS0 = zeros(length(dLdS))
        GC.@preserve tabs dLdS begin
            check(ccall((:snh_midpoint_adjoint, lib), Cint,
                        (Ptr{Cvoid}, Int32, Ptr{SnhSpec}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                        d.handle, length(labels), specs, dLdS, grad, dS0), d.handle)
        end
        tangent = Sunny.SystemTangent(Dict(zip(labels, grad)))
        return NoTangent(), tangent, NoTangent(), NoTangent()
    end
    return S_final, trajectory_pullback
end

function Sunny.energy(d::DeviceSystem)
    E = Ref(0.0)
    check(ccall((:snh_energy, lib), Cint, (Ptr{Cvoid}, Ptr{Float64}), d.handle, E), d.handle)
    return E[]
end

# SampledCorrelations accumulated on the device (:dipole or :SUN): the measure's
# observables and pairs, dt and measperiod as Sunny computes them
# (adjusted_dt_and_downsampling_factor), then add_sample! and sc.data.
struct DeviceCorrelations
    d::DeviceSystem
    dt::Float64
    measperiod::Int
    dims::NTuple{4, Int}   # (ncorr, na, na) sizes and n_all_omega follow sc.data
    n_all_omega::Int
    ncorr::Int
end

function device_correlations(d::DeviceSystem, measure; energies, dt, window=:cosine)
    nω = length(energies)
    dt, measperiod = Sunny.adjusted_dt_and_downsampling_factor(dt, nω, energies[end])
    na = Sunny.natoms(d.sys.crystal)
    obs = zeros(3, na, 3)                 # observable a of atom i: obs[:, i, a] in Julia order
    for i in 1:na, a in 1:length(measure.observables[:, 1, 1, 1, i])
        obs[:, i, a] = measure.observables[a, 1, 1, 1, i]
    end
    pairs = Int32[p[k] - 1 for p in measure.corr_pairs for k in 1:2]
    n_all = 2(nω - 1)
    check(ccall((:snh_corr_init, lib), Cint,
                (Ptr{Cvoid}, Int32, Ptr{Float64}, Int32, Ptr{Int32}, Int32, Int32, Int32),
                d.handle, size(obs, 3), obs, length(measure.corr_pairs), pairs, n_all, measperiod,
                window == :cosine ? 1 : 0), d.handle)
    return DeviceCorrelations(d, dt, measperiod, (length(measure.corr_pairs), na, na, 0), n_all,
                              length(measure.corr_pairs))
end

function Sunny.add_sample!(c::DeviceCorrelations, ::DeviceSystem; tol=1e-12, max_iters=100)
    check(ccall((:snh_corr_add_sample, lib), Cint, (Ptr{Cvoid}, Float64, Float64, Int32),
                c.d.handle, c.dt, tol, max_iters), c.d.handle)
end

function correlation_data(c::DeviceCorrelations)
    na = c.dims[2]
    data = zeros(ComplexF64, c.ncorr, na, na, c.d.sys.dims..., c.n_all_omega)
    check(ccall((:snh_corr_data, lib), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{Int64}),
                c.d.handle, data, C_NULL), c.d.handle)
    return data   # the layout of sc.data: zero-copy
end

# LocalSampler (src/MonteCarlo/Samplers.jl:133-153) on the device, :dipole
# or :SUN, without Ewald.  The proposal function is translated to the device's
# mixture: propose_uniform, propose_flip, the closure of propose_delta (its
# captured `magnitude`) and the closure of @mix_proposals (its `cumprobs`
# and `fns`).
function device_proposal(f)
    f === Sunny.propose_uniform && return [(Int32(0), 0.0, 1.0)]
    f === Sunny.propose_flip && return [(Int32(1), 0.0, 1.0)]
    hasfield(typeof(f), :magnitude) && return [(Int32(2), Float64(f.magnitude), 1.0)]
    if hasfield(typeof(f), :fns) && hasfield(typeof(f), :cumprobs)
        probs = diff([0.0; collect(f.cumprobs)])
        return [(k, m, p * w) for (g, p) in zip(f.fns, probs) for (k, m, w) in device_proposal(g)]
    end
    error("the device LocalSampler supports propose_uniform, propose_flip, propose_delta and @mix_proposals of them")
end

function Sunny.step!(d::DeviceSystem, sampler::Sunny.LocalSampler)
    terms = device_proposal(sampler.propose)
    kinds = Int32[t[1] for t in terms]
    mags = Float64[t[2] for t in terms]
    wts = Float64[t[3] for t in terms]
    kT = Ref(sampler.kT)
    out = zeros(4)
    check(ccall((:snh_step_local, lib), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Int32, Ptr{Int32}, Ptr{Float64}, Ptr{Float64}, UInt64, UInt64, Float64,
                 Ptr{Float64}),
                d.handle, kT, length(kinds), kinds, mags, wts, d.seed, d.step, sampler.nsweeps, out), d.handle)
    d.step += ceil(UInt64, sampler.nsweeps)
    sampler.ΔE += out[1]
    sampler.ΔS += Sunny.Vec3(out[2], out[3], out[4])
    return
end

end
