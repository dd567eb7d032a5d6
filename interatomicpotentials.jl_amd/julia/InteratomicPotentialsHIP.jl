"""
    InteratomicPotentialsHIP

Julia front end of libipgpu.so: the GPU (MI355X, gfx950) implementation of
InteratomicPotentials.jl's neighbor list, empirical pair potentials and SNAP
descriptors/potentials, reached with `ccall` through the C ABI declared in
include/ipgpu.h.  Methods are added for the wrapper type `HIP{P}` so that
the reference's own types and generic functions can be used without type
piracy:

    using InteratomicPotentials, InteratomicPotentialsHIP
    e, f = energy_and_force(system, HIP(LennardJones(ε, σ, rcut, [:Ar])))
    B    = compute_local_descriptors(system, HIP(snap))

Units follow the reference: empirical inputs are `austrip`ped (hartree,
bohr), SNAP is unit-naive (`ustrip`).  Atom indices returned by the library
are 0-based; this module adds 1.
"""
module InteratomicPotentialsHIP

using AtomsBase
using InteratomicPotentials
using StaticArrays
using Unitful
using UnitfulAtomic
import ChainRulesCore

const libipgpu = get(ENV, "IPGPU_LIB", joinpath(@__DIR__, "..", "lib", "libipgpu.so"))

const IPG_LIST_HALF = Cint(0)
const IPG_FP64 = Cint(0)
const IPG_FP32 = Cint(1)
const IPG_SNAP_LIST_REFERENCE = Cint(0)
const IPG_SNAP_PLAIN_FP64 = Cint(0)
const IPG_SNAP_ADJOINT_FP64 = Cint(1)
const IPG_SNAP_ADJOINT_FP32 = Cint(2)
const IPG_ERR_CAPACITY = Cint(-2)

struct IPGSystem
    natoms::Int64
    positions::Ptr{Float64}
    species::Ptr{Int32}
    box::NTuple{3,Float64}
    periodic::NTuple{3,Int32}
end

struct IPGPairTerm
    kind::Int32
    reserved::Int32
    coeff::Float64
    rcut::Float64
    p::NTuple{6,Float64}
    species_mask::UInt64
end

struct IPGSnapParams
    twojmax::Int32
    nspecies::Int32
    rcutfac::Float64
    rmin0::Float64
    rfac0::Float64
    radii::Ptr{Float64}
    weight::Ptr{Float64}
    chem_flag::Int32
    bzero_flag::Int32
    bnorm_flag::Int32
    switch_flag::Int32
    wselfall_flag::Int32
end

"Wrapper selecting the GPU implementation of a reference potential or basis."
struct HIP{P}
    p::P
end

check(rc) = rc == 0 ? nothing :
    error(unsafe_string(ccall((:ipg_last_error, libipgpu), Cstring, ())) * " (status $rc)")

const CTX = Ref{Ptr{Cvoid}}(C_NULL)
function context()
    if CTX[] == C_NULL
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:ipg_ctx_create, libipgpu), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, h))
        CTX[] = h[]
    end
    CTX[]
end

# -- systems ---------------------------------------------------------------
function box_and_periodicity(s, strip)
    per = periodicity(s)
    L = ntuple(i -> per[i] ? strip(bounding_box(s)[i][i]) : 0.0, 3)
    L, ntuple(i -> Int32(per[i]), 3)
end

positions_matrix(s, strip) = [strip(position(s, i)[k]) for k in 1:3, i in 1:length(s)]

function with_system(f, s, strip, species::Vector{Int32})
    X = positions_matrix(s, strip)          # 3 x N, contiguous AoS
    L, per = box_and_periodicity(s, strip)
    GC.@preserve X species begin
        sys = IPGSystem(length(s), pointer(X), isempty(species) ? Ptr{Int32}(C_NULL) : pointer(species), L, per)
        f(Ref(sys), X)
    end
end

# -- neighbor list (src/nnlist.jl:23-46) -----------------------------------
function InteratomicPotentials.neighborlist(s::AbstractSystem{3}, rcutoff::Float64, ::Type{HIP})
    with_system(s, austrip, Int32[]) do sys, X
        n = length(s)
        off = zeros(Int64, n + 1)
        cap = Int64(64n)
        while true
            nbr = zeros(Int32, cap)
            disp = zeros(Float64, 3, cap)
            np = Ref{Int64}(0)
            rc = ccall((:ipg_neighborlist, libipgpu), Cint,
                       (Ptr{Cvoid}, Ref{IPGSystem}, Cdouble, Cint, Ptr{Int64}, Ptr{Int32}, Ptr{Float64}, Int64,
                        Ref{Int64}), context(), sys, rcutoff, IPG_LIST_HALF, off, nbr, disp, cap, np)
            if rc == IPG_ERR_CAPACITY
                cap = np[]
                continue
            end
            check(rc)
            nl = NeighborList(n)
            for i in 1:n
                r = (off[i] + 1):off[i+1]
                nl.j[i] = Int.(nbr[r]) .+ 1
                nl.r[i] = [SVector{3}(disp[:, q]) for q in r]
            end
            return nl
        end
    end
end

# -- empirical potentials --------------------------------------------------
pair_kind(::LennardJones) = 0
pair_kind(::Morse) = 1
pair_kind(::ZBL) = 2
pair_kind(::BornMayer) = 3
pair_kind(::Coulomb) = 4
pair_params(p::LennardJones) = (p.ϵ, p.σ)
pair_params(p::Morse) = (p.D, p.α, p.σ)
pair_params(p::ZBL) = (Float64(p.Z₁), Float64(p.Z₂), p.e)
pair_params(p::BornMayer) = (p.A, p.ρ, p.σ, p.C, p.D)
pair_params(p::Coulomb) = (p.q₁, p.q₂)

function term(p::EmpiricalPotential, c, table)
    sp = get_species(p)
    mask = ismissing(sp) ? typemax(UInt64) :
           reduce(|, (UInt64(1) << (k - 1) for (k, s) in enumerate(table) if s in sp); init=UInt64(0))
    q = collect(pair_params(p))
    IPGPairTerm(pair_kind(p), 0, Float64(c), get_rcutoff(p), ntuple(k -> k <= length(q) ? q[k] : 0.0, 6), mask)
end

flat_terms(p::EmpiricalPotential, c=1.0) = [(p, c)]
flat_terms(p::LinearCombinationPotential, c=1.0) =
    reduce(vcat, (flat_terms(q, c * a) for (q, a) in zip(p.potentials, p.coefficients)))

function pair_eval(s::AbstractSystem, p; precision=IPG_FP64)
    table = unique(atomic_symbol(s))
    species = Int32[findfirst(==(x), table) - 1 for x in atomic_symbol(s)]
    terms = [term(q, c, table) for (q, c) in flat_terms(p)]
    with_system(s, austrip, species) do sys, X
        e = Ref{Float64}(0.0)
        f = zeros(Float64, 3, length(s))
        v = zeros(Float64, 6)
        check(ccall((:ipg_pair_energy_forces, libipgpu), Cint,
                    (Ptr{Cvoid}, Ref{IPGSystem}, Ptr{IPGPairTerm}, Cint, Cint, Ref{Float64}, Ptr{Float64},
                     Ptr{Float64}), context(), sys, terms, length(terms), precision, e, f, v))
        e[], f, v
    end
end

function InteratomicPotentials.energy_and_force(s::AbstractSystem, h::HIP{<:Union{EmpiricalPotential,LinearCombinationPotential}})
    e, f, _ = pair_eval(s, h.p)
    (; e=e * u"hartree", f=[SVector{3}(f[:, i]) for i in 1:length(s)] * u"hartree/bohr")
end
InteratomicPotentials.potential_energy(s::AbstractSystem, h::HIP) = energy_and_force(s, h).e
InteratomicPotentials.force(s::AbstractSystem, h::HIP) = energy_and_force(s, h).f
function InteratomicPotentials.virial_stress(s::AbstractSystem, h::HIP{<:Union{EmpiricalPotential,LinearCombinationPotential}})
    _, _, v = pair_eval(s, h.p)
    SVector{6}(v) * u"hartree"
end
# abstract_potential.jl:10: the trace of the virial stress
InteratomicPotentials.virial(s::AbstractSystem, h::HIP{<:Union{EmpiricalPotential,LinearCombinationPotential}}) =
    sum(virial_stress(s, h)[1:3])

# -- SNAP -------------------------------------------------------------------
mutable struct SnapHandle
    ptr::Ptr{Cvoid}
    ncoeff::Int
    radii::Vector{Float64}
    weight::Vector{Float64}
end

const SNAP_HANDLES = IdDict{Any,SnapHandle}()

function handle(snap::SNAP)
    get!(SNAP_HANDLES, snap) do
        radii = Vector{Float64}(snap.radii)
        weight = Vector{Float64}(snap.weight)
        prm = IPGSnapParams(snap.twojmax, length(snap.species), snap.rcutfac, snap.rmin0, snap.rfac0,
                            pointer(radii), pointer(weight), snap.chem_flag, snap.bzero_flag, snap.bnorm_flag,
                            snap.switch_flag, snap.wselfall_flag)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        GC.@preserve radii weight check(ccall((:ipg_snap_create, libipgpu), Cint,
                                              (Ptr{Cvoid}, Ref{IPGSnapParams}, Ref{Ptr{Cvoid}}), context(), prm, h))
        nc = Ref{Int32}(0)
        check(ccall((:ipg_snap_info, libipgpu), Cint, (Ptr{Cvoid}, Ref{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}),
                    h[], nc, C_NULL, C_NULL, C_NULL))
        SnapHandle(h[], nc[], radii, weight)
    end
end

snap_species(s, snap) = Int32[findfirst(==(x), snap.species) - 1 for x in atomic_symbol(s)]
function snap_appear(s)
    els = unique(atomic_symbol(s))
    Int32[findfirst(==(x), els) - 1 for x in atomic_symbol(s)]
end

function InteratomicPotentials.compute_local_descriptors(s::AbstractSystem, h::HIP{<:SNAP}; mode=IPG_SNAP_LIST_REFERENCE)
    snap = h.p
    H = handle(snap)
    appear = snap_appear(s)
    with_system(s, ustrip, snap_species(s, snap)) do sys, X
        B = zeros(Float64, H.ncoeff, length(s))
        check(ccall((:ipg_snap_local_descriptors, libipgpu), Cint,
                    (Ptr{Cvoid}, Ref{IPGSystem}, Cint, Ptr{Int32}, Ptr{Float64}), H.ptr, sys, mode, appear, B))
        [B[:, i] for i in 1:length(s)]
    end
end

function InteratomicPotentials.compute_force_descriptors(s::AbstractSystem, h::HIP{<:SNAP}; mode=IPG_SNAP_LIST_REFERENCE)
    snap = h.p
    H = handle(snap)
    M = H.ncoeff * length(snap.species)
    appear = snap_appear(s)
    with_system(s, ustrip, snap_species(s, snap)) do sys, X
        dB = zeros(Float64, 3, M, length(s))   # per atom a 3 x M column-major matrix
        check(ccall((:ipg_snap_force_descriptors, libipgpu), Cint,
                    (Ptr{Cvoid}, Ref{IPGSystem}, Cint, Ptr{Int32}, Ptr{Float64}), H.ptr, sys, mode, appear, dB))
        [dB[:, :, i] for i in 1:length(s)]
    end
end

# snap.jl:226-285: (B, dB, W) with W the 6 x M virial rows summed over atoms.
function InteratomicPotentials.compute_all_descriptors(s::AbstractSystem, h::HIP{<:SNAP}; mode=IPG_SNAP_LIST_REFERENCE)
    snap = h.p
    H = handle(snap)
    M = H.ncoeff * length(snap.species)
    appear = snap_appear(s)
    with_system(s, ustrip, snap_species(s, snap)) do sys, X
        B = zeros(Float64, H.ncoeff, length(s))
        dB = zeros(Float64, 3, M, length(s))
        W = zeros(Float64, 6, M)            # column c holds the library's row c of M x 6
        check(ccall((:ipg_snap_all_descriptors, libipgpu), Cint,
                    (Ptr{Cvoid}, Ref{IPGSystem}, Cint, Ptr{Int32}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                    H.ptr, sys, mode, appear, B, dB, W))
        [B[:, i] for i in 1:length(s)], [dB[:, :, i] for i in 1:length(s)], W
    end
end

function snap_energy_forces(s, lbp, path; mode=IPG_SNAP_LIST_REFERENCE)
    snap = lbp.basis
    H = handle(snap)
    appear = snap_appear(s)
    β = Vector{Float64}(lbp.β)
    with_system(s, ustrip, snap_species(s, snap)) do sys, X
        e = Ref{Float64}(0.0)
        f = zeros(Float64, 3, length(s))
        check(ccall((:ipg_snap_energy_forces, libipgpu), Cint,
                    (Ptr{Cvoid}, Ref{IPGSystem}, Cint, Ptr{Int32}, Ptr{Float64}, Int64, Cdouble, Cint,
                     Ref{Float64}, Ptr{Float64}), H.ptr, sys, mode, appear, β, length(β), lbp.β0[1], path, e, f))
        e[], [SVector{3}(f[:, i]) for i in 1:length(s)]
    end
end

# chem_flag = true and twojmax > 8 have only the reference's β . dB route on the GPU
default_path(snap) = (snap.chem_flag || snap.twojmax > 8) ? IPG_SNAP_PLAIN_FP64 : IPG_SNAP_ADJOINT_FP64

function InteratomicPotentials.energy_and_force(s::AbstractSystem, h::HIP{<:LBasisPotential};
                                                path=default_path(h.p.basis))
    e, f = snap_energy_forces(s, h.p, path)
    (; e, f)
end

# basis_potentials.jl:43-50 with linear_potential.jl:37-42: SVector{6}(W * β),
# W the virial descriptors of compute_all_descriptors (snap.jl:267-281).
function InteratomicPotentials.virial_stress(s::AbstractSystem, h::HIP{<:LBasisPotential}; path=default_path(h.p.basis),
                                             mode=IPG_SNAP_LIST_REFERENCE)
    lbp = h.p
    snap = lbp.basis
    H = handle(snap)
    appear = snap_appear(s)
    β = Vector{Float64}(lbp.β)
    with_system(s, ustrip, snap_species(s, snap)) do sys, X
        v = zeros(Float64, 6)
        check(ccall((:ipg_snap_virial_stress, libipgpu), Cint,
                    (Ptr{Cvoid}, Ref{IPGSystem}, Cint, Ptr{Int32}, Ptr{Float64}, Int64, Cint, Ptr{Float64}),
                    H.ptr, sys, mode, appear, β, length(β), path, v))
        SVector{6}(v)
    end
end
# basis_potentials.jl:52-57 and linear_potential.jl:44-49: the sum of all six entries
InteratomicPotentials.virial(s::AbstractSystem, h::HIP{<:LBasisPotential}; kw...) = sum(virial_stress(s, h; kw...))

"""
    rrule(force, s, h::HIP{<:LBasisPotential})

Reverse rule of the SNAP forces for force-matching fits: the pullback of a
force cotangent returns the gradients with respect to β and to the
positions (exact forward-mode tangents on the GPU, ipg_snap_force_vjp).
"""
function ChainRulesCore.rrule(::typeof(InteratomicPotentials.force), s::AbstractSystem, h::HIP{<:LBasisPotential};
                              precision=IPG_FP64, mode=IPG_SNAP_LIST_REFERENCE)
    lbp = h.p
    _, F = snap_energy_forces(s, lbp, IPG_SNAP_ADJOINT_FP64; mode)
    function force_pullback(ΔF)
        snap = lbp.basis
        H = handle(snap)
        β = Vector{Float64}(lbp.β)
        g = reduce(hcat, [Vector{Float64}(ChainRulesCore.unthunk(x)) for x in ΔF])
        gβ = zeros(Float64, length(β))
        gx = zeros(Float64, 3, length(s))
        with_system(s, ustrip, snap_species(s, snap)) do sys, X
            check(ccall((:ipg_snap_force_vjp, libipgpu), Cint,
                        (Ptr{Cvoid}, Ref{IPGSystem}, Cint, Ptr{Int32}, Ptr{Float64}, Int64, Ptr{Float64}, Cint,
                         Ptr{Float64}, Ptr{Float64}), H.ptr, sys, mode, snap_appear(s), β, length(β), g, precision,
                        gβ, gx))
        end
        dh = ChainRulesCore.Tangent{typeof(h)}(p=ChainRulesCore.Tangent{typeof(lbp)}(β=gβ))
        ds = [SVector{3}(gx[:, i]) for i in 1:length(s)]
        return ChainRulesCore.NoTangent(), ds, dh
    end
    return F, force_pullback
end

"""
    compute_batch_descriptors(systems, HIP(snap); precision=IPG_FP64)

Descriptors of a data set in one GPU call (CUTOFF list mode), for fitting:
per configuration `B` (`ncoeff x N_c`), `dB` (`3 x M x N_c`), the energy row
(sum of B in species blocks, length `M`) and `W` (`6 x M`), as
`compute_all_descriptors` returns them per configuration.
"""
function compute_batch_descriptors(systems::AbstractVector, h::HIP{<:SNAP}; precision=IPG_FP64)
    snap = h.p
    H = handle(snap)
    M = H.ncoeff * length(snap.species)
    nc = length(systems)
    Xs = [positions_matrix(s, ustrip) for s in systems]
    sps = [snap_species(s, snap) for s in systems]
    sys = Vector{IPGSystem}(undef, nc)
    n = sum(length, systems)
    B = zeros(Float64, H.ncoeff, n)
    dB = zeros(Float64, 3, M, n)
    erow = zeros(Float64, M, nc)
    W = zeros(Float64, 6, M, nc)
    GC.@preserve Xs sps begin
        for c in 1:nc
            L, per = box_and_periodicity(systems[c], ustrip)
            sys[c] = IPGSystem(length(systems[c]), pointer(Xs[c]), pointer(sps[c]), L, per)
        end
        check(ccall((:ipg_snap_batch_descriptors, libipgpu), Cint,
                    (Ptr{Cvoid}, Int64, Ptr{IPGSystem}, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                    H.ptr, nc, sys, precision, B, dB, erow, W))
    end
    starts = cumsum([0; length.(systems)])
    [(B=B[:, starts[c]+1:starts[c+1]], dB=dB[:, :, starts[c]+1:starts[c+1]], e=erow[:, c], W=W[:, :, c])
     for c in 1:nc]
end

export HIP, compute_batch_descriptors

end # module
