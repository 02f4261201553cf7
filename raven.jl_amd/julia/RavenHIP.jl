# RavenHIP.jl -- Julia binding of libravenhip (include/ravenhip.h) for Raven.jl.
#
# Drop-in for Raven's DGSEM right-hand side (src/balancelaws/dgsem.jl:84-126)
# and LSRK time stepping (src/balancelaws/odesolvers.jl:9-61) on one MI355X.
# Grids, LobattoCell operators, GridArray state and balance laws stay Raven's;
# this module only hands their arrays to the library through ccall.
#
# Status: written against the C ABI; Julia is not available where this
# library is built and tested, so this file is not exercised by the test
# suite (the Python mirror in raven.jl_amd/ravenhip is).
module RavenHIP

using Raven
using Raven.BalanceLaws
using Raven.BalanceLaws.Euler: EulerLaw
using Raven.BalanceLaws.Advection: AdvectionLaw

const libravenhip = joinpath(@__DIR__, "..", "ravenhip", "libravenhip.so")

const RV_LAW_ADVECTION = Int32(1)
const RV_LAW_EULER = Int32(2)
const RV_FLUX_RUSANOV = Int32(1)
const RV_FLUX_CENTRAL = Int32(2)
const RV_FLUX_EC = Int32(3)
const RV_FORM_WEAK = Int32(0)
const RV_FORM_FLUX_DIFFERENCING = Int32(1)
const RV_F64 = Int32(1)
const RV_F32 = Int32(2)

# Field order and types mirror `rv_desc` in include/ravenhip.h.
struct RvDesc
    abi_version::Int32
    dim::Int32
    nq::Int32
    law::Int32
    flux::Int32
    dtype::Int32
    geometry::Int32
    directions::Int32
    ncells::Int64
    gamma::Float64
    velocity::NTuple{3,Float64}
    D::Ptr{Float64}
    weights::Ptr{Float64}
    metrics::Ptr{Float64}
    surface_metrics::Ptr{Float64}
    vmapM::Ptr{Int64}
    vmapP::Ptr{Int64}
    boundary_codes::Ptr{Int64}
    dtoc::Ptr{Int64}
    ncontinuous::Int64
    stream::Ptr{Cvoid}
    volume_form::Int32
    volume_flux::Int32
end

mutable struct HIPDGSEM{L}
    law::L
    handle::Ptr{Cvoid}
    function HIPDGSEM(law, handle)
        dg = new{typeof(law)}(law, handle)
        finalizer(d -> ccall((:rv_destroy, libravenhip), Cvoid, (Ptr{Cvoid},), d.handle), dg)
        return dg
    end
end

check(rc, h) = rc == 0 || error("libravenhip: ", unsafe_string(
    ccall((:rv_last_error, libravenhip), Cstring, (Ptr{Cvoid},), h)))

lawid(::EulerLaw) = RV_LAW_EULER
lawid(::AdvectionLaw) = RV_LAW_ADVECTION
fluxid(::RusanovFlux) = RV_FLUX_RUSANOV
fluxid(::CentralFlux) = RV_FLUX_CENTRAL
fluxid(::EntropyConservativeFlux) = RV_FLUX_EC
formid(::WeakForm) = (RV_FORM_WEAK, Int32(0))
formid(f::FluxDifferencingForm) = (RV_FORM_FLUX_DIFFERENCING, fluxid(f.volume_numericalflux))

"""
    HIPDGSEM(; law, grid, surface_numericalflux, volume_form = WeakForm(),
             dtype = Float64, stream = C_NULL)

Same keywords as `DGSEM` (dgsem.jl:46-81).  Host copies of the grid arrays
are read once (`Array(...)` of device arrays is fine); the library keeps
device copies and per-face neighbour records.
"""
function HIPDGSEM(; law, grid, surface_numericalflux, volume_form = WeakForm(),
                  directions = 1:ndims(law), dtype = Float64, stream = C_NULL)
    cell = referencecell(grid)
    dim = ndims(cell)
    Nq = size(cell, 1)
    D = Array{Float64}(derivatives_1d(cell)[1])
    w = Array{Float64}(vec(weights_1d(cell)[1]))
    metrics = Array{Float64}(parent(first(volumemetrics(grid))))
    surf = Array{Float64}(parent(first(surfacemetrics(grid))))
    fm = facemaps(grid)
    vmapM = Array{Int64}(fm.vmapM)
    vmapP = Array{Int64}(fm.vmapP)
    bc = Array{Int64}(boundarycodes(grid))
    # not exported by Raven (grids.jl:68, 86)
    dtoc = Array{Int64}(Raven.discontinuoustocontinuous(grid))
    u = law isa AdvectionLaw ? Tuple(Float64.(constants(law).u⃗)) : ()
    vel = ntuple(i -> i <= length(u) ? u[i] : 0.0, 3)
    γ = law isa EulerLaw ? Float64(constants(law).γ) : 1.4
    handle = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve D w metrics surf vmapM vmapP bc dtoc begin
        form, vflux = formid(volume_form)
        # DGSEM's `directions` (dgsem.jl:46-81) as the descriptor's bit mask
        dirmask = Int32(sum(1 << (d - 1) for d in directions))
        desc = RvDesc(2, dim, Nq, lawid(law), fluxid(surface_numericalflux),
                      dtype == Float64 ? RV_F64 : RV_F32, 0, dirmask, length(grid), γ, vel,
                      pointer(D), pointer(w), pointer(metrics), pointer(surf),
                      pointer(vmapM), pointer(vmapP), pointer(bc), pointer(dtoc),
                      Raven.numcnodes(grid), stream, form, vflux)
        check(ccall((:rv_create, libravenhip), Cint, (Ref{RvDesc}, Ref{Ptr{Cvoid}}),
                    desc, handle), C_NULL)
    end
    return HIPDGSEM(law, handle[])
end

# dg(dq, q, t; increment) (dgsem.jl:84): q and dq are device GridArrays whose
# parent memory is HIP device memory in Raven's layout.
function (dg::HIPDGSEM)(dq, q, time; increment = true)
    check(ccall((:rv_rhs, libravenhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cint),
                dg.handle, pointer(parent(dq)), pointer(parent(q)), time, increment), dg.handle)
end

# solve!(q, timeend, LSRK54(dg, q, dt); adjust_final) (odesolvers.jl:9-30) in
# one call.
function lsrk_solve!(dg::HIPDGSEM, q, dq, dt, timeend; t0 = 0.0, scheme = 54,
                     adjust_final = true)
    tout = Ref{Cdouble}(0)
    nsteps = Ref{Int64}(0)
    check(ccall((:rv_lsrk_solve, libravenhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cdouble, Cdouble, Cdouble, Cint,
                 Ref{Cdouble}, Ref{Int64}),
                dg.handle, pointer(parent(q)), pointer(parent(dq)), scheme, t0, dt, timeend,
                adjust_final ? 1 : 0, tout, nsteps), dg.handle)
    return tout[], nsteps[]
end

# Couple the hanging faces of an adapted grid through mortars, from Raven's own
# non-conforming tables facemaps(grid).vmapNC / .nctoface
# (lobattocells.jl:2121-2276): rv_mortars_from_nctables, then rv_enable_mortars.
function enable_mortars!(dg::HIPDGSEM, grid)
    cell = referencecell(grid)
    dim = ndims(cell)
    Nq = size(cell, 1)
    K = Int64(numcells(grid))
    fm = facemaps(grid)
    vmapNC = [Array{Int64}(v) for v in fm.vmapNC]
    nctoface = [Array{Int8}(v) for v in fm.nctoface]
    nunique = Int64[size(v, ndims(v)) for v in vmapNC]
    kind = zeros(Int32, 2dim, K)
    partner = zeros(Int64, 4, 2dim, K)
    GC.@preserve vmapNC nctoface begin
        vptrs = Ptr{Int64}[pointer(v) for v in vmapNC]
        fptrs = Ptr{Int8}[pointer(v) for v in nctoface]
        check(ccall((:rv_mortars_from_nctables, libravenhip), Cint,
                    (Int32, Int32, Int64, Ptr{Int64}, Ptr{Ptr{Int64}}, Ptr{Ptr{Int8}},
                     Ptr{Int32}, Ptr{Int64}),
                    dim, Nq, K, nunique, vptrs, fptrs, kind, partner), C_NULL)
    end
    check(ccall((:rv_enable_mortars, libravenhip), Cint, (Ptr{Cvoid}, Ptr{Int32}, Ptr{Int64}),
                dg.handle, kind, partner), dg.handle)
end

# entropyintegral(dg, q) / entropyproduct(dg, p, q) (measures.jl:23-33)
function entropyintegral(dg::HIPDGSEM, q)
    out = Ref{Cdouble}(0)
    check(ccall((:rv_entropy_integral, libravenhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{Cdouble}),
                dg.handle, pointer(parent(q)), out), dg.handle)
    return out[]
end
function entropyproduct(dg::HIPDGSEM, p, q)
    out = Ref{Cdouble}(0)
    check(ccall((:rv_entropy_product, libravenhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{Cdouble}),
                dg.handle, pointer(parent(p)), pointer(parent(q)), out), dg.handle)
    return out[]
end

# dostep!(q, rlsrk::RLSRK) (odesolvers.jl:117-151) for an RLSRK built on a
# HIPDGSEM: the solver's dq, k, q0 and gamma are used in place, so Raven's
# solve! drives it unchanged through this method.
function rlsrk_dostep!(q, rlsrk, dg::HIPDGSEM; scheme = 54)
    γ = Ref{Cdouble}(rlsrk.γ)
    iters = Ref{Int32}(0)
    check(ccall((:rv_rlsrk_step, libravenhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cdouble,
                 Cdouble, Ref{Cdouble}, Ref{Int32}),
                dg.handle, pointer(parent(q)), pointer(parent(rlsrk.dq)), pointer(parent(rlsrk.k)),
                pointer(parent(rlsrk.q0)), scheme, rlsrk.time, rlsrk.dt, γ, iters), dg.handle)
    rlsrk.γ = γ[]
    rlsrk.time += rlsrk.dt
    return iters[]
end

# Resident device states for a host without AMDGPU.jl (rv_state_*): a state
# of the operator's shape and precision, uploaded from / downloaded to a host
# array in Raven's parent layout.
function state_alloc(dg::HIPDGSEM)
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:rv_state_alloc, libravenhip), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
                dg.handle, p), dg.handle)
    return p[]
end
state_free(dg::HIPDGSEM, p::Ptr{Cvoid}) =
    check(ccall((:rv_state_free, libravenhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), dg.handle, p),
          dg.handle)
state_upload!(dg::HIPDGSEM, p::Ptr{Cvoid}, host::Array) =
    check(ccall((:rv_state_upload, libravenhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                dg.handle, p, host), dg.handle)
state_download!(dg::HIPDGSEM, host::Array, p::Ptr{Cvoid}) =
    check(ccall((:rv_state_download, libravenhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                dg.handle, host, p), dg.handle)

# min_node_distance(grid; dims) (grids.jl:94-145) on the device: `points` is
# the device copy of parent(points(grid)) (Np, dim, K), FP64.
function min_node_distance(points_device::Ptr{Cvoid}, dim, Nq, K; dims = 1:dim,
                           stream = C_NULL)
    out = Ref{Cdouble}(0)
    mask = Int32(sum(1 << (d - 1) for d in dims))
    check(ccall((:rv_min_node_distance, libravenhip), Cint,
                (Cint, Cint, Int64, Ptr{Cvoid}, Int32, Ref{Cdouble}, Ptr{Cvoid}),
                dim, Nq, K, points_device, mask, out, stream), C_NULL)
    return out[]
end

end # module
