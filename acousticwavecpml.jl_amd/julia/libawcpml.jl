# libawcpml.jl -- ccall layer over libawcpml.so (C ABI: include/awcpml.h).
#
# Included by the Acoustic2D_HIP / Acoustic3D_HIP backend modules, the way
# the reference's backends include src/shared/acoustic_*_xPU.jl.  Every
# ccall names its argument types in a literal tuple; Julia arrays are passed
# directly (column-major, x fastest: the layout the library expects) and
# positions stay 1-based, as the reference's possrcs/posrecs are.

# path of the shared library: ENV["AWCPML_LIB"], else the in-tree build
const libawcpml = get(ENV, "AWCPML_LIB",
                      joinpath(@__DIR__, "..", "lib", "libawcpml.so"))

const AWCPML_F64 = Cint(0)
const AWCPML_F32 = Cint(1)

"""
    AWCPMLError(code, msg)

A libawcpml call returned a negative status (`msg` from `awcpml_last_error`).
"""
struct AWCPMLError <: Exception
    code::Cint
    msg::String
end

Base.showerror(io::IO, e::AWCPMLError) = print(io, "libawcpml error ", e.code, ": ", e.msg)

function hipcheck(rc::Cint)
    if rc != 0
        msg = unsafe_string(ccall((:awcpml_last_error, libawcpml), Cstring, ()))
        throw(AWCPMLError(rc, msg))
    end
    return nothing
end

precision_code(p::Symbol) = p === :f64 ? AWCPML_F64 :
                            p === :f32 ? AWCPML_F32 :
                            throw(ArgumentError("precision must be :f64 or :f32 (got $(p))"))

"""
    HIPPlan(dims, halo; precision=:f64, nshots=1)

One velocity model on the GPU with `nshots` independent shots (the device
arrays the reference allocates with `@zeros`, acoustic_2D_xPU.jl:208-215).
Freed by its finalizer or by `destroy!`.
"""
mutable struct HIPPlan
    handle::Ptr{Cvoid}
    ndim::Int
    dims::NTuple{3,Int}
    nt::Int
    nrecs::Vector{Int}

    function HIPPlan(dims::Tuple, halo::Integer; precision::Symbol=:f64, nshots::Integer=1)
        ndim = length(dims)
        nz = ndim == 3 ? dims[3] : 1
        h = Ref{Ptr{Cvoid}}(C_NULL)
        hipcheck(ccall((:awcpml_plan_create, libawcpml), Cint,
                       (Cint, Int64, Int64, Int64, Cint, Cint, Cint, Ref{Ptr{Cvoid}}),
                       ndim, dims[1], dims[2], nz, halo, precision_code(precision), nshots, h))
        plan = new(h[], ndim, (Int(dims[1]), Int(dims[2]), Int(nz)), 0, zeros(Int, nshots))
        finalizer(destroy!, plan)
        return plan
    end
end

function destroy!(plan::HIPPlan)
    if plan.handle != C_NULL
        ccall((:awcpml_plan_destroy, libawcpml), Cint, (Ptr{Cvoid},), plan.handle)
        plan.handle = C_NULL
    end
    return nothing
end

fieldsize(plan::HIPPlan) = plan.ndim == 3 ? plan.dims : plan.dims[1:2]

"fact = dt^2 vel^2 and the inverse spacings (_dx, _dy[, _dz]) and their squares."
function set_model!(plan::HIPPlan, fact::AbstractArray{<:Real}, inv_d::Vector{Float64},
                    inv_d2::Vector{Float64})
    f = Array{Float64}(fact)
    hipcheck(ccall((:awcpml_set_model, libawcpml), Cint,
                   (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                   plan.handle, f, inv_d, inv_d2))
    return plan
end

"The eight calc_Kab_CPML vectors of one axis (1-based here, 0-based in the ABI)."
function set_cpml!(plan::HIPPlan, axis::Integer, a_l, a_r, b_K_l, b_K_r,
                   a_hl, a_hr, b_K_hl, b_K_hr)
    v = [Vector{Float64}(x) for x in (a_l, a_r, b_K_l, b_K_r, a_hl, a_hr, b_K_hl, b_K_hr)]
    hipcheck(ccall((:awcpml_set_cpml, libawcpml), Cint,
                   (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                    Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                   plan.handle, axis - 1, v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8]))
    return plan
end

"Sources of one shot: 1-based grid positions (nsrcs x ndim) and dt2srctf (nt x nsrcs)."
function set_sources!(plan::HIPPlan, shot::Integer, possrcs::AbstractMatrix{<:Integer},
                      dt2srctf::AbstractMatrix{<:Real})
    pos = Matrix{Int64}(possrcs)
    w = Matrix{Float64}(dt2srctf)
    hipcheck(ccall((:awcpml_set_sources, libawcpml), Cint,
                   (Ptr{Cvoid}, Cint, Cint, Ptr{Int64}, Int64, Ptr{Float64}),
                   plan.handle, shot - 1, size(pos, 1), pos, size(w, 1), w))
    plan.nt = size(w, 1)
    return plan
end

"Receivers of one shot: 1-based grid positions (nrecs x ndim)."
function set_receivers!(plan::HIPPlan, shot::Integer, posrecs::AbstractMatrix{<:Integer})
    pos = Matrix{Int64}(posrecs)
    hipcheck(ccall((:awcpml_set_receivers, libawcpml), Cint,
                   (Ptr{Cvoid}, Cint, Cint, Ptr{Int64}),
                   plan.handle, shot - 1, size(pos, 1), pos))
    plan.nrecs[shot] = size(pos, 1)
    return plan
end

"Two time steps per stencil launch where that applies (bit-identical results)."
function set_time_block!(plan::HIPPlan, steps::Integer; rows::Integer=0, march::Integer=0,
                         prefetch::Integer=0)
    hipcheck(ccall((:awcpml_set_time_block, libawcpml), Cint,
                   (Ptr{Cvoid}, Cint, Cint, Cint, Cint),
                   plan.handle, steps, rows, march, prefetch))
    return plan
end

"3D two-step launches: aligned 64-node x tiles where no CPML slab reaches (bit-identical)."
function set_two_step_align!(plan::HIPPlan, enable::Bool)
    hipcheck(ccall((:awcpml_set_two_step_align, libawcpml), Cint, (Ptr{Cvoid}, Cint),
                   plan.handle, enable))
    return plan
end

"Capture the time loop in HIP graphs (replayed on later solves)."
function set_graph!(plan::HIPPlan, enable::Bool)
    hipcheck(ccall((:awcpml_set_graph, libawcpml), Cint, (Ptr{Cvoid}, Cint),
                   plan.handle, enable ? 1 : 0))
    return plan
end

"Zero the wavefields, psi/xi and traces (the @zeros state)."
reset!(plan::HIPPlan) =
    hipcheck(ccall((:awcpml_reset, libawcpml), Cint, (Ptr{Cvoid},), plan.handle))

"Run the next `nsteps` time steps of every shot (the `for it=1:nt` loop)."
advance!(plan::HIPPlan, nsteps::Integer) =
    hipcheck(ccall((:awcpml_advance, libawcpml), Cint, (Ptr{Cvoid}, Int64), plan.handle, nsteps))

"One forward! call at time step `it` (1-based), as the reference's do_bench times it."
forward!(plan::HIPPlan, it::Integer) =
    hipcheck(ccall((:awcpml_forward, libawcpml), Cint, (Ptr{Cvoid}, Int64), plan.handle, it))

"Receiver traces of one shot, nt x nrecs (recs.seismograms)."
function traces(plan::HIPPlan, shot::Integer=1)
    out = zeros(Float64, plan.nt, plan.nrecs[shot])
    hipcheck(ccall((:awcpml_get_traces, libawcpml), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}),
                   plan.handle, shot - 1, out))
    return out
end

"A wavefield of one shot: which = 1 is pcur (the field solve2D/solve3D returns)."
function field(plan::HIPPlan, shot::Integer=1, which::Integer=1)
    out = zeros(Float64, fieldsize(plan)...)
    hipcheck(ccall((:awcpml_get_field, libawcpml), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}),
                   plan.handle, shot - 1, which, out))
    return out
end

"Forward run keeping checkpoints every `every` steps (0: round(sqrt(nt))) for `fact_gradient`."
checkpoint_run!(plan::HIPPlan, every::Integer=0) =
    hipcheck(ccall((:awcpml_checkpoint_run, libawcpml), Cint, (Ptr{Cvoid}, Int64),
                   plan.handle, every))

"Shots per thread of the reverse step (0: sized to the grid)."
set_adjoint_group!(plan::HIPPlan, shots::Integer) =
    hipcheck(ccall((:awcpml_set_adjoint_group, libawcpml), Cint, (Ptr{Cvoid}, Cint),
                   plan.handle, shots))

"Reverse steps summed in the plan's precision before the FP64 fold (0: a checkpoint segment, -1: never)."
set_gradient_fold!(plan::HIPPlan, steps::Integer) =
    hipcheck(ccall((:awcpml_set_gradient_fold, libawcpml), Cint, (Ptr{Cvoid}, Int64),
                   plan.handle, steps))

"dJ/dfact summed over the shots for trace cotangents (one nt x nrecs matrix per shot)."
function fact_gradient(plan::HIPPlan, trace_bar::Vector{<:AbstractMatrix})
    flat = reduce(vcat, [vec(Matrix{Float64}(t)) for t in trace_bar])
    out = zeros(Float64, fieldsize(plan)...)
    hipcheck(ccall((:awcpml_adjoint, libawcpml), Cint,
                   (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}),
                   plan.handle, flat, length(flat), out))
    return out
end
