@doc raw"""
Module with the FWI gradient of the HIP backend as a ChainRulesCore `rrule`.

    prob = FWIProblem(plan, dt, inv_d, inv_d2; every=0)
    tr = hip_traces(prob, vel)          # one (nt, nrecs) matrix per shot

`plan` is a `HIPPlan` whose CPML coefficients, sources and receivers are
set (dt, the coefficients and the wavelets stay those of the starting
model); `hip_traces` sets fact = dt^2 vel^2 and runs the forward solve,
keeping a checkpoint every `every` steps (0: round(sqrt(nt))).  The rrule's
pullback runs the hand-written adjoint (`awcpml_adjoint`) and returns
dJ/dvel = dJ/dfact * 2 dt^2 vel.  The reference lists the adjoint and
checkpointing as work in progress (README.md:297).  A pullback is only valid
until the next `hip_traces` on the same problem (the checkpoints are
replaced); a stale pullback raises.
"""
module AcousticFWI_HIP

using ChainRulesCore

include("libawcpml.jl")

export FWIProblem, hip_traces

mutable struct FWIProblem
    plan::HIPPlan
    dt::Float64
    inv_d::Vector{Float64}
    inv_d2::Vector{Float64}
    every::Int
    runs::Int
end

FWIProblem(plan::HIPPlan, dt::Real, inv_d, inv_d2; every::Integer=0) =
    FWIProblem(plan, Float64(dt), Vector{Float64}(inv_d), Vector{Float64}(inv_d2), Int(every), 0)

function hip_traces(prob::FWIProblem, vel::AbstractArray{<:Real})
    set_model!(prob.plan, (prob.dt^2) .* (vel .^ 2), prob.inv_d, prob.inv_d2)
    checkpoint_run!(prob.plan, prob.every)
    prob.runs += 1
    return [traces(prob.plan, s) for s = 1:length(prob.plan.nrecs)]
end

function ChainRulesCore.rrule(::typeof(hip_traces), prob::FWIProblem, vel::AbstractArray{<:Real})
    tr = hip_traces(prob, vel)
    run = prob.runs
    function hip_traces_pullback(tbar)
        run == prob.runs || error("stale pullback: hip_traces ran again on this problem " *
                                  "after the forward call it belongs to")
        tb = unthunk(tbar)
        cot = Matrix{Float64}[]
        for s = 1:length(tr)
            t = tb isa AbstractZero ? tb : unthunk(tb[s])
            push!(cot, t isa AbstractZero ? zeros(size(tr[s])) : Matrix{Float64}(t))
        end
        gf = fact_gradient(prob.plan, cot)
        return NoTangent(), NoTangent(), gf .* (2 * prob.dt^2) .* vel
    end
    return tr, hip_traces_pullback
end

end # module AcousticFWI_HIP
