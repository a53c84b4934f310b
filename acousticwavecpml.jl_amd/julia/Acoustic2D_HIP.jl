@doc raw"""
Module with the acoustic 2D xPU solver on AMD GPUs: the time loop runs in
libawcpml (hand-written HIP kernels for gfx950) through `ccall`.

    solve2D(lx, ly, lt, vel, srcs, recs; halo=20, rcoef=0.0001, ppw=10.0,
            freetop=true, do_bench=false, do_vis=false, nvis=5,
            gif_name="acoustic2D_xPU", plims=[-1, 1], threshold=0.01,
            precision=:f64)

Same entry point, keyword arguments, assertions and outputs as the
`Acoustic2D_Threads` / `Acoustic2D_CUDA` backends
(src/shared/acoustic_2D_xPU.jl:133-444): returns the final pressure field and
fills `recs.seismograms` with the (nt, nrecs) traces.  The host setup (CFL dt,
calc_Kab_CPML, fact, source wavelets, position snapping) is done here in
Float64 exactly as there; the per-step kernels, the buffer rotation and the
receiver recording run on the GPU.  `precision=:f32` runs the FP32 kernels
(setup still in Float64).  `do_vis` (Plots.jl animation) is not supported.
"""
module Acoustic2D_HIP

using BenchmarkTools
using Printf

import ..AcousticWaveCPML: calc_Kab_CPML, check_trial, Sources, Receivers

include("libawcpml.jl")

export solve2D

# round(Int, x, RoundNearestTiesUp) of the reference's position snapping
snap(x::Real) = round(Int, x, RoundNearestTiesUp)

function solve2D(lx::Real, ly::Real, lt::Real, vel::Matrix{<:Real}, srcs::Sources,
                 recs::Receivers;
                 halo::Integer=20, rcoef::Real=0.0001, ppw::Real=10.0, freetop::Bool=true,
                 do_bench::Bool=false, do_vis::Bool=false, nvis::Integer=5,
                 gif_name::String="acoustic2D_xPU", plims::Vector{<:Real}=[-1, 1],
                 threshold::Real=0.01, precision::Symbol=:f64)
    do_vis && throw(ArgumentError("do_vis is not supported by the HIP backend"))
    # --- host setup in Float64, as acoustic_2D_xPU.jl:155-201 ---------------
    f0 = srcs.freqdomain
    vel_max = maximum(vel)
    nx, ny = size(vel)
    dx, dy = lx / (nx - 1), ly / (ny - 1)
    dt = sqrt(2) / (vel_max * (1 / dx + 1 / dy)) / 2     # CFL limit
    nt = ceil(Int, lt / dt)
    times = collect(range(0.0, step=dt, length=nt))
    npower, K_max, alpha_max = 2.0, 1.0, π * f0
    d0_x = -(npower + 1) * vel_max * log(rcoef) / (2.0 * (halo * dx))
    d0_y = -(npower + 1) * vel_max * log(rcoef) / (2.0 * (halo * dy))
    cx = (calc_Kab_CPML(halo, dt, npower, d0_x, alpha_max, K_max, "ongrd")...,
          calc_Kab_CPML(halo, dt, npower, d0_x, alpha_max, K_max, "halfgrd")...)
    cy = [calc_Kab_CPML(halo, dt, npower, d0_y, alpha_max, K_max, "ongrd")...,
          calc_Kab_CPML(halo, dt, npower, d0_y, alpha_max, K_max, "halfgrd")...]
    if freetop  # no absorption on the top (y-low) side
        cy[1] = zeros(halo); cy[5] = zeros(halo + 1)
        cy[3] = ones(halo);  cy[7] = ones(halo + 1)
    end
    inv_d = [1.0 / dx, 1.0 / dy]
    inv_d2 = [1.0 / (dx^2), 1.0 / (dy^2)]
    fact = (dt^2) .* (vel .^ 2)
    @assert sqrt(dx^2 + dy^2) <= vel_max / (ppw * f0) "Not enough points per wavelength!"
    nsrcs = srcs.n
    dt2srctf = zeros(nt, nsrcs)
    for s = 1:nsrcs
        dt2srctf[:, s] .= (dt^2 / (dx * dy)) .* srcs.srctfs[s].(times, srcs.t0s[s], f0)
    end
    possrcs = zeros(Int, size(srcs.positions))
    for s = 1:nsrcs
        possrcs[s, :] .= snap.([srcs.positions[s, 1] / dx + 1, srcs.positions[s, 2] / dy + 1])
    end
    @assert all(1 .<= possrcs[:, 1] .<= nx) && all(1 .<= possrcs[:, 2] .<= ny) "At least one source is not inside the model!"
    posrecs = zeros(Int, size(recs.positions))
    for r = 1:recs.n
        posrecs[r, :] .= snap.([recs.positions[r, 1] / dx + 1, recs.positions[r, 2] / dy + 1])
    end
    @assert all(1 .<= posrecs[:, 1] .<= nx) && all(1 .<= posrecs[:, 2] .<= ny) "At least one receiver is not inside the model!"

    # --- the GPU plan: device arrays, coefficients, sources, receivers ------
    plan = HIPPlan((nx, ny), halo; precision=precision)
    try
        set_model!(plan, fact, inv_d, inv_d2)
        set_cpml!(plan, 1, cx...)
        set_cpml!(plan, 2, cy...)
        set_sources!(plan, 1, possrcs, dt2srctf)
        set_receivers!(plan, 1, posrecs)

        if do_bench
            # one forward! at it = 1, timed like acoustic_2D_xPU.jl:263-297
            forward!(plan, 1)
            trial = @benchmark forward!($plan, 1)
            confidence, med_range = 0.95, 0.05
            pass, ci, tol_range, t_med = check_trial(trial, confidence, med_range)
            t_it = minimum(trial).time / 1e9
            if !pass
                @printf("Statistical trial check not passed!\nmedian = %g [sec]\n%d%% tolerance range = (%g, %g) [sec]\n%d%% CI = (%g, %g) [sec]\n",
                        t_med, round(Int, med_range * 100), tol_range[1], tol_range[2],
                        round(Int, confidence * 100), ci[1], ci[2])
            end
            bytes = precision === :f64 ? sizeof(Float64) : sizeof(Float32)
            A_eff = ((halo + 1) * ny * 2 * 3 + (halo + 1) * nx * 2 * 3 + (halo + 1) * ny * 2 * 2 +
                     (halo + 1) * nx * 2 * 2 + 4 * nx * ny) * bytes / 1e9
            @printf("size = %dx%d, time = %1.3e sec, Teff = %1.3f GB/s\n", nx, ny, t_it, A_eff / t_it)
            return nothing
        end

        # --- the time loop (acoustic_2D_xPU.jl:327-402) on the GPU -----------
        reset!(plan)
        advance!(plan, nt)
        recs.seismograms = traces(plan, 1)
        return field(plan, 1, 1)
    finally
        destroy!(plan)
    end
end

end # module Acoustic2D_HIP
