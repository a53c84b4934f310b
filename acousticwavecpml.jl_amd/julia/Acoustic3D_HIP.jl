@doc raw"""
Module with the acoustic 3D xPU solver on AMD GPUs: the time loop runs in
libawcpml (hand-written HIP kernels for gfx950) through `ccall`.

    solve3D(lx, ly, lz, lt, vel, srcs, recs; halo=20, rcoef=0.0001, ppw=10.0,
            freetop=true, do_bench=false, do_vis=false, do_save=false, nvis=5,
            nsave=5, gif_name="acoustic3D_xPU_slice", save_name="acoustic3D_xPU",
            plims=[-1.0, 1.0], threshold=0.01, precision=:f64, time_block=2)

Same entry point, keyword arguments, assertions and outputs as the
`Acoustic3D_Threads` / `Acoustic3D_CUDA` backends
(src/shared/acoustic_3D_xPU.jl:169-544): returns the final pressure field and
fills `recs.seismograms`.  `do_save` writes every `nsave` steps the HDF5 file
`<save_name>_it<it>.h5` in `simulations/tmp` with the datasets lx, ly, lz,
halo, pcur, possrcs, posrecs, as the reference does
(acoustic_3D_xPU.jl:480-498).  `precision=:f32` runs the FP32 kernels;
`time_block=2` lets the library advance two time steps per launch where that
applies (bit-identical results).  `do_vis` is not supported.
"""
module Acoustic3D_HIP

using BenchmarkTools
using HDF5
using Printf

import ..AcousticWaveCPML: TMP_FLD, calc_Kab_CPML, check_trial, Sources, Receivers

include("libawcpml.jl")

export solve3D

snap(x::Real) = round(Int, x, RoundNearestTiesUp)

# CPML coefficients of one axis: calc_Kab_CPML "ongrd" then "halfgrd"
function axis_coefficients(halo, dt, npower, d0, alpha_max, K_max)
    return [calc_Kab_CPML(halo, dt, npower, d0, alpha_max, K_max, "ongrd")...,
            calc_Kab_CPML(halo, dt, npower, d0, alpha_max, K_max, "halfgrd")...]
end

function grid_positions(positions, d, n, what)
    pos = zeros(Int, size(positions))
    for r = 1:size(positions, 1)
        pos[r, :] .= snap.([positions[r, a] / d[a] + 1 for a = 1:3])
    end
    @assert all(1 .<= pos[:, 1] .<= n[1]) && all(1 .<= pos[:, 2] .<= n[2]) && all(1 .<= pos[:, 3] .<= n[3]) "At least one $(what) is not inside the model!"
    return pos
end

function save_snapshot(it, lx, ly, lz, halo, pcur, possrcs, posrecs, save_name)
    path = joinpath(TMP_FLD, "$(save_name)_it$(it).h5")
    isfile(path) && rm(path)
    h5write(path, "lx", lx)
    h5write(path, "ly", ly)
    h5write(path, "lz", lz)
    h5write(path, "halo", halo)
    h5write(path, "pcur", pcur)
    h5write(path, "possrcs", possrcs)
    h5write(path, "posrecs", posrecs)
    return path
end

function solve3D(lx::Real, ly::Real, lz::Real, lt::Real, vel::Array{<:Real,3}, srcs::Sources,
                 recs::Receivers;
                 halo::Integer=20, rcoef::Real=0.0001, ppw::Real=10.0, freetop::Bool=true,
                 do_bench::Bool=false, do_vis::Bool=false, do_save::Bool=false,
                 nvis::Integer=5, nsave::Integer=5, gif_name::String="acoustic3D_xPU_slice",
                 save_name::String="acoustic3D_xPU", plims::Vector{<:Real}=[-1.0, 1.0],
                 threshold::Real=0.01, precision::Symbol=:f64, time_block::Integer=2)
    do_vis && throw(ArgumentError("do_vis is not supported by the HIP backend"))
    # --- host setup in Float64, as acoustic_3D_xPU.jl:191-285 ---------------
    f0 = srcs.freqdomain
    vel_max = maximum(vel)
    nx, ny, nz = size(vel)
    d = (lx / (nx - 1), ly / (ny - 1), lz / (nz - 1))
    dt = sqrt(3) / (vel_max * (1 / d[1] + 1 / d[2] + 1 / d[3])) / 2    # CFL limit
    nt = ceil(Int, lt / dt)
    times = collect(range(0.0, step=dt, length=nt))
    npower, K_max, alpha_max = 2.0, 1.0, π * f0
    coefs = [axis_coefficients(halo, dt, npower,
                               -(npower + 1) * vel_max * log(rcoef) / (2.0 * (halo * d[a])),
                               alpha_max, K_max) for a = 1:3]
    if freetop  # no absorption on the top (y-low) side
        coefs[2][1] = zeros(halo); coefs[2][5] = zeros(halo + 1)
        coefs[2][3] = ones(halo);  coefs[2][7] = ones(halo + 1)
    end
    inv_d = [1.0 / d[a] for a = 1:3]
    inv_d2 = [1.0 / (d[a]^2) for a = 1:3]
    fact = (dt^2) .* (vel .^ 2)
    @assert sqrt(d[1]^2 + d[2]^2 + d[3]^2) <= vel_max / (ppw * f0) "Not enough points per wavelength!"
    nsrcs = srcs.n
    dt2srctf = zeros(nt, nsrcs)
    for s = 1:nsrcs
        dt2srctf[:, s] .= (dt^2 / (d[1] * d[2] * d[3])) .* srcs.srctfs[s].(times, srcs.t0s[s], f0)
    end
    possrcs = grid_positions(srcs.positions, d, (nx, ny, nz), "source")
    posrecs = grid_positions(recs.positions, d, (nx, ny, nz), "receiver")

    # --- the GPU plan ---------------------------------------------------------
    plan = HIPPlan((nx, ny, nz), halo; precision=precision)
    try
        set_model!(plan, fact, inv_d, inv_d2)
        for a = 1:3
            set_cpml!(plan, a, coefs[a]...)
        end
        set_sources!(plan, 1, possrcs, dt2srctf)
        set_receivers!(plan, 1, posrecs)

        if do_bench
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
            A_eff = ((halo + 1) * ny * nz * 2 * 3 + (halo + 1) * nx * nz * 2 * 3 +
                     (halo + 1) * nx * ny * 2 * 3 + (halo + 1) * ny * nz * 2 * 2 +
                     (halo + 1) * nx * nz * 2 * 2 + (halo + 1) * nx * ny * 2 * 2 +
                     4 * nx * ny * nz) * bytes / 1e9
            @printf("size = %dx%dx%d, time = %1.3e sec, Teff = %1.3f GB/s\n", nx, ny, nz, t_it, A_eff / t_it)
            return nothing
        end

        # --- the time loop (acoustic_3D_xPU.jl:392-499) on the GPU -----------
        set_time_block!(plan, time_block)
        reset!(plan)
        if do_save
            mkpath(TMP_FLD)
            done = 0
            while done < nt
                n = min(nsave - done % nsave, nt - done)
                advance!(plan, n)
                done += n
                if done % nsave == 0
                    save_snapshot(done, lx, ly, lz, halo, field(plan, 1, 1), possrcs, posrecs,
                                  save_name)
                end
            end
        else
            advance!(plan, nt)
        end
        recs.seismograms = traces(plan, 1)
        return field(plan, 1, 1)
    finally
        destroy!(plan)
    end
end

end # module Acoustic3D_HIP
