"""Throughput of the CPML time-stepping hot path on one MI355X.

Workload: the 3D plain CPML forward solve -- 256^3 interior + 20-point CPML
halo (298^3 nodes), free top, dx = 10 m, c(depth) = 1500 + 0.6 depth with a
+-5 % seeded smoothed perturbation, Ricker f0 = 10 Hz, 64 receivers,
nt = 2000 time steps (tests/cases.py: plain3d).

One bench "step" is one whole solve of that workload: zero the state and run
all nt time steps of every shot (inputs already resident in HBM).  Three
variants run on the same model:
  fp64_plain    FP64, one shot per launch (the drop-in path)
  fp32          FP32, one shot
  fp32_batched  FP32, S shots per launch (fact and CPML work shared)
`value` is the FP32-batched throughput in shot*Gpts/s (shot x nodes x nt /
time); the FP64 plain figure and the ratio are reported beside it.

The roofline entry is for the dominant kernel (the fused stencil) of the
`value` variant: algorithmic bytes per launch
  B = sizeof(T) * [S*(3N + C) + N],  C = sum over absorbing sides of
      (4*halo + 2) * N / n_axis
divided by its average launch time from HIP events on the plan's stream.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "acousticwavecpml.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

METRIC = ("shot·Gpts/s + effective HBM GB/s (% of peak), FP32-batched vs FP64-plain, "
          "same GPU")
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(n, halo, freetop, esize, nshots):
    N = 1
    for v in n:
        N *= v
    C = 0
    for a, na in enumerate(n):
        sides = 1 if (freetop and a == 1) else 2
        if halo > 0:
            C += sides * (4 * halo + 2) * (N // na)
    return esize * (nshots * (3 * N + C) + N), N


def build_plan(S, dtype, nshots):
    """Plan of `nshots` shots on the workload's model: shot 0 at the
    configured source, the others spread along x at the source depth, like
    a line survey."""
    import awcpml
    P = awcpml.Plan(S.n, S.halo, dtype=dtype, nshots=nshots)
    P.set_model(S.fact, S.inv_d, S.inv_d2)
    for a in range(S.ndim):
        P.set_cpml(a, S.coef[a])
    for s in range(nshots):
        pos = S.possrcs.copy()
        if s > 0:
            pos[0, 0] = 1 + S.halo + 1 + (s * (S.n[0] - 2 * S.halo - 4)) // nshots
        P.set_sources(s, pos, S.dt2srctf)
        P.set_receivers(s, S.posrecs)
    return P


def time_variant(S, dtype, nshots, steps, warmup, freetop, kernel_launches=50, tune=True):
    import awcpml
    P = build_plan(S, dtype, nshots)
    esize = 8 if dtype == "f64" else 4
    if tune:
        awcpml.autotune(P, launches=5)
    P.set_graph(True)  # the whole nt-step loop replays as one HIP graph
    for _ in range(warmup):
        P.reset()
        P.advance(S.nt)
    t0 = time.perf_counter()
    for _ in range(steps):
        P.reset()
        P.advance(S.nt)
    t = (time.perf_counter() - t0) / max(steps, 1)
    traces0, field0 = P.traces(0), P.field(0, 1)
    # dominant kernel alone, HIP events on the plan's stream: the two-step
    # kernel (two time steps per launch) when the tuner chose it
    tb = P.time_block()
    P.reset()
    if tb[0] == 2:
        nl = max(1, kernel_launches // 2)
        ms = P.time_stencil2(nl) / nl
    else:
        ms = P.time_stencil(kernel_launches) / kernel_launches
    tiling = P.tiling() + (P.shot_group(),)
    P.close()
    B, N = algorithmic_bytes(S.n, S.halo, freetop, esize, nshots)
    B_launch = B * tb[0]
    pts = nshots * N * S.nt
    return {
        "dtype": dtype, "shots": nshots, "s_per_solve": t,
        "shot_gpts_per_s": pts / t / 1e9,
        "effective_gbs": B * S.nt / t / 1e9,
        "kernel": "stencil3d_two" if tb[0] == 2 else "stencil3d",
        "steps_per_launch": tb[0],
        "kernel_ms": ms, "kernel_gbs": B_launch / (ms * 1e-3) / 1e9,
        "bytes_per_launch": B_launch,
        "tiling": {"block_y": tiling[0], "march_len": tiling[1], "prefetch": tiling[2],
                   "shot_group": tiling[3]},
        "time_block": {"steps": tb[0], "block_rows": tb[1], "march_len": tb[2], "prefetch": tb[3]},
        "_traces0": traces0, "_field0": field0,
    }


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, gpu, case2d=None, gpu2d=None):
    """The oracle (C/OpenMP FP64 restatement of the reference's Threads
    path) on the host cores: the 3D workload in full on all cores, a short
    sample on one core, and the 2D workload in full.  As the checker, its
    outputs also give the GPU results' errors: FP64 bit for bit, FP32 as
    relative L2 of the traces and of the final field."""
    import oracle
    from cases import oracle_problem, rel_l2
    N = int(np.prod(case["n"]))
    cores = oracle.set_threads(0)
    o = oracle_problem(case)
    t0 = time.perf_counter()
    done = o.run()
    t = time.perf_counter() - t0
    tr, pend = o.traces(), o.pend()
    o.close()
    parity = {
        "fp64_plain_bitwise_traces": bool(np.array_equal(gpu["f64"]["_traces0"], tr)),
        "fp64_plain_bitwise_field": bool(np.array_equal(gpu["f64"]["_field0"], pend)),
    }
    for key, v in gpu.items():
        if key == "f64":
            continue
        parity[f"{key}_trace_rel_l2"] = rel_l2(v["_traces0"], tr)
        parity[f"{key}_field_rel_l2"] = rel_l2(v["_field0"], pend)
    # one core, a short sample of the same workload
    oracle.set_threads(1)
    o = oracle_problem(case)
    o.run(2)
    t1 = time.perf_counter()
    n1 = o.run(8)
    t_one = time.perf_counter() - t1
    o.close()
    oracle.set_threads(0)
    out = {
        "value": done * N / t / 1e9, "unit": "shot·Gpts/s", "cores": cores, "kind": "port",
        "sample": f"plain3d {case['n'][0]}^3 FP64 single shot, all {done} time steps ({t:.1f} s), "
                  f"C/OpenMP oracle on {cores} threads",
        "cpu_model": _cpu_model(),
        "one_core": {"value": n1 * N / t_one / 1e9, "unit": "shot·Gpts/s",
                     "sample": f"time steps 3..{2 + n1} of the same solve on 1 thread ({t_one:.1f} s)"},
        "parity_vs_oracle": parity,
    }
    if case2d is not None:
        o = oracle_problem(case2d)
        N2 = int(np.prod(case2d["n"]))
        t2 = time.perf_counter()
        d2 = o.run()
        t2 = time.perf_counter() - t2
        out["plain2d"] = {
            "value": d2 * N2 / t2 / 1e9, "unit": "shot·Gpts/s", "cores": cores,
            "sample": f"plain2d {case2d['n'][0]}^2 FP64, all {d2} time steps ({t2:.2f} s)",
            "gpu_fp64_bitwise_traces": bool(gpu2d is not None and
                                            np.array_equal(gpu2d["_traces0"], o.traces())),
            "gpu_fp64_bitwise_field": bool(gpu2d is not None and
                                           np.array_equal(gpu2d["_field0"], o.pend())),
        }
        o.close()
    return out


def bench_plain2d(steps=10):
    """The 2D drop-in workload (298^2, FP64, one shot, nt = 1000): whole
    solves with the time loop in a HIP graph, autotuned launch shape."""
    import awcpml
    from cases import plain2d
    case = plain2d()
    S = _setup(case)
    P = build_plan(S, "f64", 1)
    tune = awcpml.autotune(P, launches=5)
    P.set_graph(True)
    P.reset()
    P.advance(S.nt)
    t0 = time.perf_counter()
    for _ in range(steps):
        P.reset()
        P.advance(S.nt)
    t = (time.perf_counter() - t0) / steps
    tr, fld = P.traces(0), P.field(0, 1)
    P.reset()
    ms = P.time_stencil(50) / 50
    P.close()
    N = S.n[0] * S.n[1]
    B, _ = algorithmic_bytes(S.n, S.halo, False, 8, 1)
    return case, {"workload": f"plain2d: 2D {S.n[0]}^2 (256^2 interior + 20 halo), FP64, one shot, "
                              f"nt={S.nt}", "s_per_solve": round(t, 5),
                  "shot_gpts_per_s": round(N * S.nt / t / 1e9, 3),
                  "kernel_ms_per_step": round(ms, 5), "kernel_gbs": round(B / ms / 1e6, 1),
                  "tiling": [tune["block_y"], tune["march_len"], tune["prefetch"]],
                  "time_block": tune["time_block"], "_traces0": tr, "_field0": fld}


def _setup(case, vel=None, srcpos=None):
    import awcpml
    from awcpml.host import setup
    vel = case["vel"] if vel is None else vel
    srcpos = case["srcpos"] if srcpos is None else srcpos
    srcs = awcpml.Sources(srcpos, case["t0s"] * len(srcpos), [awcpml.rickersource1D] * len(srcpos),
                          case["f0"])
    recs = awcpml.Receivers(case["recpos"])
    return setup(case["l"], case["lt"], vel, srcs, recs, halo=case["halo"], rcoef=case["rcoef"],
                 ppw=case["ppw"], freetop=case["freetop"])


def _shot_plan(S, shots, dtype, first=0, count=None):
    """Plan over shots[first:first+count] (one source each) on setup S."""
    import awcpml
    from awcpml.host import round_ties_up
    count = len(shots) - first if count is None else count
    P = awcpml.Plan(S.n, S.halo, dtype=dtype, nshots=count)
    P.set_model(S.fact, S.inv_d, S.inv_d2)
    for a in range(S.ndim):
        P.set_cpml(a, S.coef[a])
    for q in range(count):
        pos = np.asarray(shots[first + q], dtype=np.float64)
        idx = np.array([[int(round_ties_up(pos[0, a] / S.d[a] + 1)) for a in range(S.ndim)]])
        P.set_sources(q, idx, S.dt2srctf)
        P.set_receivers(q, S.posrecs)
    return P


def bench_batch2d(steps):
    """Many shots on one 2D model in one launch per step, graph-captured
    time loop; error of two shots against the FP64 path."""
    import awcpml
    from cases import batch2d, rel_l2
    case = batch2d()
    S = _setup(case)
    P = _shot_plan(S, case["shots"], "f32")
    tune = awcpml.autotune(P, launches=5)
    P.set_graph(True)
    P.reset()
    P.advance(S.nt)
    t0 = time.perf_counter()
    for _ in range(steps):
        P.reset()
        P.advance(S.nt)
    t = (time.perf_counter() - t0) / steps
    P.reset()
    steps_per_launch = P.time_block()[0]
    if steps_per_launch == 2:
        ms = P.time_stencil2(25) / 25 / 2  # per time step
    else:
        ms = P.time_stencil(50) / 50
    P.reset()
    P.advance(S.nt)
    e = 0.0
    for q in (0, len(case["shots"]) - 1):
        R = _shot_plan(S, case["shots"], "f64", first=q, count=1)
        R.reset()
        R.advance(S.nt)
        e = max(e, rel_l2(P.traces(q), R.traces(0)))
        R.close()
    N = S.n[0] * S.n[1]
    B, _ = algorithmic_bytes(S.n, S.halo, False, 4, P.nshots)
    out = {"workload": f"2D {S.n[0]}^2 (1024^2 interior + 20 halo), {P.nshots} shots, nt={S.nt}, FP32",
           "s_per_solve": round(t, 4), "shot_gpts_per_s": round(P.nshots * N * S.nt / t / 1e9, 2),
           "kernel_ms_per_step": round(ms, 4), "kernel_gbs": round(B / ms / 1e6, 1),
           "steps_per_launch": steps_per_launch,
           "trace_rel_l2_vs_fp64_max": e,
           "tiling": [tune["block_y"], tune["march_len"], tune["prefetch"], tune["shot_group"]],
           "time_block": tune["time_block"]}
    P.close()
    return out


def bench_uneven3d():
    """12 3D models of uneven size in one launch per step, or per two steps
    when the autotuner picks two-step launches (FP32), then each model alone
    in FP64 for its trace error."""
    import awcpml
    from cases import lt_for_nt, rel_l2, uneven3d
    cases = uneven3d()
    plans, setups = [], []
    for c in cases:
        vel = c["make_vel"]()
        c = dict(c)
        c["lt"] = lt_for_nt(3, (10.0,) * 3, float(vel.max()), c["nt"])
        S = _setup(c, vel=vel)
        del vel
        P = awcpml.Plan.from_setup(S, dtype="f32")
        plans.append(P)
        setups.append(S)
    tune = awcpml.autotune_batch(plans)
    B = awcpml.Batch(plans)
    B.set_graph(True)
    B.run()
    t0 = time.perf_counter()
    B.run()
    t = time.perf_counter() - t0
    pts = sum(P.nt * int(np.prod(S.n)) for P, S in zip(plans, setups))
    steps_per_launch = B.time_block()[0]
    if steps_per_launch == 2:
        ms = B.time_stencil2(5) / 5 / 2  # per time step
    else:
        ms = B.time_stencil(10) / 10
    bytes_step = sum(algorithmic_bytes(S.n, S.halo, False, 4, 1)[0] for S in setups)
    B.run()
    errs = []
    for P, S in zip(plans, setups):
        R = awcpml.Plan.from_setup(S, dtype="f64")
        R.reset()
        R.advance(S.nt)
        errs.append(rel_l2(P.traces(0), R.traces(0)))
        R.close()
    B.close()
    for P in plans:
        P.close()
    return {"workload": f"{len(cases)} 3D models, edges {[S.n[0] for S in setups]} nodes, one shot "
                        f"each, nt={setups[0].nt}, FP32, one batch launch per "
                        f"{'two steps' if steps_per_launch == 2 else 'step'}",
            "s_per_solve": round(t, 4), "shot_gpts_per_s": round(pts / t / 1e9, 2),
            "kernel_ms_per_step": round(ms, 4), "kernel_gbs": round(bytes_step / ms / 1e6, 1),
            "steps_per_launch": steps_per_launch,
            "trace_rel_l2_vs_fp64": [float("%.3g" % e) for e in errs],
            "tiling": [tune["block_y"], tune["march_len"], tune["prefetch"]],
            "time_block": tune["time_block"]}


def bench_fwi2d():
    """FWI gradient of the L2 trace misfit (32 shots, 554^2, FP32) with
    checkpointing every 50 steps, and its error against the FP64 gradient."""
    import awcpml
    from awcpml import fwi
    from cases import fwi2d, rel_l2
    case = fwi2d()
    S = _setup(case)
    fact0 = np.asfortranarray((S.dt * S.dt) * (case["start"] * case["start"]))
    res = {}
    for dtype in ("f32", "f64"):
        # observed data from the true model at the same precision, so the
        # residual holds no rounding difference of the direct arrivals
        obs_plan = _shot_plan(S, case["shots"], dtype)
        obs_plan.reset()
        obs_plan.advance(S.nt)
        obs = [obs_plan.traces(q) for q in range(obs_plan.nshots)]
        obs_plan.close()
        P = _shot_plan(S, case["shots"], dtype)
        P.set_model(fact0, S.inv_d, S.inv_d2)
        awcpml.autotune(P, launches=4)  # forward launch shape, incl. two-step launches
        fwi.misfit_gradient(P, obs, every=case["every"])  # warm-up
        t0 = time.perf_counter()
        J, g = fwi.misfit_gradient(P, obs, every=case["every"])
        res[dtype] = (time.perf_counter() - t0, J, g)
        P.close()
    t32, J32, g32 = res["f32"]
    N = S.n[0] * S.n[1]
    gv = fwi.velocity_gradient(g32, case["start"], S.dt)
    c = S.n[0] // 2
    return {"workload": f"2D {S.n[0]}^2, {len(case['shots'])} shots, nt={S.nt}, FP32, checkpoint "
                        f"every {case['every']} steps (forward, replay, adjoint)",
            "s_per_gradient": round(t32, 4),
            "shot_gpts_per_s_forward_equiv": round(len(case["shots"]) * N * S.nt / t32 / 1e9, 2),
            "misfit": J32, "misfit_fp64": res["f64"][1],
            "grad_rel_l2_vs_fp64": rel_l2(g32, res["f64"][2]),
            "s_per_gradient_fp64": round(res["f64"][0], 4),
            "vel_grad_sign_at_anomaly": float(np.sign(gv[c, c]))}


def load_traffic(kernel, workload):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            rec = json.load(fh)
        ent = rec.get(kernel, {})
        if ent.get("workload") == workload:
            return ent.get("bytes_per_launch")
    except (OSError, ValueError):
        pass
    return None


def calibration_copy(nbytes=1 << 30, repeats=3):
    """A plain device-to-device copy of `nbytes` (`repeats` times) through
    the HIP runtime the library uses: the known-size reference that scales
    the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    from awcpml import _lib
    _lib.load()
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipDeviceSynchronize.argtypes = []
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), nbytes) == 0 and hip.hipMalloc(ctypes.byref(dst), nbytes) == 0
    assert hip.hipMemset(src, 1, nbytes) == 0
    for _ in range(repeats):
        assert hip.hipMemcpy(dst, src, nbytes, 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(f"[profile] calibration copy: {repeats} x {nbytes} bytes read + written", flush=True)


def profile_kernel(S, dtype, nshots, tiling, launches, time_block=None):
    """Profiler target: only the stencil launches of one variant at a fixed
    launch shape (no autotuning, so every dispatch of the kernel is the same
    shape), then a 1 GiB device copy (1 GiB read + 1 GiB written) that
    calibrates the FETCH_SIZE / WRITE_SIZE counters."""
    P = build_plan(S, dtype, nshots)
    if tiling:
        P.set_tiling(*tiling[:3])
        if len(tiling) > 3:
            P.set_shot_group(tiling[3])
    steps = 1
    if time_block:
        P.set_time_block(2, *time_block[:3])
        steps = 2
    P.reset()
    timer = P.time_stencil2 if steps == 2 else P.time_stencil
    timer(1)  # loads the kernel's code object outside the timed launches
    ms = timer(launches) / launches
    esize = 8 if dtype == "f64" else 4
    B, _ = algorithmic_bytes(S.n, S.halo, True, esize, nshots)
    B *= steps
    print(f"[profile] {dtype} S={nshots} tiling={P.tiling()} shot_group={P.shot_group()} "
          f"time_block={P.time_block()} {launches} launches "
          f"{ms:.4f} ms/launch, {B} algorithmic bytes/launch, {B / ms / 1e6:.1f} GB/s", flush=True)
    P.close()
    calibration_copy()


def profile_batch2d(args):
    """Profiler target: only the stencil launches of the 64-shot 2D workload
    (1066^2, FP32) at a fixed launch shape (no autotuning), then the 1 GiB
    calibration copy."""
    from cases import batch2d
    case = batch2d()
    S = _setup(case)
    P = _shot_plan(S, case["shots"], "f32")
    tiling = tuple(int(v) for v in args.tiling.split(",")) if args.tiling else (4, 64, 3, 4)
    P.set_tiling(*tiling[:3])
    if len(tiling) > 3:
        P.set_shot_group(tiling[3])
    steps = 1
    if args.time_block:
        P.set_time_block(2, *(int(v) for v in args.time_block.split(",")))
        steps = 2
    P.reset()
    n = args.launches
    timer = P.time_stencil2 if steps == 2 else P.time_stencil
    timer(1)
    ms = timer(n) / n
    B, _ = algorithmic_bytes(S.n, S.halo, False, 4, P.nshots)
    print(f"[profile] batch2d FP32 S={P.nshots} tiling={P.tiling()} shot_group={P.shot_group()} "
          f"steps/launch={steps} {n} launches {ms:.4f} ms/launch, {B * steps} algorithmic bytes/launch, "
          f"{B * steps / ms / 1e6:.1f} GB/s", flush=True)
    P.close()
    calibration_copy()


def profile_uneven(args):
    """Profiler target: only the batch stencil launches of the uneven 3D
    workload at a fixed launch shape (no autotuning)."""
    import awcpml
    from cases import lt_for_nt, uneven3d
    plans, setups = [], []
    for c in uneven3d():
        vel = c["make_vel"]()
        c = dict(c)
        c["lt"] = lt_for_nt(3, (10.0,) * 3, float(vel.max()), c["nt"])
        S = _setup(c, vel=vel)
        del vel
        plans.append(awcpml.Plan.from_setup(S, dtype="f32"))
        setups.append(S)
    tiling = tuple(int(v) for v in args.tiling.split(",")) if args.tiling else (4, 64, 2)
    for P in plans:
        P.set_tiling(*tiling[:3])
    B = awcpml.Batch(plans)
    steps = 1
    if args.time_block:
        w, ml, pf = (int(v) for v in args.time_block.split(","))
        for P in plans:
            P.set_time_block(1, 0, ml)
        B.set_time_block(2, w, pf)
        steps = 2
    B.reset()
    n = args.launches
    timer = B.time_stencil2 if steps == 2 else B.time_stencil
    timer(1)  # loads the kernel's code object outside the timed launches
    ms = timer(n) / n
    bytes_step = sum(algorithmic_bytes(S.n, S.halo, False, 4, 1)[0] for S in setups)
    print(f"[profile] uneven3d batch FP32 time_block={B.time_block()} {n} launches "
          f"{ms:.4f} ms/launch, {bytes_step * steps} algorithmic bytes/launch, "
          f"{bytes_step * steps / ms / 1e6:.1f} GB/s", flush=True)
    B.close()
    for P in plans:
        P.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--shots", type=int, default=8)
    ap.add_argument("--nt", type=int, default=2000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched-2D, uneven-3D and FWI-gradient measurements")
    ap.add_argument("--profile-kernel", action="store_true",
                    help="profiler target: stencil launches of the FP32 S-shot variant only")
    ap.add_argument("--profile-dtype", default="f32", choices=["f32", "f64"])
    ap.add_argument("--tiling", default="",
                    help="block_y,march_len,prefetch[,shot_group] for --profile-kernel")
    ap.add_argument("--time-block", default="",
                    help="block_rows,march_len,prefetch: --profile-kernel times two-step launches")
    ap.add_argument("--launches", type=int, default=20)
    ap.add_argument("--profile-batch2d", action="store_true",
                    help="profiler target: stencil launches of the 64-shot 1066^2 2D workload (FP32); "
                         "--tiling block_y,march_len,prefetch,shot_group, --time-block for two-step")
    ap.add_argument("--profile-fwi", action="store_true",
                    help="profiler target: one FP32 FWI gradient of the 554^2 32-shot workload; "
                         "--time-block rows,march_len,prefetch for two-step forward launches")
    ap.add_argument("--profile-uneven", action="store_true",
                    help="profiler target: batch stencil launches of the 12-model uneven 3D "
                         "workload (FP32); --time-block rows,march_len,prefetch for two-step "
                         "launches, --tiling block_y,march_len,prefetch for single steps")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from awcpml.host import setup
    import awcpml
    from cases import plain3d, rel_l2

    if args.profile_fwi:
        from awcpml import fwi
        from cases import fwi2d
        fc = fwi2d()
        FS = _setup(fc)
        P = _shot_plan(FS, fc["shots"], "f32")
        if args.time_block:  # two-step launches for both forward passes
            P.set_time_block(2, *(int(v) for v in args.time_block.split(",")))
        P.reset()
        P.advance(FS.nt)
        obs = [P.traces(q) for q in range(P.nshots)]
        P.set_model(np.asfortranarray((FS.dt * FS.dt) * (fc["start"] * fc["start"])), FS.inv_d, FS.inv_d2)
        t0 = time.perf_counter()
        J, g = fwi.misfit_gradient(P, obs, every=fc["every"])
        print(f"[profile] fwi2d FP32 gradient {time.perf_counter() - t0:.3f} s, misfit {J:.4e}", flush=True)
        P.close()
        return

    if args.profile_batch2d:
        profile_batch2d(args)
        return

    if args.profile_uneven:
        profile_uneven(args)
        return

    case = plain3d(nt=args.nt)
    srcs = awcpml.Sources(case["srcpos"], case["t0s"], [awcpml.rickersource1D], case["f0"])
    recs = awcpml.Receivers(case["recpos"])
    S = setup(case["l"], case["lt"], case["vel"], srcs, recs, halo=case["halo"],
              rcoef=case["rcoef"], ppw=case["ppw"], freetop=case["freetop"])
    print(f"[bench] plain3d n={S.n} nt={S.nt} halo={S.halo}", file=sys.stderr, flush=True)
    if args.profile_kernel:
        tiling = tuple(int(v) for v in args.tiling.split(",")) if args.tiling else None
        nsh = args.shots if args.profile_dtype == "f32" else 1
        tb = tuple(int(v) for v in args.time_block.split(",")) if args.time_block else None
        profile_kernel(S, args.profile_dtype, nsh, tiling, args.launches, tb)
        return

    v64 = time_variant(S, "f64", 1, args.steps, args.warmup, case["freetop"])
    print(f"[bench] fp64 plain: {v64['shot_gpts_per_s']:.2f} shot*Gpts/s", file=sys.stderr, flush=True)
    v32 = time_variant(S, "f32", 1, args.steps, args.warmup, case["freetop"])
    print(f"[bench] fp32 1 shot: {v32['shot_gpts_per_s']:.2f} shot*Gpts/s", file=sys.stderr, flush=True)
    vb = time_variant(S, "f32", args.shots, args.steps, args.warmup, case["freetop"])
    print(f"[bench] fp32 {args.shots} shots: {vb['shot_gpts_per_s']:.2f} shot*Gpts/s", file=sys.stderr,
          flush=True)
    print("[bench] plain2d (2D drop-in workload)", file=sys.stderr, flush=True)
    case2d, v2d = bench_plain2d()
    print(f"[bench] plain2d fp64: {v2d['shot_gpts_per_s']:.2f} shot*Gpts/s", file=sys.stderr, flush=True)

    main_v = vb
    kname = ("stencil3d_two<float>" if main_v["steps_per_launch"] == 2 else "stencil3d<float>")
    workload = f"plain3d_{S.n[0]}^3_nt{S.nt}_f32_S{args.shots}"
    achieved = main_v["kernel_gbs"]
    traffic = load_traffic(kname, workload)
    out = {
        "metric": METRIC,
        "value": round(main_v["shot_gpts_per_s"], 3),
        "unit": "shot·Gpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_v["s_per_solve"] * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"plain3d: 3D CPML forward solve, {S.n[0]}^3 nodes (256^3 interior + "
                        f"20-point halo), free top, nt={S.nt}, FP32, {args.shots} shots per launch",
            "nodes": list(S.n), "nt": S.nt, "halo": S.halo, "shots": args.shots,
            "effective_gbs": round(main_v["effective_gbs"], 1),
            "effective_frac_of_peak": round(main_v["effective_gbs"] / HBM_PEAK_GBS, 4),
            "fp64_plain": {k: (round(v, 4) if isinstance(v, float) else v)
                           for k, v in v64.items() if not k.startswith("_")},
            "fp32_single": {k: (round(v, 4) if isinstance(v, float) else v)
                            for k, v in v32.items() if not k.startswith("_")},
            "tiling": main_v["tiling"],
            "time_block": main_v["time_block"],
            "fp32_vs_fp64_plain": round(main_v["shot_gpts_per_s"] / v64["shot_gpts_per_s"], 3),
            "plain2d": {k: v for k, v in v2d.items() if not k.startswith("_")},
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": kname, "bytes_per_launch": main_v["bytes_per_launch"],
            "steps_per_launch": main_v["steps_per_launch"],
            "kernel_ms": round(main_v["kernel_ms"], 5),
        },
        "cpu_baseline": None,
    }
    if not args.no_extras:
        extras = {}
        for name, fn in (("batch2d", lambda: bench_batch2d(1)), ("uneven3d", bench_uneven3d),
                         ("fwi2d", bench_fwi2d)):
            print(f"[bench] extra: {name}", file=sys.stderr, flush=True)
            extras[name] = fn()
            print(f"[bench] {name}: {extras[name]}", file=sys.stderr, flush=True)
        out["config"]["extras"] = extras
    if not args.no_cpu:
        print("[bench] cpu baseline (oracle)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(
            case, {"f64": v64, "fp32_single": v32, f"fp32_{args.shots}shot_shot0": vb},
            case2d, v2d)
        # the FP32 results' errors against the FP64 oracle
        par = out["cpu_baseline"]["parity_vs_oracle"]
        out["config"]["fp32_trace_rel_l2_vs_oracle"] = par["fp32_single_trace_rel_l2"]
        out["config"]["fp32_field_rel_l2_vs_oracle"] = par["fp32_single_field_rel_l2"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
