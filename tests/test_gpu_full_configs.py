"""The benchmark workloads at their full sizes, checked against the oracle.

  config 2   plain3d: 298^3 nodes, free top, nt = 2000 (the 3D drop-in).
             FP64 through the autotuned launch shape with the time loop in
             a HIP graph, as bench.py runs it: traces and final field bit
             for bit the oracle's.  FP32 with 1 and with 8 shots per
             launch: traces within 1e-5 relative L2 of FP64; the final
             field's FP32 error is printed and bounded.
  config 3   batch2d: 1066^2, 64 shots, nt = 3000; shots 0 and 63 against
             the oracle (FP64 bitwise, FP32 within 1e-5).
  config 4   uneven3d: the 12-model batch (170^3 ... 554^3, nt = 1500) in
             FP32; the 170^3 and 234^3 models against the oracle, and
             those two models in an FP64 batch bit for bit.

The product's host setup (awcpml.host.setup) builds every plan, as in
bench.py; it equals the oracle's setup arrays (checked here first), so the
FP64 comparisons are exact.
"""
import numpy as np
import pytest

import awcpml
from awcpml.host import round_ties_up, setup
from cases import batch2d, lt_for_nt, oracle_problem, plain3d, rel_l2, uneven3d

pytestmark = pytest.mark.gpu

FP32_TRACE_TOL = 1e-5
FP32_FIELD_TOL = 1e-1   # final wavefield, FP32 vs FP64 (printed)


def _setup(case, vel=None, srcpos=None):
    vel = case["vel"] if vel is None else vel
    srcpos = case["srcpos"] if srcpos is None else srcpos
    srcs = awcpml.Sources(srcpos, case["t0s"] * len(srcpos),
                          [awcpml.rickersource1D] * len(srcpos), case["f0"])
    recs = awcpml.Receivers(case["recpos"])
    return setup(case["l"], case["lt"], vel, srcs, recs, halo=case["halo"], rcoef=case["rcoef"],
                 ppw=case["ppw"], freetop=case["freetop"])


def _same_setup(S, o):
    assert S.dt == o.dt and S.nt == o.nt
    assert np.array_equal(S.fact, o.fact())
    assert np.array_equal(S.dt2srctf, o.dt2srctf())
    for a in range(S.ndim):
        for w in range(8):
            assert np.array_equal(S.coef[a][w], o.coef(a, w)), (a, w)
    ps, pr = o.positions()
    assert np.array_equal(S.possrcs, ps) and np.array_equal(S.posrecs, pr)


def _index(S, pos):
    return np.array([[int(round_ties_up(pos[0, a] / S.d[a] + 1)) for a in range(S.ndim)]])


def _plan(S, dtype, src_idx):
    """Plan with one shot per entry of src_idx (1-based node positions)."""
    P = awcpml.Plan(S.n, S.halo, dtype=dtype, nshots=len(src_idx))
    P.set_model(S.fact, S.inv_d, S.inv_d2)
    for a in range(S.ndim):
        P.set_cpml(a, S.coef[a])
    for q, idx in enumerate(src_idx):
        P.set_sources(q, idx, S.dt2srctf)
        P.set_receivers(q, S.posrecs)
    return P


def _solve(P, nt):
    P.reset()
    P.advance(nt)


def test_config2_plain3d_full():
    case = plain3d()
    print("[gpu] config 2: plain3d 298^3 nt=2000, oracle run", flush=True)
    S = _setup(case)
    o = oracle_problem(case)
    _same_setup(S, o)
    o.run()
    tr_ref, p_ref = o.traces(), o.pend()
    o.close()
    assert S.nt == 2000 and np.abs(tr_ref).max() > 0

    # FP64 drop-in, autotuned and graph-captured as bench.py runs it
    P = _plan(S, "f64", [S.possrcs])
    best = awcpml.autotune(P, launches=3, cache=None)
    P.set_graph(True)
    _solve(P, S.nt)
    print(f"  f64 launch shape {best} time_block {P.time_block()}", flush=True)
    assert np.array_equal(P.traces(0), tr_ref)
    assert np.array_equal(P.field(0, 1), p_ref)
    # the other pick: one step per launch if the tuner chose two, and back
    tb = P.time_block()
    P.set_time_block(1 if tb[0] == 2 else 2, 8, 48, 2)
    _solve(P, S.nt)
    print(f"  f64 time_block {P.time_block()}", flush=True)
    assert np.array_equal(P.traces(0), tr_ref)
    assert np.array_equal(P.field(0, 1), p_ref)
    P.close()

    # FP32, one shot
    P = _plan(S, "f32", [S.possrcs])
    awcpml.autotune(P, launches=3, cache=None)
    P.set_graph(True)
    _solve(P, S.nt)
    tr32, p32 = P.traces(0), P.field(0, 1)
    P.close()
    e_tr, e_p = rel_l2(tr32, tr_ref), rel_l2(p32, p_ref)
    print(f"  f32 1 shot: trace rel_l2 {e_tr:.3e}, final field rel_l2 {e_p:.3e}", flush=True)
    assert e_tr <= FP32_TRACE_TOL
    assert e_p <= FP32_FIELD_TOL

    # 8 shots per launch, spread along x like bench.py; shot 0 is the
    # oracle's source.  FP64 8-shot plan: shot 0 bitwise the oracle, and it
    # is the FP64 reference of the other shots.
    shots = [S.possrcs]
    for s in range(1, 8):
        pos = S.possrcs.copy()
        pos[0, 0] = 1 + S.halo + 1 + (s * (S.n[0] - 2 * S.halo - 4)) // 8
        shots.append(pos)
    R = _plan(S, "f64", shots)
    R.set_time_block(2, 8, 48, 2)
    R.set_graph(True)
    _solve(R, S.nt)
    assert np.array_equal(R.traces(0), tr_ref)
    ref8 = [R.traces(q) for q in range(8)]
    R.close()
    P = _plan(S, "f32", shots)
    awcpml.autotune(P, launches=3, cache=None)
    P.set_graph(True)
    _solve(P, S.nt)
    errs = [rel_l2(P.traces(q), ref8[q]) for q in range(8)]
    print(f"  f32 8 shots: time_block {P.time_block()} trace rel_l2 max {max(errs):.3e}", flush=True)
    assert max(errs) <= FP32_TRACE_TOL
    # shots are independent: shot 0 equals the one-shot FP32 run bit for bit
    assert np.array_equal(P.traces(0), tr32)
    assert np.array_equal(P.field(0, 1), p32)
    P.close()


def test_config3_batch2d_full():
    case = batch2d()
    print("[gpu] config 3: batch2d 1066^2, 64 shots, nt=3000", flush=True)
    S = _setup(case)
    picks = (0, len(case["shots"]) - 1)
    refs = {}
    for q in picks:
        c = dict(case)
        c["srcpos"] = case["shots"][q]
        o = oracle_problem(c)
        if q == 0:
            _same_setup(_setup(case, srcpos=case["shots"][0]), o)
        o.run()
        refs[q] = o.traces()
        o.close()
    idx = [_index(S, pos) for pos in case["shots"]]
    R = _plan(S, "f64", [idx[q] for q in picks])
    R.set_graph(True)
    _solve(R, S.nt)
    for k, q in enumerate(picks):
        assert np.array_equal(R.traces(k), refs[q]), q
    R.close()
    P = _plan(S, "f32", idx)
    best = awcpml.autotune(P, launches=3, cache=None)
    P.set_graph(True)
    _solve(P, S.nt)
    for q in picks:
        err = rel_l2(P.traces(q), refs[q])
        print(f"  f32 shot {q} trace rel_l2 {err:.3e} (shape {best['block_y']},{best['march_len']},"
              f"{best['prefetch']},{best['shot_group']} tb {best['time_block']})", flush=True)
        assert err <= FP32_TRACE_TOL
    P.close()


def test_config4_uneven3d_full():
    models = uneven3d()
    print("[gpu] config 4: uneven3d 12 models, nt=1500", flush=True)
    cases, setups = [], []
    for c in models:
        vel = c["make_vel"]()
        c = dict(c)
        c["lt"] = lt_for_nt(3, (10.0,) * 3, float(vel.max()), c["nt"])
        setups.append(_setup(c, vel=vel))
        cases.append(c)
        del vel
    edges = [S.n[0] for S in setups]
    picks = [edges.index(170), edges.index(234)]
    refs = {}
    for m in picks:
        c = dict(cases[m])
        c["vel"] = c["make_vel"]()
        o = oracle_problem(c)
        _same_setup(setups[m], o)
        o.run()
        refs[m] = (o.traces(), o.pend())
        o.close()
    # FP64 batch of the two checked models: bitwise
    plans = [awcpml.Plan.from_setup(setups[m], dtype="f64") for m in picks]
    B = awcpml.Batch(plans)
    B.set_graph(True)
    B.run()
    for P, m in zip(plans, picks):
        assert np.array_equal(P.traces(0), refs[m][0]), edges[m]
        assert np.array_equal(P.field(0, 1), refs[m][1]), edges[m]
    B.close()
    for P in plans:
        P.close()
    # the full FP32 batch, autotuned as bench.py runs it
    plans = [awcpml.Plan.from_setup(S, dtype="f32") for S in setups]
    best = awcpml.autotune_batch(plans)
    B = awcpml.Batch(plans)
    B.set_graph(True)
    B.run()
    print(f"  f32 batch time_block {B.time_block()} shape {best['time_block']}", flush=True)
    for m in picks:
        e_tr = rel_l2(plans[m].traces(0), refs[m][0])
        e_p = rel_l2(plans[m].field(0, 1), refs[m][1])
        print(f"  f32 model {edges[m]}^3: trace rel_l2 {e_tr:.3e}, final field {e_p:.3e}", flush=True)
        assert e_tr <= FP32_TRACE_TOL
        assert e_p <= FP32_FIELD_TOL
    B.close()
    for P in plans:
        P.close()
