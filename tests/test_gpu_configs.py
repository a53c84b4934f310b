"""Reduced versions of the extension workloads, checked against the oracle.

  batched 2D   many shots on one layered model in one launch per step
  uneven 3D    several models of different sizes in one batch launch
  FWI 2D       the misfit gradient through the GPU adjoint

Same construction as the full workloads (tests/cases.py) at sizes the
oracle finishes in seconds.  FP64 results are bitwise the oracle's forward
(traces) or within 1e-10 of its adjoint (gradient); FP32 traces are within
1e-5 and the FP32 gradient within 1e-3 relative L2 of the FP64 oracle.
"""
import numpy as np
import pytest

import awcpml
from awcpml import fwi
from awcpml.host import round_ties_up, setup
from cases import batch2d, fwi2d, lt_for_nt, rel_l2, uneven3d

pytestmark = pytest.mark.gpu


def _setup(case, vel=None):
    vel = case["vel"] if vel is None else vel
    srcs = awcpml.Sources(case["srcpos"], case["t0s"], [awcpml.rickersource1D], case["f0"])
    recs = awcpml.Receivers(case["recpos"])
    return setup(case["l"], case["lt"], vel, srcs, recs, halo=case["halo"], rcoef=case["rcoef"],
                 ppw=case["ppw"], freetop=case["freetop"])


def _index(S, pos):
    return np.array([[int(round_ties_up(pos[0, a] / S.d[a] + 1)) for a in range(S.ndim)]])


def _plan(S, shots, dtype, fact=None):
    P = awcpml.Plan(S.n, S.halo, dtype=dtype, nshots=len(shots))
    P.set_model(S.fact if fact is None else fact, S.inv_d, S.inv_d2)
    for a in range(S.ndim):
        P.set_cpml(a, S.coef[a])
    for q, pos in enumerate(shots):
        P.set_sources(q, _index(S, pos), S.dt2srctf)
        P.set_receivers(q, S.posrecs)
    return P


def _oracle_traces(case, S, pos, vel=None):
    """Oracle traces of one shot at `pos` (same setup path as the oracle tests)."""
    from cases import oracle_problem
    c = dict(case)
    c["srcpos"] = pos
    if vel is not None:
        c["vel"] = vel
    o = oracle_problem(c)
    o.run()
    tr = o.traces()
    o.close()
    return tr


def test_batched_2d_shots_match_oracle():
    case = batch2d(nt=240, interior=160, nshots=8, nrec=24, seed=5)
    print("[gpu] batched 2D", flush=True)
    S = _setup(case)
    for dtype, tol in (("f64", 0.0), ("f32", 1e-5)):
        P = _plan(S, case["shots"], dtype)
        awcpml.autotune(P, launches=2, cache=None)
        P.set_graph(True)
        P.reset()
        P.advance(S.nt)
        for q in (0, 3, 7):
            ref = _oracle_traces(case, S, case["shots"][q])
            got = P.traces(q)
            if tol == 0.0:
                assert np.array_equal(got, ref), (dtype, q)
            else:
                err = rel_l2(got, ref)
                print(f"  {dtype} shot {q} trace rel_l2 {err:.2e}", flush=True)
                assert err <= tol, (dtype, q, err)
        P.close()


def test_uneven_3d_batch_matches_oracle():
    models = uneven3d(nt=50, count=4, seed=4, edges=(14, 20, 26))
    print("[gpu] uneven 3D batch", flush=True)
    cases, setups = [], []
    for c in models:
        vel = c["make_vel"]()
        c = dict(c)
        c["vel"] = vel
        c["lt"] = lt_for_nt(3, (10.0,) * 3, float(vel.max()), c["nt"])
        cases.append(c)
        setups.append(_setup(c))
    for dtype in ("f64", "f32"):
        plans = [awcpml.Plan.from_setup(S, dtype=dtype) for S in setups]
        B = awcpml.Batch(plans)
        B.set_graph(True)
        B.run()
        for c, P in zip(cases, plans):
            ref = _oracle_traces(c, None, c["srcpos"])
            if dtype == "f64":
                assert np.array_equal(P.traces(0), ref)
            else:
                err = rel_l2(P.traces(0), ref)
                print(f"  f32 model {c['n'][0]} trace rel_l2 {err:.2e}", flush=True)
                assert err <= 1e-5
        B.close()
        for P in plans:
            P.close()


def test_fwi_gradient_matches_oracle_adjoint():
    case = fwi2d(nt=180, interior=72, nshots=3, nrec=16, every=13)
    print("[gpu] FWI 2D gradient", flush=True)
    S = _setup(case)
    from cases import oracle_problem
    fact0 = np.asfortranarray((S.dt * S.dt) * (case["start"] * case["start"]))
    # observed data from the true model (FP64 plan, bitwise the oracle's forward)
    T = _plan(S, case["shots"], "f64")
    T.reset()
    T.advance(S.nt)
    obs = [T.traces(q) for q in range(T.nshots)]
    T.close()
    J_ref, g_ref = 0.0, np.zeros(S.n)
    for q, pos in enumerate(case["shots"]):
        c = dict(case)
        c["srcpos"] = pos
        o = oracle_problem(c)
        o.set_fact(fact0)
        J, g = o.fwi_gradient(obs[q])
        J_ref += J
        g_ref += g
        o.close()
    for dtype, tol in (("f64", 1e-10), ("f32", 1e-3)):
        P = _plan(S, case["shots"], dtype, fact=fact0)
        J, g = fwi.misfit_gradient(P, obs, every=case["every"])
        err = rel_l2(g, g_ref)
        print(f"  {dtype} J rel {abs(J - J_ref) / J_ref:.2e} grad rel_l2 {err:.2e}", flush=True)
        assert abs(J - J_ref) <= (1e-12 if dtype == "f64" else 1e-3) * J_ref
        assert err <= tol
        P.close()
