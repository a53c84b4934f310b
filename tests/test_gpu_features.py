"""GPU tests of the extensions around the drop-in path: multi-shot plans,
uneven batches in one launch per step, hipGraph-captured time loops, the
autotuner, and the ABI's error behaviour.

Every FP64 result is compared bit for bit with an independent run (the
single-shot plan, or the CPU oracle); FP32 traces within 1e-5 relative L2.
"""
import numpy as np
import pytest

import awcpml
from awcpml import _lib
from awcpml.autotune import two_step_batch_candidates
from cases import oracle_problem, rel_l2, small2d, small3d

pytestmark = pytest.mark.gpu


def plan_from_oracle(o, dtype="f64", nshots=1, pos_list=None):
    P = awcpml.Plan(o.n, o.halo, dtype=dtype, nshots=nshots)
    sp = o._spacing
    P.set_model(o.fact(), [1 / d for d in sp], [1 / (d * d) for d in sp])
    for a in range(o.ndim):
        P.set_cpml(a, [o.coef(a, w) for w in range(8)])
    ps, pr = o.positions()
    for s in range(nshots):
        P.set_sources(s, ps if pos_list is None else pos_list[s], o.dt2srctf())
        P.set_receivers(s, pr)
    return P


def oracle_for(case):
    o = oracle_problem(case)
    o._spacing = [case["l"][a] / (case["n"][a] - 1) for a in range(len(case["n"]))]
    return o


def test_graph_replays_match_eager():
    case = small3d(freetop=True, nt=90)
    print("[gpu] graph replay small3d", flush=True)
    o = oracle_for(case)
    o.run()
    P = plan_from_oracle(o)
    P.reset()
    P.advance(o.nt)
    eager_tr, eager_p = P.traces(), P.field()
    P.set_graph(True)
    for rep in range(3):
        P.reset()
        P.advance(o.nt)  # first: capture + launch; then replays
        assert np.array_equal(P.traces(), eager_tr), rep
        assert np.array_equal(P.field(), eager_p), rep
        # eager work between replays
        P.set_graph(False)
        P.reset()
        P.advance(5)
        P.set_graph(True)
    # split runs replay per (start, length)
    P.reset()
    P.advance(40)
    P.advance(o.nt - 40)
    assert np.array_equal(P.traces(), o.traces())
    assert np.array_equal(P.field(), o.pend())
    P.close()


def test_multishot_plan_equals_single_shots():
    case = small3d(freetop=False, nt=80)
    print("[gpu] multishot small3d", flush=True)
    o = oracle_for(case)
    ps, _ = o.positions()
    shifts = [0, 3, -4]
    pos = []
    for d in shifts:
        q = ps.copy()
        q[:, 0] += d
        pos.append(q)
    M = plan_from_oracle(o, nshots=3, pos_list=pos)
    M.reset()
    M.advance(o.nt)
    for s, q in enumerate(pos):
        S = plan_from_oracle(o, nshots=1, pos_list=[q])
        S.reset()
        S.advance(o.nt)
        assert np.array_equal(M.traces(s), S.traces()), s
        assert np.array_equal(M.field(s), S.field()), s
        S.close()
    M.close()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_uneven_batch_equals_separate_plans(dtype):
    cases = [small3d(n=(37, 29, 23), halo=5, nt=70, seed=1),
             small3d(n=(45, 41, 33), halo=6, nt=95, seed=2, freetop=False, nsrc=2),
             small3d(n=(29, 31, 27), halo=4, nt=50, seed=3)]
    print(f"[gpu] uneven batch {dtype}", flush=True)
    oracles = [oracle_for(c) for c in cases]
    plans = [plan_from_oracle(o, dtype=dtype) for o in oracles]
    for graph in (False, True):
        B = awcpml.Batch(plans)
        B.set_graph(graph)
        for _ in range(2):
            B.run()
            for P, o in zip(plans, oracles):
                ref = plan_from_oracle(o, dtype=dtype)
                ref.reset()
                ref.advance(o.nt)
                assert np.array_equal(P.traces(), ref.traces())
                assert np.array_equal(P.field(), ref.field())
                ref.close()
        B.close()
    for o in oracles:
        o.run()
    for P, o in zip(plans, oracles):
        if dtype == "f64":
            assert np.array_equal(P.traces(), o.traces())
        else:
            assert rel_l2(P.traces(), o.traces()) <= 1e-5
        P.close()


def test_batch_2d_with_multishot_plan():
    cases = [small2d(nt=150), small2d(n=(61, 47), halo=7, nt=120, freetop=False, seed=9)]
    print("[gpu] batch 2d", flush=True)
    oracles = [oracle_for(c) for c in cases]
    plans = [plan_from_oracle(oracles[0], nshots=2), plan_from_oracle(oracles[1])]
    B = awcpml.Batch(plans)
    B.run()
    for o in oracles:
        o.run()
    assert np.array_equal(plans[0].traces(0), oracles[0].traces())
    assert np.array_equal(plans[0].traces(1), oracles[0].traces())
    assert np.array_equal(plans[1].traces(0), oracles[1].traces())
    B.close()
    for P in plans:
        P.close()


def test_autotune_keeps_results(tmp_path):
    case = small3d(n=(61, 53, 47), halo=8, nt=60)
    print("[gpu] autotune", flush=True)
    o = oracle_for(case)
    o.run()
    P = plan_from_oracle(o)
    cache = str(tmp_path / "tune.json")
    best = awcpml.autotune(P, launches=3, cache=cache)
    assert not best["cached"] and best["tried"] > 10
    assert P.tiling() == (best["block_y"], best["march_len"], best["prefetch"])
    P.reset()
    P.advance(o.nt)
    assert np.array_equal(P.traces(), o.traces())
    again = awcpml.autotune(P, cache=cache)
    assert again["cached"]
    P.close()


def test_every_tiling_is_exact():
    """All launch shapes give the same bits (the autotuner may pick any)."""
    case = small3d(n=(41, 37, 35), halo=6, nt=40, freetop=False, nsrc=2)
    print("[gpu] tilings", flush=True)
    o = oracle_for(case)
    o.run()
    P = plan_from_oracle(o)
    for by in (1, 2, 4):
        for ml in (1, 3, 7, 33):
            for pf in (1, 2, 3, 4, 5):
                P.set_tiling(by, ml, pf)
                P.reset()
                P.advance(o.nt)
                assert np.array_equal(P.field(), o.pend()), (by, ml, pf)
    P.close()


def test_forward_and_errors():
    case = small2d(nt=30)
    print("[gpu] forward/errors", flush=True)
    o = oracle_for(case)
    P = plan_from_oracle(o)
    P.reset()
    P.advance(o.nt)
    with pytest.raises(_lib.AwcpmlError, match="past nt"):
        P.advance(1)
    with pytest.raises(_lib.AwcpmlError, match="not inside"):
        P.set_receivers(0, [[1, 10_000]])
    with pytest.raises(_lib.AwcpmlError, match="outside 1..nt"):
        P.forward(0)
    P.forward(1)  # one forward! at it = 1, like the reference's do_bench
    assert P.steps_done() == o.nt
    P.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_shot_groups_are_exact(dim):
    """Every (block_y, shot_group) shape gives the bits of the single-shot
    plans, including shot counts that the group does not divide."""
    case = small3d(n=(41, 37, 35), halo=6, nt=40, freetop=False) if dim == 3 else \
        small2d(n=(61, 47), halo=7, nt=60, freetop=True)
    print(f"[gpu] shot groups {dim}D", flush=True)
    o = oracle_for(case)
    ps, _ = o.positions()
    pos = []
    for d in (0, 3, -4, 5, -2):
        q = ps.copy()
        q[:, 0] += d
        pos.append(q)
    refs = []
    for q in pos:
        S = plan_from_oracle(o, nshots=1, pos_list=[q])
        S.reset()
        S.advance(o.nt)
        refs.append((S.traces(), S.field()))
        S.close()
    M = plan_from_oracle(o, nshots=len(pos), pos_list=pos)
    for by, g in ((1, 1), (2, 2), (4, 2), (4, 4), (8, 1), (8, 4), (8, 8)):
        for ml, pf in ((5, 2), (16, 3), (6, 4), (13, 5)):
            M.set_tiling(by, ml, pf)
            M.set_shot_group(g)
            assert M.shot_group() == g
            M.reset()
            M.advance(o.nt)
            for s, (tr, fld) in enumerate(refs):
                assert np.array_equal(M.traces(s), tr), (by, g, ml, pf, s)
                assert np.array_equal(M.field(s), fld), (by, g, ml, pf, s)
    with pytest.raises(_lib.AwcpmlError, match="does not divide"):
        M.set_tiling(2, 8, 2)
        M.set_shot_group(4)
    M.close()


def test_autotune_batch_keeps_results():
    cases = [small3d(n=(37, 29, 23), halo=5, nt=40, seed=1),
             small3d(n=(45, 41, 33), halo=6, nt=40, seed=2, freetop=False)]
    print("[gpu] autotune batch", flush=True)
    oracles = [oracle_for(c) for c in cases]
    plans = [plan_from_oracle(o) for o in oracles]
    best = awcpml.autotune_batch(plans, launches=1)
    assert best["tried"] == 18 + len(two_step_batch_candidates(plans))
    assert all(P.tiling() == (best["block_y"], best["march_len"], best["prefetch"]) for P in plans)
    B = awcpml.Batch(plans)
    B.run()
    for P, o in zip(plans, oracles):
        o.run()
        assert np.array_equal(P.traces(), o.traces())
        P.close()
    B.close()
