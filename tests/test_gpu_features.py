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
from cases import OVERLAP_CASES, oracle_problem, overlap_case, rel_l2, small2d, small3d

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


@pytest.mark.parametrize("two", [False, True])
def test_batch_follows_plan_reconfiguration(two):
    """A plan of a batch gets new sources (more of them, and another nt) and
    its traces are read, which replaces its device source/receiver buffers;
    the batch must notice: advancing without a reset raises, and after a
    reset the batch rebuilds its argument tables and matches the oracle."""
    cases = [small3d(n=(37, 29, 23), halo=5, nt=70, seed=1),
             small3d(n=(45, 41, 33), halo=6, nt=95, seed=2, freetop=False, nsrc=2)]
    print(f"[gpu] batch reconfiguration two={two}", flush=True)
    oracles = [oracle_for(c) for c in cases]
    plans = [plan_from_oracle(o) for o in oracles]
    if two:
        for P in plans:
            P.set_time_block(2, 4, 6, 2)
    B = awcpml.Batch(plans)
    B.set_graph(True)
    B.run()
    for P, o in zip(plans, oracles):
        o.run()
        assert np.array_equal(P.traces(), o.traces())
    # new acquisition for plan 0: three sources and a longer run
    c2 = small3d(n=(37, 29, 23), halo=5, nt=110, seed=1, nsrc=3)
    o2 = oracle_for(c2)
    o2.run()
    ps, pr = o2.positions()
    plans[0].set_sources(0, ps, o2.dt2srctf())
    plans[0].set_receivers(0, pr)
    plans[0].traces()  # uploads the new buffers, frees the old ones
    with pytest.raises(_lib.AwcpmlError, match="reconfigured"):
        B.advance(1)
    for _ in range(2):
        B.run()
        assert np.array_equal(plans[0].traces(), o2.traces())
        assert np.array_equal(plans[0].field(), o2.pend())
        assert np.array_equal(plans[1].traces(), oracles[1].traces())
    B.close()
    for P in plans:
        P.close()


@pytest.mark.parametrize("n,halo,freetop", OVERLAP_CASES)
def test_overlapping_slabs_bitwise(n, halo, freetop):
    """n < 2*halo+2 on some axes: the two CPML slabs overlap and the low
    side wins, as the reference's if/elseif (acoustic_2D_xPU.jl:45-56).
    Every single-step tiling and the two-step launches match the oracle."""
    case = overlap_case(n, halo, freetop=freetop)
    print(f"[gpu] overlapping slabs n={n} halo={halo} freetop={freetop}", flush=True)
    o = oracle_for(case)
    o.run()
    P = plan_from_oracle(o)
    for by, ml, pf in ((4, 7, 2), (1, 16, 3), (2, 40, 5), (1, 3, 1)):
        P.set_tiling(by, ml, pf)
        P.reset()
        P.advance(o.nt)
        assert np.array_equal(P.traces(), o.traces()), (by, ml, pf)
        assert np.array_equal(P.field(), o.pend()), (by, ml, pf)
    for rows, ml in ((4, 5), (8, 64)):
        P.set_time_block(2, rows, ml, 2)
        assert P.time_block()[0] == 2
        P.reset()
        P.advance(o.nt)
        assert np.array_equal(P.traces(), o.traces()), ("two", rows, ml)
        assert np.array_equal(P.field(), o.pend()), ("two", rows, ml)
    P.close()
    # FP32 against the FP64 oracle
    P = plan_from_oracle(o, dtype="f32")
    P.reset()
    P.advance(o.nt)
    assert rel_l2(P.traces(), o.traces()) <= 1e-5
    P.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_wide_offsets_bitwise(dim):
    """The 64-bit-offset kernels (taken automatically for fields of 4 GiB
    and more per shot) forced on small grids: bitwise the oracle, alone and
    in a batch; two-step launches fall back to single steps."""
    cases = ([small3d(freetop=True, nsrc=2, nt=80), small3d(n=(45, 41, 33), halo=6, nt=60, seed=2,
                                                           freetop=False)] if dim == 3 else
             [small2d(nt=150), small2d(n=(61, 47), halo=7, nt=120, freetop=False, seed=9)])
    print(f"[gpu] wide offsets {dim}D", flush=True)
    oracles = [oracle_for(c) for c in cases]
    for o in oracles:
        o.run()
    P = plan_from_oracle(oracles[0])
    P.set_wide_offsets(True)
    P.set_time_block(2, 4, 8, 2)
    assert P.time_block()[0] == 1
    for pf in (1, 2, 3):
        P.set_tiling(2, 9, pf)
        P.reset()
        P.advance(oracles[0].nt)
        assert np.array_equal(P.traces(), oracles[0].traces()), pf
        assert np.array_equal(P.field(), oracles[0].pend()), pf
    P.set_tiling(4, 0, 2)
    plans = [P] + [plan_from_oracle(o) for o in oracles[1:]]
    for Q in plans[1:]:
        Q.set_tiling(4, 0, 2)
    B = awcpml.Batch(plans)
    B.run()
    for Q, o in zip(plans, oracles):
        assert np.array_equal(Q.traces(), o.traces())
        assert np.array_equal(Q.field(), o.pend())
    B.close()
    for Q in plans:
        Q.close()


def test_many_sources_fall_back_to_single_steps():
    """A shot with more sources than a two-step launch stages (32) takes
    single steps, bitwise; 32 sources still take two-step launches."""
    for nsrc, steps in ((33, 1), (32, 2)):
        case = small3d(n=(41, 37, 35), halo=5, nt=60, freetop=False)
        rng = np.random.Generator(np.random.PCG64(nsrc))
        pts = [[(6 + int(rng.integers(0, 29))) * 10.0, (6 + int(rng.integers(0, 25))) * 10.0,
                (6 + int(rng.integers(0, 23))) * 10.0] for _ in range(nsrc)]
        pts[1] = pts[0]  # two sources on one node, added in order
        case["srcpos"] = np.array(pts)
        case["t0s"] = [0.02 + 0.001 * s for s in range(nsrc)]
        print(f"[gpu] {nsrc} sources per shot", flush=True)
        o = oracle_for(case)
        o.run()
        P = plan_from_oracle(o)
        P.set_time_block(2, 4, 10, 2)
        P.reset()
        P.advance(o.nt)
        assert P.time_block()[0] == steps  # known once the sources are uploaded
        assert np.array_equal(P.traces(), o.traces()), nsrc
        assert np.array_equal(P.field(), o.pend()), nsrc
        P.close()
