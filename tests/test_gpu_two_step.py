"""GPU tests of two time steps per stencil launch (awcpml_set_time_block),
3D (block rows exchanged through LDS) and 2D (independent waves).

A two-step launch evaluates each step with the single-step kernel's
arithmetic, so every result must equal one step per launch bit for bit,
in FP64 and FP32: traces, and the three field buffers afterwards.  The
FP64 results are also compared with the CPU oracle bit for bit.  Cases
cover several x tiles (62 owned nodes per wave, or aligned 64-node tiles
where no CPML slab reaches a block) and y tiles (rows-2 owned rows per
block), chunk lengths that leave short chunks, odd step counts
(the last step runs alone), several shots, sources on different planes and
two sources on one node, graphs, and the fallback to single steps for a
source on a Dirichlet node.
"""
import numpy as np
import pytest

import awcpml
from cases import oracle_problem, small2d, small3d

pytestmark = pytest.mark.gpu


def _oracle(case):
    o = oracle_problem(case)
    o._spacing = [case["l"][a] / (case["n"][a] - 1) for a in range(len(case["n"]))]
    return o


def _plan(o, dtype="f64", nshots=1, srcs=None, wav=None):
    P = awcpml.Plan(o.n, o.halo, dtype=dtype, nshots=nshots)
    sp = o._spacing
    P.set_model(o.fact(), [1 / d for d in sp], [1 / (d * d) for d in sp])
    for a in range(o.ndim):
        P.set_cpml(a, [o.coef(a, w) for w in range(8)])
    ps, pr = o.positions()
    for s in range(nshots):
        P.set_sources(s, ps if srcs is None else srcs[s], o.dt2srctf() if wav is None else wav)
        P.set_receivers(s, pr)
    return P


def _state(P, nshots):
    return [(P.traces(s), P.field(s, 0), P.field(s, 1), P.field(s, 2)) for s in range(nshots)]


def _run(P, nt, steps, rows=0, march=0, splits=None):
    P.set_time_block(steps, rows, march)
    P.reset()
    if splits:
        done = 0
        for n in splits:
            P.advance(n)
            done += n
        P.advance(nt - done)
    else:
        P.advance(nt)


def _same(a, b):
    for x, y in zip(a, b):
        for u, v in zip(x, y):
            assert np.array_equal(u, v)


CASES = {
    "freetop": dict(n=(37, 29, 23), halo=5, freetop=True, nt=61),
    "open_2src": dict(n=(37, 29, 23), halo=5, freetop=False, nt=57, nsrc=2),
    "xtiles": dict(n=(150, 33, 26), halo=6, freetop=True, nt=45),
    "ytiles": dict(n=(40, 71, 24), halo=4, freetop=False, nt=39),
    # x tiles, row groups and march chunks both with and without CPML nodes:
    # every block class of a launch is non-empty
    "classes": dict(n=(197, 64, 61), halo=5, freetop=False, nt=31),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("rows,march", [(8, 0), (4, 5), (16, 7)])
def test_two_step_fp64_bitwise(name, rows, march):
    print("[gpu] two-step fp64 %s rows %d march %d" % (name, rows, march), flush=True)
    c = CASES[name]
    case = small3d(n=c["n"], halo=c["halo"], freetop=c["freetop"], nt=c["nt"],
                   nsrc=c.get("nsrc", 1))
    o = _oracle(case)
    o.run()
    P = _plan(o)
    _run(P, o.nt, 1)
    one = _state(P, 1)
    _run(P, o.nt, 2, rows, march)
    assert P.time_block()[0] == 2
    two = _state(P, 1)
    _same(two, one)
    assert np.array_equal(two[0][0], o.traces())
    assert np.array_equal(two[0][2], o.pend())
    P.close()


@pytest.mark.parametrize("rows", [8, 16])
def test_two_step_fp32_bitwise(rows):
    print("[gpu] two-step fp32 rows %d" % rows, flush=True)
    case = small3d(n=(150, 40, 29), halo=6, freetop=True, nt=70, nsrc=2)
    o = _oracle(case)
    P = _plan(o, dtype="f32")
    _run(P, o.nt, 1)
    one = _state(P, 1)
    _run(P, o.nt, 2, rows, 0)
    two = _state(P, 1)
    _same(two, one)
    P.close()


# Aligned 64-node x tiles (awcpml_set_two_step_align) for the blocks no CPML
# slab reaches: several tiles, tile counts that differ between the two
# tilings (190: 4 tiles of 62, 3 of 64), and a grid whose tile-edge node
# x = 191 lies in the x slab (the launch keeps 62-node tiles).  Sources and
# receivers sit on the nodes just outside aligned tiles (0-based x = 63,
# 64, 127, 128), which the halo-row waves evaluate for lanes 0 and 63.
ALIGN_CASES = {
    "multi": dict(n=(180, 61, 40), halo=6, freetop=False, nt=41),
    "tiles_differ": dict(n=(190, 45, 30), halo=5, freetop=True, nt=37),
    "xslab_fallback": dict(n=(197, 40, 30), halo=6, freetop=False, nt=29),
}


@pytest.mark.parametrize("name", list(ALIGN_CASES))
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_two_step_aligned_tiles(name, dtype):
    print("[gpu] two-step aligned tiles %s %s" % (name, dtype), flush=True)
    c = ALIGN_CASES[name]
    case = small3d(n=c["n"], halo=c["halo"], freetop=c["freetop"], nt=c["nt"])
    o = _oracle(case)
    # 1-based positions on and next to the aligned tiles' edge nodes, in
    # rows and planes without CPML
    ym, zm = c["n"][1] // 2, c["n"][2] // 2
    src = np.array([[64, ym, zm], [65, ym + 1, zm - 1], [128, ym - 2, zm + 1], [129, ym, zm]])
    w = o.dt2srctf()
    wav = np.asfortranarray(np.concatenate([w * (1.0 + 0.1 * e) for e in range(len(src))], axis=1))
    P = _plan(o, dtype=dtype, srcs=[src], wav=wav)
    rec = np.array([[64, ym, zm], [65, ym, zm], [63, ym + 3, zm], [128, ym, zm + 2], [129, ym, zm],
                    [2, ym, zm], [c["n"][0] - 1, ym, zm]])
    P.set_receivers(0, rec)
    assert P.two_step_align() == (dtype == "f32")  # the default per precision
    _run(P, o.nt, 1)
    one = _state(P, 1)
    for rows, march in ((16, 0), (8, 7), (4, 11)):
        for al in (True, False):
            P.set_two_step_align(al)
            _run(P, o.nt, 2, rows, march)
            assert P.time_block()[0] == 2
            _same(_state(P, 1), one)
    P.set_two_step_align(True)
    P.close()


def test_two_step_shots_graphs_splits():
    print("[gpu] two-step shots + graphs + split advances", flush=True)
    case = small3d(n=(70, 35, 27), halo=5, freetop=False, nt=64)
    o = _oracle(case)
    ps, _ = o.positions()
    srcs = []
    for d in (0, 4, -5):
        q = ps.copy()
        q[:, 0] += d
        srcs.append(q)
    P = _plan(o, nshots=3, srcs=srcs)
    _run(P, o.nt, 1)
    one = _state(P, 3)
    _run(P, o.nt, 2, 8, 6, splits=[5, 20])  # odd first chunk: single steps interleave
    _same(_state(P, 3), one)
    P.set_graph(True)
    for rep in range(3):
        _run(P, o.nt, 2, 8, 6)
        _same(_state(P, 3), one)
        _run(P, o.nt, 2, 8, 6, splits=[7])
        _same(_state(P, 3), one)
    P.set_graph(False)
    P.close()


def test_two_step_sources_on_one_node():
    print("[gpu] two-step two sources on one node", flush=True)
    case = small3d(n=(37, 29, 23), halo=5, freetop=True, nt=50)
    o = _oracle(case)
    ps, _ = o.positions()
    w = o.dt2srctf()
    wav = np.asfortranarray(np.concatenate([w, 0.37 * w[::-1]], axis=1))
    P = _plan(o, srcs=[np.concatenate([ps, ps], axis=0)], wav=wav)
    _run(P, o.nt, 1)
    one = _state(P, 1)
    _run(P, o.nt, 2, 8, 0)
    assert P.time_block()[0] == 2
    _same(_state(P, 1), one)
    P.close()


def test_two_step_falls_back_for_dirichlet_source():
    print("[gpu] two-step fallback: source on a Dirichlet node", flush=True)
    case = small3d(n=(37, 29, 23), halo=5, freetop=True, nt=40)
    o = _oracle(case)
    ps, _ = o.positions()
    q = ps.copy()
    q[:, 2] = 1  # 1-based z index 1: the top boundary plane
    P = _plan(o, srcs=[q])
    _run(P, o.nt, 1)
    one = _state(P, 1)
    _run(P, o.nt, 2, 8, 0)
    assert P.time_block()[0] == 1
    _same(_state(P, 1), one)
    P.close()


def test_time_stencil2():
    print("[gpu] time_stencil2", flush=True)
    case = small3d(n=(70, 35, 27), halo=5, freetop=True, nt=20)
    o = _oracle(case)
    P = _plan(o, dtype="f32")
    P.set_time_block(2)
    P.reset()
    assert P.time_stencil2(3) > 0.0
    P.reset()
    P.advance(o.nt)
    P.close()


def _interior_2d(n=(57, 43), halo=6, freetop=True, nt=61):
    """small2d without its source on the top row (a Dirichlet node)."""
    case = small2d(n=n, halo=halo, freetop=freetop, nt=nt)
    case["srcpos"] = case["srcpos"][:1]
    case["t0s"] = case["t0s"][:1]
    return case


CASES_2D = {
    "freetop": dict(n=(57, 43), halo=6, freetop=True, nt=61),
    "open": dict(n=(57, 43), halo=6, freetop=False, nt=58),
    "xtiles": dict(n=(150, 47), halo=5, freetop=True, nt=45),
}


@pytest.mark.parametrize("name", list(CASES_2D))
@pytest.mark.parametrize("rows,march", [(4, 0), (8, 5), (4, 13)])
def test_two_step_2d_fp64_bitwise(name, rows, march):
    print("[gpu] two-step 2D fp64 %s rows %d march %d" % (name, rows, march), flush=True)
    o = _oracle(_interior_2d(**CASES_2D[name]))
    o.run()
    P = _plan(o)
    _run(P, o.nt, 1)
    one = _state(P, 1)
    _run(P, o.nt, 2, rows, march)
    assert P.time_block()[0] == 2
    two = _state(P, 1)
    _same(two, one)
    assert np.array_equal(two[0][0], o.traces())
    assert np.array_equal(two[0][2], o.pend())
    P.close()


# 2D aligned tiles: lanes 0 and 63 of every wave evaluate their row's
# tile-edge nodes; sources and receivers on those nodes (0-based x = 63,
# 64, 127, 128), a grid with an x slab over a tile edge (x = 191: the launch
# keeps 62-node tiles) and tile counts that differ between the tilings.
ALIGN_CASES_2D = {
    "multi": dict(n=(180, 71), halo=6, freetop=False, nt=45),
    "tiles_differ": dict(n=(190, 57), halo=5, freetop=True, nt=41),
    "xslab_fallback": dict(n=(197, 50), halo=6, freetop=False, nt=33),
}


@pytest.mark.parametrize("name", list(ALIGN_CASES_2D))
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_two_step_2d_aligned_tiles(name, dtype):
    print("[gpu] two-step 2D aligned tiles %s %s" % (name, dtype), flush=True)
    o = _oracle(_interior_2d(**ALIGN_CASES_2D[name]))
    n = ALIGN_CASES_2D[name]["n"]
    ym = n[1] // 2
    src = np.array([[64, ym], [65, ym + 1], [128, ym - 2], [129, ym]])
    w = o.dt2srctf()
    wav = np.asfortranarray(np.concatenate([w * (1.0 + 0.1 * e) for e in range(len(src))], axis=1))
    P = _plan(o, dtype=dtype, srcs=[src], wav=wav)
    P.set_receivers(0, np.array([[64, ym], [65, ym + 3], [128, ym], [129, ym - 1], [2, ym],
                                 [n[0] - 1, ym]]))
    assert P.two_step_align() == (dtype == "f32")
    _run(P, o.nt, 1)
    one = _state(P, 1)
    for rows, march in ((4, 0), (8, 5), (4, 13)):
        for al in (True, False):
            P.set_two_step_align(al)
            _run(P, o.nt, 2, rows, march)
            assert P.time_block()[0] == 2
            _same(_state(P, 1), one)
    P.close()


def test_two_step_2d_fp32_shots_graphs():
    print("[gpu] two-step 2D fp32, 3 shots, graphs", flush=True)
    o = _oracle(_interior_2d(n=(150, 61), halo=6, freetop=False, nt=73))
    ps, _ = o.positions()
    srcs = []
    for d in (0, 7, -9):
        q = ps.copy()
        q[:, 0] += d
        srcs.append(q)
    P = _plan(o, dtype="f32", nshots=3, srcs=srcs)
    _run(P, o.nt, 1)
    one = _state(P, 3)
    for pf in (1, 2, 3):
        P.set_time_block(2, 4, 9, pf)
        P.reset()
        P.advance(6)  # the split leaves an odd remainder: a single step at the end
        P.advance(o.nt - 6)
        _same(_state(P, 3), one)
    P.set_graph(True)
    for rep in range(2):
        _run(P, o.nt, 2, 8, 0)
        _same(_state(P, 3), one)
    P.set_graph(False)
    P.close()


def test_two_step_2d_falls_back_for_dirichlet_source():
    print("[gpu] two-step 2D fallback: source on the top row", flush=True)
    o = _oracle(small2d(nt=40))  # its second source lies on the top row
    o.run()
    P = _plan(o)
    _run(P, o.nt, 2, 4, 0)
    assert P.time_block()[0] == 1
    assert np.array_equal(P.traces(), o.traces())
    assert np.array_equal(P.field(), o.pend())
    P.close()


def test_two_step_fwi_gradient_unchanged():
    print("[gpu] FWI gradient with two-step forward launches", flush=True)
    from awcpml import fwi
    o = _oracle(_interior_2d(n=(57, 43), halo=6, freetop=True, nt=60))
    o.run()
    obs = o.traces() * 0.5
    P = _plan(o)
    P.set_time_block(1)
    J1, g1 = fwi.misfit_gradient(P, [obs], every=7)
    P.set_time_block(2, 4, 9)
    J2, g2 = fwi.misfit_gradient(P, [obs], every=7)
    assert J1 == J2
    assert np.array_equal(g1, g2)
    P.close()


# ---------------------------------------------------------------------
# Batches of plans, two steps of every plan per launch
# (awcpml_batch_set_time_block): bit-identical to single-step batches.

BATCH_3D = [dict(n=(37, 29, 23), halo=5, nt=70, seed=1),
            dict(n=(45, 41, 33), halo=6, nt=95, seed=2, freetop=False, nsrc=2),
            dict(n=(29, 31, 27), halo=4, nt=50, seed=3)]


def _batch_state(plans):
    out = []
    for P in plans:
        out += _state(P, P.nshots)
    return out


def _batch_run(B, nt, splits=None):
    B.reset()
    done = 0
    for n in (splits or ()):
        B.advance(n)
        done += n
    B.advance(nt - done)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_batch_two_step_bitwise(dtype):
    print("[gpu] batch two-step 3D %s" % dtype, flush=True)
    oracles = [_oracle(small3d(**c)) for c in BATCH_3D]
    plans = [_plan(o, dtype=dtype) for o in oracles]
    nt = max(o.nt for o in oracles)
    B = awcpml.Batch(plans)
    assert B.time_block()[0] == 1
    B.run()
    one = _batch_state(plans)
    B.close()
    # each plan keeps its own chunk length
    for P, m in zip(plans, (0, 5, 7)):
        P.set_time_block(1, 0, m)
    for rows, pf in ((16, 2), (8, 1), (4, 3)):
        print("[gpu]   rows %d prefetch %d" % (rows, pf), flush=True)
        for P in plans:
            P.reset()  # a batch run may leave the buffers in other roles
        B = awcpml.Batch(plans)
        B.set_time_block(2, rows, pf)
        assert B.time_block() == (2, rows, pf)
        for graph in (False, True):
            B.set_graph(graph)
            for splits in (None, (3, 8), (1,)):
                _batch_run(B, nt, splits)
                _same(_batch_state(plans), one)
        B.close()
    if dtype == "f64":
        for P, o in zip(plans, oracles):
            o.run()
            assert np.array_equal(P.traces(), o.traces())
            assert np.array_equal(P.field(), o.pend())
    for P in plans:
        P.close()


def test_batch_of_plans_in_traded_roles_is_refused():
    print("[gpu] a plan whose buffers traded roles must be reset before batching", flush=True)
    oracles = [_oracle(small3d(**c)) for c in BATCH_3D[:2]]
    plans = [_plan(o, dtype="f32") for o in oracles]
    B = awcpml.Batch(plans)
    B.set_time_block(2, 8, 2)
    B.reset()
    B.advance(2)  # one two-step launch: the spares now hold step 2
    B.close()
    with pytest.raises(awcpml._lib.AwcpmlError, match="reset it before batching"):
        awcpml.Batch(plans)
    for P in plans:
        P.reset()
    awcpml.Batch(plans).close()
    for P in plans:
        P.close()


def test_batch_two_step_inherits_and_replays():
    print("[gpu] batch two-step: setting taken from the plans, graph replays", flush=True)
    oracles = [_oracle(small3d(**c)) for c in BATCH_3D[:2]]
    plans = [_plan(o, dtype="f32") for o in oracles]
    nt = max(o.nt for o in oracles)
    B = awcpml.Batch(plans)
    B.run()
    one = _batch_state(plans)
    B.close()
    for P in plans:
        P.set_time_block(2, 8, 0, 2)
        P.reset()
    B = awcpml.Batch(plans)
    assert B.time_block() == (2, 8, 2)
    B.set_graph(True)
    for rep in range(3):
        _batch_run(B, nt, (4,))
        _same(_batch_state(plans), one)
        # the plans read and run on their own between replays
        assert np.array_equal(plans[0].traces(), one[0][0])
    B.close()
    # after the batch is gone the plans run alone as before
    for P, o in zip(plans, oracles):
        P.reset()
        P.advance(o.nt)
    _same(_batch_state(plans), one)
    for P in plans:
        P.close()


def test_batch_two_step_2d_bitwise():
    print("[gpu] batch two-step 2D fp64, one plan with 2 shots", flush=True)
    oracles = [_oracle(_interior_2d(nt=61)),
               _oracle(_interior_2d(n=(150, 47), halo=7, freetop=False, nt=44))]
    plans = [_plan(oracles[0], nshots=2), _plan(oracles[1])]
    nt = max(o.nt for o in oracles)
    B = awcpml.Batch(plans)
    B.run()
    one = _batch_state(plans)
    B.close()
    for rows, pf in ((4, 2), (8, 3)):
        for P in plans:
            P.reset()  # a batch run may leave the buffers in other roles
        B = awcpml.Batch(plans)
        B.set_time_block(2, rows, pf)
        assert B.time_block()[0] == 2
        for graph in (False, True):
            B.set_graph(graph)
            _batch_run(B, nt, (5,))
            _same(_batch_state(plans), one)
        B.close()
    for o in oracles:
        o.run()
    assert np.array_equal(plans[0].traces(1), oracles[0].traces())
    assert np.array_equal(plans[1].traces(), oracles[1].traces())
    for P in plans:
        P.close()


def test_batch_two_step_falls_back_for_dirichlet_source():
    print("[gpu] batch two-step fallback: a plan with a source on the top row", flush=True)
    oracles = [_oracle(small2d(nt=40)), _oracle(_interior_2d(nt=40))]
    plans = [_plan(o) for o in oracles]
    B = awcpml.Batch(plans)
    B.set_time_block(2, 4, 2)
    assert B.time_block()[0] == 1
    B.run()
    for P, o in zip(plans, oracles):
        o.run()
        assert np.array_equal(P.traces(), o.traces())
    B.close()
    for P in plans:
        P.close()


def test_autotune_batch_two_step_keeps_results():
    print("[gpu] autotune_batch with two-step candidates", flush=True)
    oracles = [_oracle(small3d(**c)) for c in BATCH_3D]
    plans = [_plan(o, dtype="f32") for o in oracles]
    nt = max(o.nt for o in oracles)
    B = awcpml.Batch(plans)
    B.run()
    one = _batch_state(plans)
    B.close()
    for P in plans:
        P.reset()
    best = awcpml.autotune_batch(plans, cands=[(4, 32, 2)], launches=2)
    B = awcpml.Batch(plans)
    assert B.time_block()[0] == (2 if best["time_block"] else 1)
    _batch_run(B, nt)
    _same(_batch_state(plans), one)
    B.close()
    for P in plans:
        P.close()


def test_time_advance_equals_advance():
    print("[gpu] time_advance (plan and batch, one and two steps, graphs) equals advance",
          flush=True)
    oracles = [_oracle(small3d(**c)) for c in BATCH_3D[:2]]
    plans = [_plan(o, dtype="f32") for o in oracles]
    nt = max(o.nt for o in oracles)
    B = awcpml.Batch(plans)
    B.run()
    one = _batch_state(plans)
    B.close()
    for steps in (1, 2):
        for graph in (False, True):
            for P in plans:
                P.reset()
            B = awcpml.Batch(plans)
            B.set_time_block(steps, 8, 2)
            B.set_graph(graph)
            B.reset()
            ms = B.time_advance(3)
            ms += B.time_advance(nt - 3)
            assert ms > 0.0
            _same(_batch_state(plans), one)
            B.close()
    # a plan alone, with its own time blocking and graphs
    P, o = plans[0], oracles[0]
    for steps in (1, 2):
        P.set_time_block(steps, 16, 0, 2)
        P.set_graph(True)
        P.reset()
        P.time_advance(5)
        P.time_advance(o.nt - 5)
        assert np.array_equal(P.traces(), one[0][0])
        assert np.array_equal(P.field(0, 1), one[0][2])
    P.set_graph(False)
    for P in plans:
        P.close()


@pytest.mark.parametrize("nts", [(1, 7), (2, 3), (5,)])
def test_batch_two_step_short_runs(nts):
    print("[gpu] batch two-step, short runs nt=%s" % (nts,), flush=True)
    oracles = [_oracle(small3d(**dict(BATCH_3D[q], nt=n))) for q, n in enumerate(nts)]
    plans = [_plan(o) for o in oracles]
    nt = max(nts)
    B = awcpml.Batch(plans)
    B.run()
    one = _batch_state(plans)
    B.close()
    for P in plans:
        P.reset()
    B = awcpml.Batch(plans)
    B.set_time_block(2, 8, 2)
    for graph in (False, True):
        B.set_graph(graph)
        _batch_run(B, nt)
        _same(_batch_state(plans), one)
    B.close()
    for P, o in zip(plans, oracles):
        o.run()
        assert np.array_equal(P.traces(), o.traces())
        P.close()
