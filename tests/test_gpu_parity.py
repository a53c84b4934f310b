"""GPU parity of the HIP path against the CPU FP64 oracle and the
reference's golden fields.

FP64 bar: bit-for-bit equality of traces and final field when the GPU plan is
fed the oracle's own setup arrays (same coefficients, wavelets, positions);
<= 1e-12 relative L2 when the product's host setup (numpy) is used end to
end.  FP32 bar: receiver traces within 1e-5 relative L2 of the FP64 oracle.
"""
import numpy as np
import pytest

import awcpml
from cases import (fixture_case, load_fixtures, oracle_problem, plain2d, rel_l2, small2d,
                   small3d)

pytestmark = pytest.mark.gpu

FP64_E2E_TOL = 1e-12   # end to end, product host setup vs oracle setup
FP32_TRACE_TOL = 1e-5  # FP32 receiver traces vs the FP64 oracle
SQRT_EPS = 1.4901161193847656e-08  # the reference's isapprox default


def plan_from_oracle(o, dtype="f64", nshots=1):
    P = awcpml.Plan(o.n, o.halo, dtype=dtype, nshots=nshots)
    inv_d = [1.0 / d for d in _spacing(o)]
    inv_d2 = [1.0 / (d * d) for d in _spacing(o)]
    P.set_model(o.fact(), inv_d, inv_d2)
    for a in range(o.ndim):
        P.set_cpml(a, [o.coef(a, w) for w in range(8)])
    ps, pr = o.positions()
    for s in range(nshots):
        P.set_sources(s, ps, o.dt2srctf())
        P.set_receivers(s, pr)
    return P


def _spacing(o):
    return o._spacing


def run_oracle(case, nsteps=None):
    o = oracle_problem(case)
    o._spacing = [case["l"][a] / (case["n"][a] - 1) for a in range(len(case["n"]))]
    o.run(nsteps)
    return o


def run_gpu(o, nsteps=None, dtype="f64"):
    P = plan_from_oracle(o, dtype=dtype)
    P.reset()
    P.advance(o.it_done if nsteps is None else nsteps)
    tr, pend = P.traces(0), P.field(0, 1)
    P.close()
    return tr, pend


def solve_case(case, dtype="f64"):
    srcs = awcpml.Sources(case["srcpos"], case["t0s"], [awcpml.rickersource1D] * len(case["t0s"]),
                          case["f0"])
    recs = awcpml.Receivers(case["recpos"])
    kw = dict(halo=case["halo"], rcoef=case["rcoef"], ppw=case.get("ppw", 10.0),
              freetop=case["freetop"], dtype=dtype)
    if len(case["n"]) == 2:
        pend = awcpml.solve2D(case["l"][0], case["l"][1], case["lt"], case["vel"], srcs, recs, **kw)
    else:
        pend = awcpml.solve3D(case["l"][0], case["l"][1], case["l"][2], case["lt"], case["vel"],
                              srcs, recs, **kw)
    return pend, recs.seismograms


@pytest.mark.parametrize("name", ["acoustic2D_center_halo20", "acoustic2D_center_freetop_halo20",
                                  "acoustic2D_gradient_freetop_halo20"])
def test_golden_fields_2d(name):
    """The reference's three golden fields (test/test2D_xPU.jl:108-195)."""
    case, ref = load_fixtures()[name]
    print(f"[gpu] golden {name} n={case['n']}", flush=True)
    pend, traces = solve_case(case)
    o = run_oracle(case)
    assert rel_l2(pend, ref) <= SQRT_EPS
    assert rel_l2(pend, o.pend()) <= FP64_E2E_TOL
    assert rel_l2(traces, o.traces()) <= FP64_E2E_TOL
    # kernel bitwise: same setup arrays as the oracle
    tr_b, pend_b = run_gpu(o)
    assert np.array_equal(pend_b, o.pend())
    assert np.array_equal(tr_b, o.traces())


@pytest.mark.parametrize("freetop", [False, True])
def test_small2d_bitwise(freetop):
    case = small2d(freetop=freetop)
    print(f"[gpu] small2d freetop={freetop}", flush=True)
    o = run_oracle(case)
    tr, pend = run_gpu(o)
    assert np.array_equal(tr, o.traces())
    assert np.array_equal(pend, o.pend())


@pytest.mark.parametrize("freetop,nsrc", [(True, 1), (False, 2)])
def test_small3d_bitwise(freetop, nsrc):
    case = small3d(freetop=freetop, nsrc=nsrc)
    print(f"[gpu] small3d freetop={freetop} nsrc={nsrc}", flush=True)
    o = run_oracle(case)
    tr, pend = run_gpu(o)
    assert np.array_equal(tr, o.traces())
    assert np.array_equal(pend, o.pend())


def test_plain2d_config_traces():
    case = plain2d()
    print("[gpu] plain2d 298^2 nt=1000", flush=True)
    o = run_oracle(case)
    assert o.nt == 1000
    pend, traces = solve_case(case)
    assert rel_l2(traces, o.traces()) <= FP64_E2E_TOL
    assert rel_l2(pend, o.pend()) <= FP64_E2E_TOL
    tr_b, pend_b = run_gpu(o)
    assert np.array_equal(tr_b, o.traces())
    assert np.array_equal(pend_b, o.pend())


def test_fp32_traces_2d():
    case = plain2d()
    print("[gpu] plain2d fp32", flush=True)
    o = run_oracle(case)
    _, traces = solve_case(case, dtype="f32")
    assert rel_l2(traces, o.traces()) <= FP32_TRACE_TOL


def test_fp32_traces_3d_small():
    case = small3d(n=(61, 53, 47), halo=8, nt=300)
    print("[gpu] small3d fp32", flush=True)
    o = run_oracle(case)
    tr, _ = run_gpu(o, dtype="f32")
    assert rel_l2(tr, o.traces()) <= FP32_TRACE_TOL


def _case_2d_narrow():
    """nx = 2*halo + 2: the two x slabs touch, so the x slabs take the
    per-lane path instead of the lane-shuffle one."""
    c = small2d(n=(14, 40), halo=6, nt=120, freetop=False)
    dx = 10.0
    c["srcpos"] = np.array([[7 * dx, 12 * dx], [3 * dx, 0.0]])
    c["recpos"] = np.array([[(2 + r) * dx, 8 * dx] for r in range(10)] + [[0.0, 20 * dx]])
    return c


@pytest.mark.parametrize("name", ["3d_straddle", "2d_straddle", "2d_narrow", "3d_fast"])
def test_x_slab_paths_bitwise(name):
    """Both x-slab code paths against the oracle: grids whose high x slab
    crosses a 64-wide tile boundary or whose slabs touch use the per-lane
    path, the others the lane-shuffle path."""
    case = {"3d_straddle": lambda: small3d(n=(67, 21, 19), halo=6, nt=50, freetop=False),
            "2d_straddle": lambda: small2d(n=(70, 30), halo=6, nt=120),
            "2d_narrow": _case_2d_narrow,
            "3d_fast": lambda: small3d(n=(60, 23, 21), halo=8, nt=50)}[name]()
    print(f"[gpu] x slab paths {name}", flush=True)
    o = oracle_problem(case)
    o.run()
    sp = [case["l"][a] / (case["n"][a] - 1) for a in range(len(case["n"]))]
    P = awcpml.Plan(o.n, o.halo, dtype="f64")
    P.set_model(o.fact(), [1 / d for d in sp], [1 / (d * d) for d in sp])
    for a in range(o.ndim):
        P.set_cpml(a, [o.coef(a, w) for w in range(8)])
    ps, pr = o.positions()
    P.set_sources(0, ps, o.dt2srctf())
    P.set_receivers(0, pr)
    for by, ml, pf in ((4, 7, 2), (1, 16, 3), (2, 40, 5)):
        P.set_tiling(by, ml, pf)
        P.reset()
        P.advance(o.nt)
        assert np.array_equal(P.traces(), o.traces()), (by, ml, pf)
        assert np.array_equal(P.field(), o.pend()), (by, ml, pf)
    P.close()


def test_do_save_snapshots_match_oracle(tmp_path):
    """solve3D(do_save=True, nsave=k): a snapshot after every k-th step
    (acoustic_3D_xPU.jl:480-498), whose pcur is bitwise the oracle's field
    at that step."""
    case = small3d(freetop=True, nsrc=2, nt=47)
    nsave = 10
    print("[gpu] do_save snapshots small3d", flush=True)
    srcs = awcpml.Sources(case["srcpos"], case["t0s"], [awcpml.rickersource1D] * len(case["t0s"]),
                          case["f0"])
    recs = awcpml.Receivers(case["recpos"])
    pend = awcpml.solve3D(case["l"][0], case["l"][1], case["l"][2], case["lt"], case["vel"], srcs,
                          recs, halo=case["halo"], rcoef=case["rcoef"], ppw=case["ppw"],
                          freetop=case["freetop"], do_save=True, nsave=nsave, save_name="snap",
                          save_dir=str(tmp_path))
    o = oracle_problem(case)
    assert o.nt == 47
    ps, pr = o.positions()
    saved = sorted(p.name for p in tmp_path.iterdir())
    assert saved == [f"snap_it{it}.npz" for it in (10, 20, 30, 40)]
    for it in (10, 20, 30, 40):
        o.run(nsave)
        with np.load(tmp_path / f"snap_it{it}.npz", allow_pickle=False) as z:
            assert np.array_equal(z["pcur"], o.pend()), it
            assert float(z["lx"]) == case["l"][0] and float(z["lz"]) == case["l"][2]
            assert int(z["halo"]) == case["halo"]
            assert np.array_equal(z["possrcs"], ps) and np.array_equal(z["posrecs"], pr)
    o.run()
    assert np.array_equal(pend, o.pend())
    assert np.array_equal(recs.seismograms, o.traces())
