"""Checks that pin the oracle where the reference's golden data stop.

* calc_Kab_CPML: sign/range invariants and the exact left/right mirror the
  reference's construction implies (src/utils.jl:39-54).
* rickersource1D: peak value, exact symmetry about t0, zero crossings.
* RoundNearestTiesUp snapping on ties and near-ties.
* 2D: transposing the model transposes the result bit for bit (IEEE + is
  commutative and the damping terms are added x then y).
* 3D: swapping x and y is exact for the same reason; a cyclic permutation
  of the axes agrees to rounding (the z term is added last).
* 3D and 2D: a vectorised numpy restatement (tests/numpy_step.py) agrees
  with the C oracle bit for bit on tiny heterogeneous grids, including
  Dirichlet-node sources/receivers and duplicate sources, and on grids
  whose two CPML slabs of an axis overlap (n < 2*halo+2).
* free top: psi_y/xi_y of the top side stay exactly zero.
* halo = 0: no CPML term at all (the textbook leapfrog with Dirichlet
  edges), checked against a direct numpy leapfrog.
* the reference's assertion failures raise.
"""
import math

import numpy as np
import pytest

import oracle
import numpy_step
from cases import OVERLAP_CASES, oracle_problem, overlap_case, rel_l2, small2d, small3d


def test_calc_kab_invariants():
    for halo in (1, 5, 20, 37):
        dt, d0, amax = 1e-3, 400.0, math.pi * 8.0
        a_l, a_r, bK_l, bK_r = oracle.calc_kab_cpml(halo, dt, 2.0, d0, amax, 1.0, "ongrd")
        ah_l, ah_r, bKh_l, bKh_r = oracle.calc_kab_cpml(halo, dt, 2.0, d0, amax, 1.0, "halfgrd")
        assert len(a_l) == halo and len(ah_l) == halo + 1
        for a in (a_l, a_r, ah_l, ah_r):
            assert np.all(a <= 0.0)
        for b in (bK_l, bK_r, bKh_l, bKh_r):
            assert np.all((b > 0.0) & (b <= 1.0))
        # left is the mirror of right, bit for bit
        assert np.array_equal(a_l, a_r[::-1]) and np.array_equal(bK_l, bK_r[::-1])
        assert np.array_equal(ah_l, ah_r[::-1]) and np.array_equal(bKh_l, bKh_r[::-1])
        # no damping at the inner edge (normalised distance 0)
        assert a_r[0] == 0.0 and ah_r[0] == 0.0
        # |a| grows towards the outer edge (b/K is not monotone: alpha
        # falls while d rises from 0); the outermost b/K is the smallest
        assert np.all(np.diff(a_r) < 0)
        assert bK_r[-1] == bK_r.min() and bKh_r[-1] == bKh_r.min()
        # K = 1 (K_max = 1): b/K is exp(-(d + alpha) dt), the innermost
        # node has d = 0, alpha = alpha_max
        assert bK_r[0] == pytest.approx(math.exp(-amax * dt), rel=1e-15)
    with pytest.raises(ValueError, match="onwhere"):
        oracle.calc_kab_cpml(5, 1e-3, 2.0, 1.0, 1.0, 1.0, "nowhere")


def test_ricker_properties():
    f0, t0 = 8.0, 0.5
    assert oracle.ricker(t0, t0, f0) == 1.0
    for x in (0.25, 0.125, 0.0625, 0.03125, 0.015625):  # t0 +- x exact in binary
        assert oracle.ricker(t0 + x, t0, f0) == oracle.ricker(t0 - x, t0, f0)
    tz = t0 + 1.0 / (math.pi * f0 * math.sqrt(2.0))
    assert abs(oracle.ricker(tz, t0, f0)) < 1e-15
    tmin = t0 + math.sqrt(1.5) / (math.pi * f0)  # minimum: -2 exp(-3/2)
    assert oracle.ricker(tmin, t0, f0) == pytest.approx(-2 * math.exp(-1.5), rel=1e-14)


def test_round_ties_up():
    assert oracle.round_ties_up(1.5) == 2
    assert oracle.round_ties_up(2.5) == 3
    assert oracle.round_ties_up(2.4999999999) == 2
    assert oracle.round_ties_up(0.49999999999999994) == 0
    assert oracle.round_ties_up(101.0) == 101


def _sym_case_2d(n=41, halo=6, nt=160):
    rng = np.random.Generator(np.random.PCG64(11))
    vel = 1800.0 + 300.0 * rng.random((n, n))
    L = (n - 1) * 10.0
    src = [(17 * 10.0, 23 * 10.0), (9 * 10.0, 30 * 10.0)]
    rec = [(12 * 10.0, 7 * 10.0), (30 * 10.0, 20 * 10.0)]
    return vel, L, src, rec, nt


def test_transpose_2d_exact():
    vel, L, src, rec, nt = _sym_case_2d()
    kw = dict(halo=6, rcoef=1e-4, ppw=2.0, freetop=False)
    dt_probe = math.sqrt(2) / (vel.max() * (0.2)) / 2
    lt = (nt - 0.5) * dt_probe
    a = oracle.OracleProblem(vel.shape, (L, L), lt, np.asfortranarray(vel), src, [0.03, 0.05], 20.0,
                             rec, **kw)
    b = oracle.OracleProblem(vel.shape, (L, L), lt, np.asfortranarray(vel.T),
                             [(s[1], s[0]) for s in src], [0.03, 0.05], 20.0,
                             [(r[1], r[0]) for r in rec], **kw)
    a.run()
    b.run()
    assert a.nt == nt
    assert np.array_equal(a.pend(), b.pend().T)
    assert np.array_equal(a.traces(), b.traces())


def _cube_case(n=27, halo=5, nt=90, seed=5):
    rng = np.random.Generator(np.random.PCG64(seed))
    vel = 1800.0 + 300.0 * rng.random((n, n, n))
    L = (n - 1) * 10.0
    dt_probe = math.sqrt(3) / (vel.max() * 0.3) / 2
    return vel, L, (nt - 0.5) * dt_probe, nt


def _run3(vel, L, lt, src, rec):
    o = oracle.OracleProblem(vel.shape, (L, L, L), lt, np.asfortranarray(vel), src, [0.02] * len(src),
                             25.0, rec, halo=5, rcoef=1e-4, ppw=2.0, freetop=False)
    o.run()
    return o


def test_swap_xy_3d_exact():
    vel, L, lt, nt = _cube_case()
    src = [(11 * 10.0, 14 * 10.0, 9 * 10.0)]
    rec = [(6 * 10.0, 20 * 10.0, 15 * 10.0), (21 * 10.0, 3 * 10.0, 4 * 10.0)]
    a = _run3(vel, L, lt, src, rec)
    b = _run3(np.transpose(vel, (1, 0, 2)), L, lt, [(s[1], s[0], s[2]) for s in src],
              [(r[1], r[0], r[2]) for r in rec])
    assert a.nt == nt
    assert np.array_equal(a.pend(), np.transpose(b.pend(), (1, 0, 2)))
    assert np.array_equal(a.traces(), b.traces())


def test_cyclic_axes_3d():
    vel, L, lt, nt = _cube_case()
    src = [(11 * 10.0, 14 * 10.0, 9 * 10.0)]
    rec = [(6 * 10.0, 20 * 10.0, 15 * 10.0), (21 * 10.0, 3 * 10.0, 4 * 10.0)]
    a = _run3(vel, L, lt, src, rec)
    # new axes (x', y', z') = (z, x, y)
    b = _run3(np.transpose(vel, (2, 0, 1)), L, lt, [(s[2], s[0], s[1]) for s in src],
              [(r[2], r[0], r[1]) for r in rec])
    assert rel_l2(a.traces(), b.traces()) < 1e-12
    assert rel_l2(a.pend(), np.transpose(b.pend(), (1, 2, 0))) < 1e-12
    assert not np.array_equal(a.pend(), np.zeros_like(a.pend()))


@pytest.mark.parametrize("case_fn,kw", [
    (small3d, dict(freetop=True, nsrc=1)),
    (small3d, dict(freetop=False, nsrc=2)),
    (small3d, dict(n=(19, 17, 23), halo=3, freetop=False, nsrc=1, nt=60)),
    (small2d, dict(freetop=True)),
    (small2d, dict(freetop=False, n=(31, 29), halo=4)),
])
def test_numpy_restatement_bitwise(case_fn, kw):
    _numpy_vs_oracle(case_fn(**kw))


@pytest.mark.parametrize("n,halo,freetop", OVERLAP_CASES)
def test_overlapping_slabs_bitwise(n, halo, freetop):
    """Grids with n < 2*halo+2: the numpy restatement (boolean masks
    left = q <= halo, right = not left and q >= n-halo-1) and the oracle's
    if/elseif agree bit for bit, and the overlap region gets damped."""
    case = overlap_case(n, halo, freetop=freetop)
    o = _numpy_vs_oracle(case)
    assert np.abs(o.xi(0, 0)).max() > 0


def _numpy_vs_oracle(case):
    o = oracle_problem(case)
    spacing = [case["l"][a] / (case["n"][a] - 1) for a in range(len(case["n"]))]
    st = numpy_step.from_oracle(o, spacing)
    for it in range(1, o.nt + 1):
        numpy_step.step(st, it)
    o.run()
    assert np.array_equal(st["pcur"], o.pend())
    assert np.array_equal(st["traces"], o.traces())
    assert np.abs(o.traces()).max() > 0
    return o


def test_duplicate_and_boundary_sources_bitwise():
    """Two sources on one node are added in order; a source on a Dirichlet
    node accumulates there; a receiver there records it."""
    case = small2d(freetop=False)
    case["srcpos"] = np.array([[200.0, 150.0], [200.0, 150.0], [0.0, 100.0]])
    case["t0s"] = [0.03, 0.04, 0.05]
    case["recpos"] = np.array([[0.0, 100.0], [200.0, 150.0]])
    o = oracle_problem(case)
    spacing = [case["l"][a] / (case["n"][a] - 1) for a in range(2)]
    st = numpy_step.from_oracle(o, spacing)
    for it in range(1, o.nt + 1):
        numpy_step.step(st, it)
    o.run()
    assert np.array_equal(st["pcur"], o.pend())
    assert np.array_equal(st["traces"], o.traces())
    assert np.abs(o.traces()[:, 0]).max() > 0


@pytest.mark.parametrize("ndim", [2, 3])
def test_free_top_stays_zero(ndim):
    case = small2d(freetop=True) if ndim == 2 else small3d(freetop=True)
    o = oracle_problem(case)
    o.run()
    assert np.all(o.psi(1, 0) == 0.0) and np.all(o.xi(1, 0) == 0.0)
    assert np.abs(o.psi(1, 1)).max() > 0 and np.abs(o.xi(1, 1)).max() > 0


def test_halo0_is_plain_leapfrog():
    case = small2d(freetop=False)
    case["halo"] = 0
    o = oracle_problem(case)
    dx = case["l"][0] / (case["n"][0] - 1)
    fact = o.fact()
    ps, pr = o.positions()
    w = o.dt2srctf()
    # three storage buffers rotated like the reference's time loop; the
    # Dirichlet nodes of the buffer written as pnew keep its old contents
    pold, pcur, pnew = np.zeros(o.n), np.zeros(o.n), np.zeros(o.n)
    tr = np.zeros((o.nt, len(pr)))
    for it in range(o.nt):
        lap = ((pcur[2:, 1:-1] - 2.0 * pcur[1:-1, 1:-1] + pcur[:-2, 1:-1]) * (1.0 / (dx * dx))
               + (pcur[1:-1, 2:] - 2.0 * pcur[1:-1, 1:-1] + pcur[1:-1, :-2]) * (1.0 / (dx * dx)))
        pnew[1:-1, 1:-1] = 2.0 * pcur[1:-1, 1:-1] - pold[1:-1, 1:-1] + fact[1:-1, 1:-1] * lap + 0.0
        for s in range(len(ps)):
            pnew[ps[s, 0] - 1, ps[s, 1] - 1] += w[it, s]
        for r in range(len(pr)):
            tr[it, r] = pnew[pr[r, 0] - 1, pr[r, 1] - 1]
        pold, pcur, pnew = pcur, pnew, pold
    o.run()
    assert np.array_equal(o.traces(), tr)
    assert np.array_equal(o.pend(), pcur)


def test_assertions():
    case = small2d()
    bad = dict(case)
    bad["ppw"] = 1000.0
    with pytest.raises(oracle.OracleError, match="points per wavelength"):
        oracle_problem(bad)
    bad = dict(case)
    bad["srcpos"] = np.array([[1e6, 10.0]])
    bad["t0s"] = [0.01]
    with pytest.raises(oracle.OracleError, match="source is not inside"):
        oracle_problem(bad)
    bad = dict(case)
    bad["recpos"] = np.array([[10.0, -100.0]])
    with pytest.raises(oracle.OracleError, match="receiver is not inside"):
        oracle_problem(bad)
