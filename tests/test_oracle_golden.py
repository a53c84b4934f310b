"""The CPU FP64 oracle against the reference's own golden data.

Golden fields: the three BSON fixtures of test/test2D_xPU.jl:108-195
(converted to raw float64 by tests/golden/make_fixtures.py), compared with
the reference's criterion `isapprox` (rtol = sqrt(eps) in relative L2,
test/test2D_xPU.jl:9-10).  The oracle also has to beat a tighter bar that a
one-ulp setup difference still passes but a real error would not.

psi known-answer tests: test/test2D.jl:13-43 and test/test3D.jl:13-71,
exact equality.
"""
import numpy as np
import pytest

import oracle
from cases import load_fixtures, oracle_problem, rel_l2

SQRT_EPS = 1.4901161193847656e-08
TIGHT = 1e-10   # one-ulp setup differences give <= ~5e-12 on these fields


@pytest.mark.parametrize("name", ["acoustic2D_center_halo20", "acoustic2D_center_freetop_halo20",
                                  "acoustic2D_gradient_freetop_halo20"])
def test_oracle_matches_golden_field(name):
    case, ref = load_fixtures()[name]
    o = oracle_problem(case)
    o.run()
    pend = o.pend()
    assert pend.shape == ref.shape
    err = rel_l2(pend, ref)
    assert err <= SQRT_EPS
    assert err <= TIGHT, err


def test_gradient_fixture_boundary_quirk():
    """Sources on the Dirichlet row accumulate per buffer (the reference
    never overwrites those nodes): the golden field has a non-zero
    boundary, and so must the oracle's."""
    case, ref = load_fixtures()["acoustic2D_gradient_freetop_halo20"]
    o = oracle_problem(case)
    o.run()
    pend = o.pend()
    assert np.abs(ref[:, 0]).max() > 0
    assert np.abs(pend[:, 0]).max() > 0
    assert rel_l2(pend[:, 0], ref[:, 0]) < 1e-6
    # the other edges stay exactly zero
    for edge in (pend[0, :], pend[-1, :], pend[:, -1]):
        assert np.all(edge == 0.0)


def test_psi_kat_2d():
    """test/test2D.jl:13-43"""
    nx = ny = 101
    halo = 20
    pcur = np.ones((nx, ny), order="F")
    pcur[[0, -1], :] = 0.0
    pcur[:, [0, -1]] = 0.0
    psi = [np.ones((halo + 1, ny), order="F"), np.ones((halo + 1, ny), order="F"),
           np.ones((nx, halo + 1), order="F"), np.ones((nx, halo + 1), order="F")]
    h = halo + 1
    coefs = [0.5 * np.ones(h), 2 * np.ones(h), 0.5 * np.ones(h), 2 * np.ones(h),
             0.5 * np.ones(h), 2 * np.ones(h), 0.5 * np.ones(h), 2 * np.ones(h)]
    oracle.update_psi_2d(pcur, psi, coefs, halo, 1 / 10.0, 1 / 10.0)
    x_l, x_r, y_l, y_r = psi
    assert np.array_equal(x_l[0, :], [0.5] + [0.55] * (ny - 2) + [0.5])
    assert np.array_equal(x_l[1:, :], 0.5 * np.ones((halo, ny)))
    assert np.array_equal(x_r[:-1, :], 2.0 * np.ones((halo, ny)))
    assert np.array_equal(x_r[-1, :], [2.0] + [1.8] * (ny - 2) + [2.0])
    assert np.array_equal(y_l[:, 0], [0.5] + [0.55] * (nx - 2) + [0.5])
    assert np.array_equal(y_l[:, 1:], 0.5 * np.ones((nx, halo)))
    assert np.array_equal(y_r[:, :-1], 2.0 * np.ones((nx, halo)))
    assert np.array_equal(y_r[:, -1], [2.0] + [1.8] * (nx - 2) + [2.0])


def test_psi_kat_3d():
    """test/test3D.jl:13-71 (at a reduced size; the values do not depend
    on it)"""
    nx = ny = nz = 41
    halo = 20
    pcur = np.ones((nx, ny, nz), order="F")
    pcur[[0, -1], :, :] = 0.0
    pcur[:, [0, -1], :] = 0.0
    pcur[:, :, [0, -1]] = 0.0
    h = halo + 1
    psi = [np.ones((h, ny, nz), order="F"), np.ones((h, ny, nz), order="F"),
           np.ones((nx, h, nz), order="F"), np.ones((nx, h, nz), order="F"),
           np.ones((nx, ny, h), order="F"), np.ones((nx, ny, h), order="F")]
    coefs = []
    for _ in range(3):
        coefs += [0.5 * np.ones(h), 2 * np.ones(h), 0.5 * np.ones(h), 2 * np.ones(h)]
    oracle.update_psi_3d(pcur, psi, coefs, halo, 0.1, 0.1, 0.1)

    def face(v_in, v_edge, n1, n2):
        a = v_in * np.ones((n1, n2))
        a[[0, -1], :] = v_edge
        a[:, [0, -1]] = v_edge
        return a

    x_l, x_r, y_l, y_r, z_l, z_r = psi
    assert np.array_equal(x_l[0], face(0.55, 0.5, ny, nz))
    assert np.array_equal(x_l[1:], 0.5 * np.ones((halo, ny, nz)))
    assert np.array_equal(x_r[:-1], 2.0 * np.ones((halo, ny, nz)))
    assert np.array_equal(x_r[-1], face(1.8, 2.0, ny, nz))
    assert np.array_equal(y_l[:, 0, :], face(0.55, 0.5, nx, nz))
    assert np.array_equal(y_l[:, 1:, :], 0.5 * np.ones((nx, halo, nz)))
    assert np.array_equal(y_r[:, :-1, :], 2.0 * np.ones((nx, halo, nz)))
    assert np.array_equal(y_r[:, -1, :], face(1.8, 2.0, nx, nz))
    assert np.array_equal(z_l[:, :, 0], face(0.55, 0.5, nx, ny))
    assert np.array_equal(z_l[:, :, 1:], 0.5 * np.ones((nx, ny, halo)))
    assert np.array_equal(z_r[:, :, :-1], 2.0 * np.ones((nx, ny, halo)))
    assert np.array_equal(z_r[:, :, -1], face(1.8, 2.0, nx, ny))
