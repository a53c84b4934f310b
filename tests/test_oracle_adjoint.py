"""The oracle's discrete adjoint (FWI gradient of the trace misfit with
respect to fact = vel^2 dt^2) against central finite differences of the
oracle's own forward solve.

J(fact) = 1/2 sum_{it, r} (trace[it, r] - obs[it, r])^2.  For a random
direction d, (J(f + e d) - J(f - e d)) / 2e must equal <grad, d>; the
central difference has O(e^2) truncation error and J is a polynomial of
degree 2 nt in fact, so e = 1e-3 of a 1 % perturbation leaves a relative
gap around 1e-9.  Directions confined to the CPML slabs pin the psi/xi
transpose terms separately from the interior stencil, whose contribution
otherwise dominates.
"""
import numpy as np
import pytest

from cases import oracle_problem, small2d, small3d


def _fd_check(case, region, obs_scale, seed=5):
    o = oracle_problem(case)
    f0 = o.fact()
    base = f0 * 1.02
    rng = np.random.Generator(np.random.PCG64(seed))
    o.run()
    obs = o.traces() * obs_scale  # observed data from the unperturbed model
    o.set_fact(base)
    J, g = o.fwi_gradient(obs)
    assert g.shape == tuple(case["n"])
    d = rng.standard_normal(o.n) * f0 * 0.01
    mask = np.zeros(o.n, dtype=bool)
    h = case["halo"]
    if region == "all":
        mask[...] = True
    else:  # CPML slabs only (every absorbing side, top included when open)
        for a in range(len(o.n)):
            sl = [slice(None)] * len(o.n)
            sl[a] = slice(0, h + 1)
            mask[tuple(sl)] = True
            sl[a] = slice(o.n[a] - h - 1, o.n[a])
            mask[tuple(sl)] = True
    d = np.where(mask, d, 0.0)
    eps = 1e-3
    o.set_fact(base + eps * d)
    Jp = o.misfit(obs)
    o.set_fact(base - eps * d)
    Jm = o.misfit(obs)
    o.close()
    fd = (Jp - Jm) / (2 * eps)
    ad = float((g * d).sum())
    assert J > 0 and abs(fd) > 0
    return fd, ad


CASES = {
    "2d_freetop": lambda: small2d(freetop=True, nt=200),
    "2d_open": lambda: small2d(freetop=False, n=(41, 37), halo=5, nt=150),
    "3d_freetop": lambda: small3d(freetop=True, nt=60),
    "3d_open_2src": lambda: small3d(freetop=False, nt=50, nsrc=2),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("region", ["all", "cpml"])
def test_adjoint_matches_central_differences(name, region):
    fd, ad = _fd_check(CASES[name](), region, obs_scale=0.5)
    assert abs(fd - ad) <= 1e-7 * abs(fd), (fd, ad)


def test_adjoint_zero_residual_gives_zero_gradient():
    case = small2d(nt=80)
    o = oracle_problem(case)
    o.run()
    obs = o.traces()
    J, g = o.fwi_gradient(obs)
    assert J == 0.0
    assert not g.any()
    o.close()


def test_adjoint_is_linear_in_the_residual():
    """grad(obs) is affine in obs: grad(2r) = 2 grad(r) with r = trace - obs."""
    case = small3d(nt=40)
    o = oracle_problem(case)
    o.run()
    tr = o.traces()
    rng = np.random.Generator(np.random.PCG64(11))
    r = rng.standard_normal(tr.shape) * np.abs(tr).max()
    J1, g1 = o.fwi_gradient(tr - r)
    J2, g2 = o.fwi_gradient(tr - 2 * r)
    assert np.isclose(J2, 4 * J1, rtol=1e-12)
    assert np.allclose(g2, 2 * g1, rtol=1e-10, atol=1e-12 * np.abs(g2).max())
    o.close()
