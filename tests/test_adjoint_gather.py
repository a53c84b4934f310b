"""The gather form of the adjoint that the GPU kernel implements, restated
in numpy and checked against the oracle's scatter-form adjoint.

The GPU reverse step (acousticwavecpml.jl_amd/csrc/adjoint_kernels.h) does
not transpose the forward statement by statement: it gathers, per node,
    lam^s  = 2 lam^{s+1} - lam^{s+2} + L(f lam^{s+1})
             + sum_sides [ G^T(ah alpha^{s+1}) + D2^T(a beta^{s+1}) ] + R^T tbar^s
    beta^s = f lam^s + bK beta^{s+1}
    alpha^{s+1} = bKh alpha^{s+2} + Dpsi^T(f lam^{s+1} + a beta^{s+1})
    grad  += lam^s q^s,   q^s = L p^{s-1} + sum (dpsi + xi^s)
This file restates exactly those equations with array slices and requires
them to reproduce oracle_fwi_gradient (the oracle's transpose, itself pinned
by finite differences) to 1e-13, on 2D and 3D, free top and open grids.
"""
import numpy as np
import pytest

from cases import oracle_problem, small2d, small3d


def _sl(nd, a, s):
    t = [slice(None)] * nd
    t[a] = s
    return tuple(t)


def _shape(nd, a, length):
    s = [1] * nd
    s[a] = length
    return s


def _sides(n, h, a):
    """(side, edge offset, 0-based slab nodes, xi layer of the first node)"""
    return [(0, 0, np.arange(1, h + 1), 1),
            (1, n[a] - h - 2, np.arange(n[a] - h - 1, n[a] - 1), n[a] - h - 1)]


def _setup(o, spacing):
    n = list(o.n)
    nd = len(n)
    coef = [[o.coef(a, w) for w in range(8)] for a in range(nd)]
    inner = tuple(slice(1, m - 1) for m in n)
    interior = np.zeros(n, bool)
    interior[inner] = True
    return n, nd, o.halo, [1 / d for d in spacing], [1 / (d * d) for d in spacing], coef, inner, interior


def forward(o, fact, spacing):
    """Traces and q^s of every step (numpy forward, the oracle's semantics)."""
    n, nd, h, inv_d, inv_d2, coef, inner, interior = _setup(o, spacing)
    ps, pr = o.positions()
    ps = ps.astype(int) - 1
    pr = pr.astype(int) - 1
    wav = o.dt2srctf()
    P = [np.zeros(n) for _ in range(3)]
    psi = [[np.zeros(n[:a] + [h + 1] + n[a + 1:]) for _ in range(2)] for a in range(nd)]
    xi = [[np.zeros(n[:a] + [h] + n[a + 1:]) for _ in range(2)] for a in range(nd)]
    Q, tr = [], np.zeros((o.nt, len(pr)))
    for it in range(o.nt):
        pold, pcur, pnew = P
        for a in range(nd):
            for side, off, _, _ in _sides(n, h, a):
                ah = coef[a][4 + side].reshape(_shape(nd, a, h + 1))
                bKh = coef[a][6 + side].reshape(_shape(nd, a, h + 1))
                dp = pcur[_sl(nd, a, slice(off + 1, off + h + 2))] - pcur[_sl(nd, a, slice(off, off + h + 1))]
                psi[a][side][...] = bKh * psi[a][side] + ah * dp * inv_d[a]
        d2 = []
        for a in range(nd):
            D = (np.roll(pcur, -1, a) - 2 * pcur + np.roll(pcur, 1, a)) * inv_d2[a]
            d2.append(np.where(interior, D, 0.0))
        lap = sum(d2)
        slab = np.zeros(n)
        for a in range(nd):
            for side, off, nodes, x0 in _sides(n, h, a):
                for i0 in nodes:
                    e = i0 - off
                    dpsi = (psi[a][side][_sl(nd, a, e)] - psi[a][side][_sl(nd, a, e - 1)]) * inv_d[a]
                    l = i0 - x0
                    xn = coef[a][2 + side][l] * xi[a][side][_sl(nd, a, l)] + \
                        coef[a][0 + side][l] * (d2[a][_sl(nd, a, i0)] + dpsi)
                    xn = np.where(interior[_sl(nd, a, i0)], xn, 0.0)
                    xi[a][side][_sl(nd, a, l)] = xn
                    slab[_sl(nd, a, i0)] += np.where(interior[_sl(nd, a, i0)], dpsi + xn, 0.0)
        q = np.where(interior, lap + slab, 0.0)
        pnew[inner] = (2 * pcur - pold + fact * q)[inner]
        for s in range(len(ps)):
            pnew[tuple(ps[s])] += wav[it, s]
        for r in range(len(pr)):
            tr[it, r] = pnew[tuple(pr[r])]
        Q.append(q)
        P = [pcur, pnew, pold]
    return tr, Q


def gather_adjoint(o, fact, spacing, tbar, Q):
    """The GPU kernel's equations (module docstring), step s = nt .. 1."""
    n, nd, h, inv_d, inv_d2, coef, inner, interior = _setup(o, spacing)
    _, pr = o.positions()
    pr = pr.astype(int) - 1
    lam1, lam2 = np.zeros(n), np.zeros(n)
    alpha = [[np.zeros(n[:a] + [h + 1] + n[a + 1:]) for _ in range(2)] for a in range(nd)]
    beta = [[np.zeros(n[:a] + [h] + n[a + 1:]) for _ in range(2)] for a in range(nd)]
    grad = np.zeros(n)
    for s in range(o.nt, 0, -1):
        u = fact * lam1
        acc = 2 * lam1 - lam2
        for a in range(nd):
            acc = acc + (np.roll(u, -1, a) - 2 * u + np.roll(u, 1, a)) * inv_d2[a]
        for a in range(nd):
            for side, off, nodes, x0 in _sides(n, h, a):
                w = np.zeros(n)
                z = np.zeros(n)
                for i0 in nodes:
                    l = i0 - x0
                    zb = coef[a][0 + side][l] * beta[a][side][_sl(nd, a, l)]
                    z[_sl(nd, a, i0)] = zb
                    w[_sl(nd, a, i0)] = u[_sl(nd, a, i0)] + zb
                w[~interior] = 0.0
                z[~interior] = 0.0
                A = alpha[a][side]
                y = np.zeros(n)
                for e in range(h + 1):  # alpha^{s+1} from alpha^{s+2}
                    A[_sl(nd, a, e)] = coef[a][6 + side][e] * A[_sl(nd, a, e)] + \
                        (w[_sl(nd, a, off + e)] - w[_sl(nd, a, off + e + 1)]) * inv_d[a]
                    y[_sl(nd, a, off + e)] = coef[a][4 + side][e] * A[_sl(nd, a, e)]
                acc = acc + (np.roll(y, 1, a) - y) * inv_d[a]
                acc = acc + (np.roll(z, -1, a) - 2 * z + np.roll(z, 1, a)) * inv_d2[a]
        lam = np.where(interior, acc, 0.0)
        for r in range(len(pr)):
            if interior[tuple(pr[r])]:
                lam[tuple(pr[r])] += tbar[s - 1, r]
        for a in range(nd):
            for side, off, nodes, x0 in _sides(n, h, a):
                for i0 in nodes:
                    l = i0 - x0
                    b = coef[a][2 + side][l] * beta[a][side][_sl(nd, a, l)] + \
                        fact[_sl(nd, a, i0)] * lam[_sl(nd, a, i0)]
                    beta[a][side][_sl(nd, a, l)] = np.where(interior[_sl(nd, a, i0)], b, 0.0)
        grad += lam * Q[s - 1]
        lam2, lam1 = lam1, lam
    return grad


@pytest.mark.parametrize("case", [
    small2d(freetop=True, nt=90),
    small2d(freetop=False, n=(41, 37), halo=5, nt=70),
    small3d(freetop=True, nt=30),
    small3d(freetop=False, nt=25, nsrc=2),
], ids=["2d_freetop", "2d_open", "3d_freetop", "3d_open"])
def test_gather_form_equals_oracle_adjoint(case):
    o = oracle_problem(case)
    spacing = [case["l"][a] / (case["n"][a] - 1) for a in range(len(case["n"]))]
    o.run()
    obs = o.traces() * 0.5
    f = o.fact() * 1.02
    o.set_fact(f)
    J, g_ref = o.fwi_gradient(obs)
    tr, Q = forward(o, f, spacing)
    g = gather_adjoint(o, f, spacing, tr - obs, Q)
    assert 0.5 * float(np.sum((tr - obs) ** 2)) == pytest.approx(J, rel=1e-12)
    assert np.linalg.norm(g - g_ref) <= 1e-13 * np.linalg.norm(g_ref)
    o.close()
