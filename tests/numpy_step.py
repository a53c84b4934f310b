"""Vectorised numpy restatement of one reference time step (test helper).

An independent second restatement of forward! (src/shared/acoustic_3D_xPU.jl
:8-167 and the 2D twin), written with array slices instead of the C oracle's
per-node loops, used on tiny grids to cross-check the C oracle bit for bit.
Arrays are indexed [i, j(, k)] with 0-based node indices.
"""
import numpy as np


def step(st, it):
    """One forward! call (it is 1-based) plus the buffer rotation.

    st: dict with pold, pcur, pnew, fact, inv_d, inv_d2, halo, coef (per axis
    [a_l, a_r, bK_l, bK_r, a_hl, a_hr, bK_hl, bK_hr]), psi/xi (per axis
    [left, right]), possrcs/posrecs (1-based), dt2srctf, traces.
    """
    p = st["pcur"]
    nd = p.ndim
    n = p.shape
    h = st["halo"]
    # update_psi_*!: full ranges, left and right
    if h > 0:
        for a in range(nd):
            c = st["coef"][a]
            for side in range(2):
                psi = st["psi"][a][side]
                ah, bKh = (c[4], c[6]) if side == 0 else (c[5], c[7])
                shift = 0 if side == 0 else n[a] - h - 2
                lo = [slice(None)] * nd
                hi = [slice(None)] * nd
                lo[a] = slice(shift, shift + h + 1)
                hi[a] = slice(shift + 1, shift + h + 2)
                shape = [1] * nd
                shape[a] = h + 1
                dp = p[tuple(hi)] - p[tuple(lo)]
                psi[...] = bKh.reshape(shape) * psi + ah.reshape(shape) * dp * st["inv_d"][a]
    inner = tuple(slice(1, m - 1) for m in n)
    pc = p[inner]
    d2 = []
    for a in range(nd):
        plus = [slice(1, m - 1) for m in n]
        minus = [slice(1, m - 1) for m in n]
        plus[a] = slice(2, n[a])
        minus[a] = slice(0, n[a] - 2)
        d2.append((p[tuple(plus)] - 2.0 * pc + p[tuple(minus)]) * st["inv_d2"][a])
    f = st["fact"][inner]
    damp = np.zeros_like(pc)
    for a in range(nd):
        c = st["coef"][a]
        q = np.arange(1, n[a] - 1)  # 0-based node index along the axis
        left = q <= h
        right = (~left) & (q >= n[a] - h - 1)
        for side, mask in ((0, left), (1, right)):
            if not mask.any():
                continue
            sel = [slice(None)] * nd
            sel[a] = np.nonzero(mask)[0]
            nodes = q[mask]
            m = nodes if side == 0 else nodes - (n[a] - h - 2)
            psi = st["psi"][a][side]
            xi = st["xi"][a][side]
            aa, bK = (c[0], c[2]) if side == 0 else (c[1], c[3])
            other = [slice(1, mm - 1) for mm in n]
            shape = [1] * nd
            shape[a] = len(m)
            ps_here = [slice(1, mm - 1) for mm in n]
            ps_prev = [slice(1, mm - 1) for mm in n]
            ps_here[a] = m
            ps_prev[a] = m - 1
            dpsi = (psi[tuple(ps_here)] - psi[tuple(ps_prev)]) * st["inv_d"][a]
            d2a = d2[a][tuple(sel)]
            xsel = [slice(1, mm - 1) for mm in n]
            xsel[a] = m - 1
            xnew = bK[m - 1].reshape(shape) * xi[tuple(xsel)] + aa[m - 1].reshape(shape) * (d2a + dpsi)
            xi[tuple(xsel)] = xnew
            damp[tuple(sel)] = damp[tuple(sel)] + f[tuple(sel)] * (dpsi + xnew)
            del other
    lap = d2[0] + d2[1]
    if nd == 3:
        lap = lap + d2[2]
    st["pnew"][inner] = 2.0 * pc - st["pold"][inner] + f * lap + damp
    pn = st["pnew"]
    for s in range(st["possrcs"].shape[0]):
        idx = tuple(int(v) - 1 for v in st["possrcs"][s])
        pn[idx] += st["dt2srctf"][it - 1, s]
    for r in range(st["posrecs"].shape[0]):
        idx = tuple(int(v) - 1 for v in st["posrecs"][r])
        st["traces"][it - 1, r] = pn[idx]
    st["pold"], st["pcur"], st["pnew"] = st["pcur"], st["pnew"], st["pold"]


def from_oracle(o, spacing):
    """Initial state with the oracle problem's setup arrays."""
    n = o.n
    nd = len(n)
    h = o.halo
    st = {
        "pold": np.zeros(n), "pcur": np.zeros(n), "pnew": np.zeros(n),
        "fact": o.fact().copy(),
        "inv_d": [1.0 / d for d in spacing], "inv_d2": [1.0 / (d * d) for d in spacing],
        "halo": h,
        "coef": [[o.coef(a, w) for w in range(8)] for a in range(nd)],
        "dt2srctf": o.dt2srctf(),
        "traces": np.zeros((o.nt, o.nrecs)),
    }
    st["possrcs"], st["posrecs"] = o.positions()
    st["psi"], st["xi"] = [], []
    for a in range(nd):
        s1 = list(n)
        s1[a] = h + 1
        s2 = list(n)
        s2[a] = h
        st["psi"].append([np.zeros(s1), np.zeros(s1)])
        st["xi"].append([np.zeros(s2), np.zeros(s2)])
    return st
