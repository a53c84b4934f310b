"""Problem definitions shared by the tests, the smoke check and bench.py.

Each case is a plain dict of the reference's solve2D/solve3D arguments
(positions in meters, rows = items).  The workloads follow the reference's
own tests (test/test2D_xPU.jl, test/test3D_xPU.jl) and the benchmark
configurations:

  plain2d   2D CPML, 256x256 interior + 20-point halo on all sides
            (298x298 nodes), dx = 10 m, c = 2000 m/s, R = 1e-3, f0 = 8 Hz,
            source at the centre, 8 receivers 20 points below the top
            halo, nt = 1000.
  plain3d   3D CPML, 256^3 interior + 20-point halo, free top, dx = 10 m,
            c(depth) = 1500 + 0.6 depth, +-5 % seeded smoothed perturbation
            (sigma = 5 points), f0 = 10 Hz, source at 0-based (149, 30, 149),
            64 receivers on an 8x8 grid 5 nodes below the surface, nt = 2000.

Depth is the reference's second index (y); the free top is the y-low side.
n = interior + 2*halo + 2 per axis, lengths = (n-1)*dx, and
lt = (nt - 0.5)*dt reproduces nt = ceil(lt/dt) exactly.
"""
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def cfl_dt(ndim, d, vmax):
    if ndim == 2:
        return math.sqrt(2) / (vmax * (1 / d[0] + 1 / d[1])) / 2
    return math.sqrt(3) / (vmax * (1 / d[0] + 1 / d[1] + 1 / d[2])) / 2


def lt_for_nt(ndim, d, vmax, nt):
    return (nt - 0.5) * cfl_dt(ndim, d, vmax)


def interior_grid(interior, halo, count):
    """0-based node indices of `count` points evenly spread over the interior."""
    first = halo + 1
    return [first + int((k + 0.5) * interior / count) for k in range(count)]


def load_fixtures():
    with open(os.path.join(GOLDEN, "fixtures.json")) as fh:
        meta = json.load(fh)
    out = {}
    for name, m in meta.items():
        ref = np.fromfile(os.path.join(GOLDEN, m["file"]), dtype="<f8").reshape(m["shape"], order="F")
        case = fixture_case(m)
        case["name"] = name
        out[name] = (case, ref)
    return out


def fixture_case(m):
    n = m["n"]
    if m["vel"]["kind"] == "constant":
        vel = np.full(n, m["vel"]["c"], order="F")
    else:  # test/test2D_xPU.jl:160-169
        vel = np.zeros(n, order="F")
        for j in range(n[1]):
            vel[:, j] = m["vel"]["v0"] + m["vel"]["dv"] * j
        jc = n[1] - 1 - m["vel"]["const_from_end"]
        vel[:, jc:] = vel[0, jc]
    return dict(n=n, l=m["l"], lt=m["lt"], vel=vel, srcpos=np.array(m["srcpos"]),
                t0s=[m["t0"]] * len(m["srcpos"]), f0=m["f0"], recpos=np.array(m["recpos"]),
                halo=m["halo"], rcoef=m["rcoef"], freetop=m["freetop"], ppw=10.0)


def plain2d(nt=1000, interior=256, halo=20, dx=10.0, c=2000.0):
    n = interior + 2 * halo + 2
    d = (dx, dx)
    L = (n - 1) * dx
    rec_i = interior_grid(interior, halo, 8)
    rec_j = halo + 1 + 20
    return dict(n=[n, n], l=[L, L], lt=lt_for_nt(2, d, c, nt), vel=np.full((n, n), c, order="F"),
                srcpos=np.array([[L / 2, L / 2]]), t0s=[4 / 8.0], f0=8.0,
                recpos=np.array([[i * dx, rec_j * dx] for i in rec_i]),
                halo=halo, rcoef=1e-3, freetop=False, ppw=10.0, nt=nt)


def smooth_random(shape, sigma, seed):
    """Seeded N(0,1) field, Gaussian-smoothed (sigma in points), max|g| = 1."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.Generator(np.random.PCG64(seed))
    g = rng.standard_normal(shape)
    g = gaussian_filter(g, sigma=sigma, mode="nearest")
    return g / np.abs(g).max()


def plain3d(nt=2000, interior=256, halo=20, dx=10.0, seed=2024, nrec_side=8):
    n = interior + 2 * halo + 2
    d = (dx, dx, dx)
    L = (n - 1) * dx
    depth = np.arange(n) * dx
    base = (1500.0 + 0.6 * depth)[None, :, None]
    vel = np.asfortranarray(base * (1.0 + 0.05 * smooth_random((n, n, n), 5.0, seed)))
    vmax = float(vel.max())
    src = [(halo + 1 + interior // 2) * dx, 30 * dx, (halo + 1 + interior // 2) * dx]
    grid = interior_grid(interior, halo, nrec_side)
    rec = [[i * dx, 5 * dx, k * dx] for k in grid for i in grid]
    return dict(n=[n, n, n], l=[L, L, L], lt=lt_for_nt(3, d, vmax, nt), vel=vel,
                srcpos=np.array([src]), t0s=[4 / 10.0], f0=10.0, recpos=np.array(rec),
                halo=halo, rcoef=1e-3, freetop=True, ppw=10.0, nt=nt)


def small3d(n=(37, 29, 23), halo=5, freetop=True, nt=120, seed=7, nsrc=1):
    """Tiny heterogeneous 3D case for exact comparisons."""
    dx = 10.0
    d = (dx, dx, dx)
    rng = np.random.Generator(np.random.PCG64(seed))
    vel = np.asfortranarray(1800.0 + 400.0 * rng.random(n))
    vmax = float(vel.max())
    L = [(v - 1) * dx for v in n]
    srcs = [[(n[0] // 2) * dx, (halo + 3) * dx, (n[2] // 2) * dx]]
    for s in range(1, nsrc):
        srcs.append([((n[0] // 2) + s) * dx, (halo + 3) * dx, (n[2] // 3) * dx])
    recs = [[(halo + 2 + 3 * r) * dx, (halo + 2) * dx, (n[2] // 2 + r - 2) * dx] for r in range(5)]
    recs.append([0.0, 0.0, 0.0])  # a Dirichlet corner node
    return dict(n=list(n), l=L, lt=lt_for_nt(3, d, vmax, nt), vel=vel, srcpos=np.array(srcs),
                t0s=[0.02] * len(srcs), f0=25.0, recpos=np.array(recs), halo=halo, rcoef=1e-4,
                freetop=freetop, ppw=2.0, nt=nt)


def overlap_case(n, halo, freetop=False, nt=150, seed=5):
    """Tiny heterogeneous 2D or 3D case for axes with n < 2*halo + 2, where
    the two CPML slabs overlap and the low side's branch wins (the
    reference's if/elseif, acoustic_2D_xPU.jl:45-56); n >= halo + 2."""
    nd = len(n)
    dx = 10.0
    rng = np.random.Generator(np.random.PCG64(seed))
    vel = np.asfortranarray(1800.0 + 400.0 * rng.random(n))
    L = [(v - 1) * dx for v in n]
    srcs = [[(v // 2) * dx for v in n], [(1 + (3 * a) % (v - 2)) * dx for a, v in enumerate(n)]]
    recs = [[(1 + (7 * r + 3 * a) % (v - 2)) * dx for a, v in enumerate(n)] for r in range(6)]
    recs.append([0.0] * nd)  # a Dirichlet corner node
    return dict(n=list(n), l=L, lt=lt_for_nt(nd, (dx,) * nd, float(vel.max()), nt), vel=vel,
                srcpos=np.array(srcs), t0s=[0.02, 0.03], f0=25.0, recpos=np.array(recs),
                halo=halo, rcoef=1e-4, freetop=freetop, ppw=2.0, nt=nt)


# (n, halo, freetop): overlapping slabs on some axes, including n = halo + 2
# (no node takes the high side) and n = 2*halo + 2 (the slabs touch)
OVERLAP_CASES = [
    ((30, 33), 20, False),
    ((22, 40), 20, True),
    ((24, 23, 26), 12, False),
    ((30, 21, 19), 12, True),
]


def small2d(n=(57, 43), halo=6, freetop=True, nt=300, seed=3):
    dx = 10.0
    d = (dx, dx)
    rng = np.random.Generator(np.random.PCG64(seed))
    vel = np.asfortranarray(1800.0 + 400.0 * rng.random(n))
    vmax = float(vel.max())
    L = [(v - 1) * dx for v in n]
    srcs = [[(n[0] // 2) * dx, (halo + 4) * dx], [(n[0] // 3) * dx, 0.0]]  # one on the top row
    recs = [[(halo + 2 + 4 * r) * dx, (halo + 2) * dx] for r in range(6)] + [[(n[0] // 3) * dx, 0.0]]
    return dict(n=list(n), l=L, lt=lt_for_nt(2, d, vmax, nt), vel=vel, srcpos=np.array(srcs),
                t0s=[0.03, 0.05], f0=20.0, recpos=np.array(recs), halo=halo, rcoef=1e-4,
                freetop=freetop, ppw=2.0, nt=nt)


def oracle_problem(case):
    import oracle
    return oracle.OracleProblem(case["n"], case["l"], case["lt"], case["vel"], case["srcpos"],
                                case["t0s"], case["f0"], case["recpos"], halo=case["halo"],
                                rcoef=case["rcoef"], ppw=case.get("ppw", 10.0),
                                freetop=case["freetop"])


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))


def batch2d(nt=3000, interior=1024, halo=20, dx=10.0, nshots=64, nrec=256, seed=3):
    """Many shots on one 2D model (1066^2 nodes): layered gradient plus
    seeded smoothed anomalies (+-8 %, sigma = 20 points), 64 Ricker shots
    (f0 = 12 Hz) at 0-based depth 21 spread evenly along x, 256 receivers
    at the same depth.  `shots` holds one source position per shot."""
    n = interior + 2 * halo + 2
    d = (dx, dx)
    L = (n - 1) * dx
    depth = np.arange(n) * dx
    layer = np.minimum(3, (4 * np.arange(n)) // n)           # four layers by depth
    base = (1800.0 + 500.0 * layer + 0.05 * depth)[None, :]
    vel = np.asfortranarray(base * (1.0 + 0.08 * smooth_random((n, n), 20.0, seed)))
    vmax = float(vel.max())
    zs = (halo + 1) * dx
    shots = [np.array([[i * dx, zs]]) for i in interior_grid(interior, halo, nshots)]
    recs = np.array([[i * dx, zs] for i in interior_grid(interior, halo, nrec)])
    return dict(n=[n, n], l=[L, L], lt=lt_for_nt(2, d, vmax, nt), vel=vel, srcpos=shots[0],
                shots=shots, t0s=[1.2 / 12.0], f0=12.0, recpos=recs, halo=halo, rcoef=1e-3,
                freetop=False, ppw=10.0, nt=nt)


def uneven3d(nt=1500, halo=20, dx=10.0, count=12, seed=11, edges=(128, 192, 256, 384, 512)):
    """`count` independent 3D models whose interior cube edges are drawn
    uniformly (seeded) from `edges`; velocity 1500 m/s at the top rising to
    4000 m/s at the bottom plus a +-500 m/s smooth analytic anomaly pattern
    (range within 1500-4500 m/s); one Ricker shot (f0 = 10 Hz) each, 16
    receivers on a line.  Returns a list of case dicts (built lazily: each
    holds a `make_vel` callable instead of the array, to keep host memory
    at one model at a time)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for m in range(count):
        e = int(rng.choice(edges))
        n = e + 2 * halo + 2
        d = (dx, dx, dx)
        L = (n - 1) * dx
        phase = float(rng.uniform(0, 2 * np.pi))

        def make_vel(n=n, phase=phase):
            x = np.linspace(0.0, 1.0, n)
            base = 1500.0 + 2500.0 * x
            an = 500.0 * (np.sin(3 * np.pi * x + phase)[:, None, None] *
                          np.sin(2 * np.pi * x)[None, :, None] *
                          np.cos(3 * np.pi * x + 0.5 * phase)[None, None, :])
            v = base[None, :, None] + an
            return np.asfortranarray(np.clip(v, 1500.0, 4500.0))

        c = (halo + 1 + e // 2) * dx
        rec = [[(halo + 1 + (k + 0.5) * e / 16) * dx, (halo + 1 + 5) * dx, c] for k in range(16)]
        # lt follows from the built model: lt_for_nt(3, (dx,)*3, vel.max(), nt)
        out.append(dict(n=[n, n, n], l=[L, L, L], lt=None, make_vel=make_vel, srcpos=np.array([[c, (halo + 1 + 10) * dx, c]]), t0s=[0.12],
                        f0=10.0, recpos=np.array(rec), halo=halo, rcoef=1e-3, freetop=False,
                        ppw=10.0, nt=nt))
    return out


def fwi2d(nt=2500, interior=512, halo=20, dx=10.0, nshots=32, nrec=256, every=50):
    """2D FWI gradient case (554^2 nodes): true model 2000 m/s with a +10 %
    Gaussian anomaly (sigma = 30 points) at the centre, starting model
    2000 m/s; 32 Ricker shots (f0 = 8 Hz) and 256 receivers at 0-based
    depth 21; forward checkpoints every 50 steps."""
    n = interior + 2 * halo + 2
    d = (dx, dx)
    L = (n - 1) * dx
    c = (n - 1) / 2
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    r2 = (ii - c) ** 2 + (jj - c) ** 2
    true = np.asfortranarray(2000.0 * (1.0 + 0.1 * np.exp(-r2 / (2 * 30.0 ** 2))))
    start = np.full((n, n), 2000.0, order="F")
    zs = (halo + 1) * dx
    shots = [np.array([[i * dx, zs]]) for i in interior_grid(interior, halo, nshots)]
    recs = np.array([[i * dx, zs] for i in interior_grid(interior, halo, nrec)])
    # dt from the true model's vmax (held fixed while the model changes)
    return dict(n=[n, n], l=[L, L], lt=lt_for_nt(2, d, float(true.max()), nt), vel=true,
                start=start, srcpos=shots[0], shots=shots, t0s=[1.2 / 8.0], f0=8.0, recpos=recs,
                halo=halo, rcoef=1e-3, freetop=False, ppw=10.0, nt=nt, every=every)
