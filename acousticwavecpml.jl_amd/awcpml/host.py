"""Host-side FP64 setup of the CPML solvers, mirroring the reference.

Everything here runs once per solve on the CPU in float64, exactly where the
reference does it before its time loop; the per-step work is in libawcpml.

  Sources, Receivers ...... src/utils.jl:88-145
  rickersource1D .......... src/utils.jl:64-66
  calc_Kab_CPML ........... src/utils.jl:19-57
  setup()  ................ src/shared/acoustic_2D_xPU.jl:151-240,
                            src/shared/acoustic_3D_xPU.jl:191-285
"""
import math

import numpy as np


class Sources:
    """Multi-source configuration of one shot (src/utils.jl:88-121).

    positions: (nsrcs, ndim) source coordinates in meters, one row per source.
    t0s: activation time of each source [s].
    srctfs: one callable per source, f(t, t0, f0) -> value (t may be an array).
    freqdomain: dominating frequency shared by all sources [Hz].
    """

    def __init__(self, positions, t0s, srctfs, freqdomain):
        positions = np.atleast_2d(np.asarray(positions, dtype=np.float64))
        assert positions.shape[0] > 0, "There must be at least one source!"
        assert positions.shape[0] == len(t0s) == len(srctfs), \
            "Number of sources do not match between positions and time-functions!"
        self.n = positions.shape[0]
        self.positions = positions
        self.t0s = [float(t) for t in t0s]
        self.srctfs = list(srctfs)
        self.freqdomain = float(freqdomain)


class Receivers:
    """Receiver configuration of one shot (src/utils.jl:126-145).

    ``seismograms`` is filled by the solvers with an (nt, nrecs) array, one
    column per receiver in the order of ``positions``.
    """

    def __init__(self, positions):
        positions = np.atleast_2d(np.asarray(positions, dtype=np.float64))
        assert positions.shape[0] > 0, "There must be at least one receiver!"
        self.n = positions.shape[0]
        self.positions = positions
        self.seismograms = np.zeros((0, 0))


def _exp(x):
    """Element-wise exp with the C library's exp (the same routine on every
    element, independent of numpy's SIMD paths)."""
    x = np.asarray(x, dtype=np.float64)
    return np.array([math.exp(v) for v in x.ravel()], dtype=np.float64).reshape(x.shape)


def rickersource1D(t, t0, f0):
    """Ricker wavelet (src/utils.jl:64-66); t may be a scalar or an array."""
    t = np.asarray(t, dtype=np.float64)
    s = math.pi * f0 * (t - t0)
    s2 = s * s
    out = (1.0 - 2.0 * s2) * _exp(-s2)
    return out if out.ndim else float(out)


def _linrange(start, stop, length):
    """Elements of Julia's LinRange(start, stop, length): (1-t)*a + t*b."""
    lendiv = max(length - 1, 1)
    t = np.arange(length, dtype=np.float64) / float(lendiv)
    return (1.0 - t) * start + t * stop


def calc_Kab_CPML(halo, dt, npower, d0, alpha_max_pml, K_max_pml, onwhere):
    """CPML coefficients (src/utils.jl:19-57).

    onwhere "ongrd": nodal coefficients (halo values); "halfgrd": staggered
    coefficients between nodes (halo+1 values).  Returns
    (a_left, a_right, b_K_left, b_K_right).
    """
    kab_size = int(halo)
    if onwhere == "halfgrd":
        kab_size += 1
        shift = 0.5
    elif onwhere == "ongrd":
        shift = 0.0
    else:
        raise RuntimeError("Wrong onwhere parameter!")
    dist = _linrange(0.0 - shift, (kab_size - shift) - 1.0, kab_size)
    if onwhere == "halfgrd" and kab_size > 0:
        dist[0] = 0.0
    out = []
    for nd in (dist[::-1] / halo if halo else dist, dist / halo if halo else dist):
        ndp = nd * nd if npower == 2.0 else nd ** npower
        d = d0 * ndp
        alpha = alpha_max_pml * (1.0 - nd)
        K = 1.0 + (K_max_pml - 1.0) * ndp
        b = _exp((-(d / K + alpha)) * dt)
        a = d * (b - 1.0) / (K * (d + K * alpha))
        out.append((a, b / K))
    (a_l, bK_l), (a_r, bK_r) = out
    return a_l, a_r, bK_l, bK_r


def round_ties_up(x):
    """round(Int, x, RoundNearestTiesUp) for finite coordinates."""
    t = math.floor(x)
    return int(t + 1 if x - t >= 0.5 else t)


class Setup:
    """Everything solve2D/solve3D compute before the time loop."""


def setup(lengths, lt, vel, srcs, recs, halo=20, rcoef=0.0001, ppw=10.0, freetop=True):
    """Host FP64 setup of one shot (acoustic_2D_xPU.jl:151-240 /
    acoustic_3D_xPU.jl:191-285).  Raises AssertionError on the reference's
    assertions (points per wavelength, sources/receivers inside the model)."""
    vel = np.asarray(vel)
    ndim = vel.ndim
    assert ndim in (2, 3), "vel must be 2D or 3D"
    assert len(lengths) == ndim
    S = Setup()
    S.ndim = ndim
    S.n = tuple(int(v) for v in vel.shape)
    S.halo = int(halo)
    f0 = srcs.freqdomain
    vel_max = float(np.max(vel))
    S.d = tuple(float(lengths[a]) / (S.n[a] - 1) for a in range(ndim))
    if ndim == 2:
        dx, dy = S.d
        S.dt = math.sqrt(2) / (vel_max * (1 / dx + 1 / dy)) / 2
    else:
        dx, dy, dz = S.d
        S.dt = math.sqrt(3) / (vel_max * (1 / dx + 1 / dy + 1 / dz)) / 2
    dt = S.dt
    S.nt = int(math.ceil(lt / dt))
    S.times = np.arange(S.nt, dtype=np.float64) * dt
    alpha_max = math.pi * f0
    npower = 2.0
    K_max = 1.0
    S.coef = []
    for a in range(ndim):
        thickness = S.halo * S.d[a]
        d0 = -(npower + 1) * vel_max * math.log(rcoef) / (2.0 * thickness) if S.halo else 0.0
        a_l, a_r, bK_l, bK_r = calc_Kab_CPML(S.halo, dt, npower, d0, alpha_max, K_max, "ongrd")
        a_hl, a_hr, bK_hl, bK_hr = calc_Kab_CPML(S.halo, dt, npower, d0, alpha_max, K_max, "halfgrd")
        S.coef.append([a_l, a_r, bK_l, bK_r, a_hl, a_hr, bK_hl, bK_hr])
    if freetop:  # the top is the y-low side (acoustic_2D_xPU.jl:180-185)
        c = S.coef[1]
        c[0] = np.zeros_like(c[0])
        c[4] = np.zeros_like(c[4])
        c[2] = np.ones_like(c[2])
        c[6] = np.ones_like(c[6])
    S.inv_d = np.array([1.0 / h for h in S.d])
    S.inv_d2 = np.array([1.0 / (h * h) for h in S.d])
    v = vel.astype(np.float64)
    S.fact = np.asfortranarray((dt * dt) * (v * v))
    h2 = sum(h * h for h in S.d) if ndim == 3 else S.d[0] * S.d[0] + S.d[1] * S.d[1]
    assert math.sqrt(h2) <= vel_max / (ppw * f0), "Not enough points per wavelength!"
    cell = S.d[0] * S.d[1] if ndim == 2 else S.d[0] * S.d[1] * S.d[2]
    scale = (dt * dt) / cell
    S.dt2srctf = np.zeros((S.nt, srcs.n), order="F")
    for s in range(srcs.n):
        S.dt2srctf[:, s] = scale * np.asarray(srcs.srctfs[s](S.times, srcs.t0s[s], f0),
                                              dtype=np.float64)
    S.possrcs = _snap(srcs.positions, S.d, ndim)
    assert all(1 <= S.possrcs[s, a] <= S.n[a] for s in range(srcs.n) for a in range(ndim)), \
        "At least one source is not inside the model!"
    S.posrecs = _snap(recs.positions, S.d, ndim)
    assert all(1 <= S.posrecs[r, a] <= S.n[a] for r in range(recs.n) for a in range(ndim)), \
        "At least one receiver is not inside the model!"
    return S


def _snap(positions, d, ndim):
    out = np.zeros((positions.shape[0], ndim), dtype=np.int64, order="F")
    for r in range(positions.shape[0]):
        for a in range(ndim):
            out[r, a] = round_ties_up(positions[r, a] / d[a] + 1)
    return out
