"""Physics of the oracle against closed-form Green's functions.

Restates the reference's analytic tests (test/test2D_xPU.jl:12-106,
test/test3D_xPU.jl:12-108): a Ricker point source in a homogeneous medium,
one receiver; the integrated absolute error of the trace must stay below 1%
of the peak analytic amplitude times the simulated time.  With halo = 0 the
window ends before any boundary reflection returns; with halo = 20 the run
is long enough that the CPML must absorb the outgoing wave.
"""
import math

import numpy as np
import pytest

import oracle


def _run(n, L, lt, c0, f0, src, rec, halo):
    vel = np.full(n, c0, order="F")
    o = oracle.OracleProblem(n, L, lt, vel, [src], [4 / f0], f0, [rec], halo=halo,
                             rcoef=0.0001, freetop=False)
    o.run()
    return o, o.traces()[:, 0]


def _check(trace, times, Gc, lt):
    assert len(trace) == len(Gc)
    err = np.trapezoid(np.abs(trace - Gc), times[1:])
    assert err <= np.abs(Gc).max() * 0.01 * lt, (err, np.abs(Gc).max() * 0.01 * lt)


def _ricker(t, t0, f0):
    s = math.pi * f0 * (t - t0)
    return (1 - 2 * s * s) * np.exp(-s * s)


@pytest.mark.parametrize("halo,lt", [(0, 10.0), (20, 20.0)])
def test_green_2d(halo, lt):
    n = (501, 501)
    L = (10.0, 10.0)
    c0, f0 = 1.0, 1.0
    src, rec = (5.0, 5.0), (2.5, 5.0)
    o, trace = _run(n, L, lt, c0, f0, src, rec, halo)
    dx = L[0] / (n[0] - 1)
    dt = math.sqrt(2) / (c0 * (1 / dx + 1 / dx)) / 2
    nt = math.ceil(lt / dt)
    assert o.nt == nt
    times = np.arange(nt + 1) * dt
    dist = math.hypot(src[0] - rec[0], src[1] - rec[1])
    G = np.zeros(nt + 1)
    live = times - dist / c0 >= 0
    G[live] = 1.0 / (2 * math.pi * c0 ** 2 * np.sqrt(times[live] ** 2 - dist ** 2 / c0 ** 2))
    Gc = np.convolve(G, _ricker(times, 4 / f0, f0) * dt)[1:nt + 1]
    _check(trace, times, Gc, lt)


@pytest.mark.parametrize("halo,lt", [(0, 26.0), (20, 50.0)])
def test_green_3d(halo, lt):
    n = (101, 101, 101)
    L = (20.0, 20.0, 20.0)
    c0, f0 = 1.0, 0.25
    src, rec = (10.0, 10.0, 10.0), (5.0, 10.0, 10.0)
    o, trace = _run(n, L, lt, c0, f0, src, rec, halo)
    dx = L[0] / (n[0] - 1)
    dt = math.sqrt(3) / (c0 * (3 / dx)) / 2
    nt = math.ceil(lt / dt)
    assert o.nt == nt
    times = np.arange(nt + 1) * dt
    dist = 5.0
    G = np.zeros(nt + 1)
    first = int(np.argmax(times - dist / c0 >= 0))
    G[first] = 1.0 / (4 * math.pi * c0 ** 2 * dist)
    Gc = np.convolve(G, _ricker(times, 4 / f0, f0))[1:nt + 1]
    _check(trace, times, Gc, lt)


def test_cpml_absorbs_2d():
    """With halo = 20 the field energy after the wave has left the model is
    a small fraction of its peak; with a wrong CPML sign or index the
    reflections (or an instability) keep it large."""
    n = (201, 201)
    L = (2000.0, 2000.0)
    vel = np.full(n, 2000.0, order="F")
    o = oracle.OracleProblem(n, L, 2.0, vel, [(1000.0, 1000.0)], [0.4], 10.0,
                             [(1000.0, 1500.0)], halo=20, rcoef=0.0001, freetop=False)
    peak = 0.0
    while o.it_done < o.nt:
        o.run(50)
        peak = max(peak, float(np.sqrt((o.pend() ** 2).sum())))
    final = float(np.sqrt((o.pend() ** 2).sum()))
    assert final < 1e-3 * peak, (final, peak)
