"""Drop-in solve2D / solve3D over libawcpml, and the Plan wrapper.

``solve2D``/``solve3D`` keep the reference's entry points, keyword arguments,
outputs and errors (src/shared/acoustic_2D_xPU.jl:133-444,
src/shared/acoustic_3D_xPU.jl:169-544): the host setup runs in FP64 on the
CPU (``host.setup``), the time loop runs on the GPU through the C ABI, the
final pressure field is returned and ``recs.seismograms`` is filled.
"""
import ctypes
import os
import time

import numpy as np

from . import _lib
from .host import setup

DTYPES = {"f64": _lib.F64, "float64": _lib.F64, np.float64: _lib.F64,
          "f32": _lib.F32, "float32": _lib.F32, np.float32: _lib.F32}


def _dtype_code(dtype):
    try:
        return DTYPES[dtype]
    except (KeyError, TypeError):
        raise ValueError(f"unsupported dtype {dtype!r} (use 'f64' or 'f32')")


class Plan:
    """One velocity model with ``nshots`` independent shots on the GPU.

    Thin owner of an ``awcpml_plan`` handle; see include/awcpml.h for the
    semantics of each call.
    """

    def __init__(self, n, halo, dtype="f64", nshots=1):
        L = _lib.load()
        self.n = tuple(int(v) for v in n)
        self.ndim = len(self.n)
        self.halo = int(halo)
        self.nshots = int(nshots)
        self.dtype = _dtype_code(dtype)
        nz = self.n[2] if self.ndim == 3 else 1
        h = ctypes.c_void_p()
        _lib.check(L.awcpml_plan_create(self.ndim, self.n[0], self.n[1], nz, self.halo,
                                        self.dtype, self.nshots, ctypes.byref(h)))
        self._h = h
        self.nt = None

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().awcpml_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- configuration ------------------------------------------------
    def set_model(self, fact, inv_d, inv_d2):
        fact = np.asfortranarray(np.asarray(fact, dtype=np.float64))
        assert fact.shape == self.n, (fact.shape, self.n)
        inv_d = np.ascontiguousarray(inv_d, dtype=np.float64)
        inv_d2 = np.ascontiguousarray(inv_d2, dtype=np.float64)
        _lib.check(_lib.load().awcpml_set_model(self._h, fact.ctypes.data_as(_lib._dp),
                                                _lib.dptr(inv_d), _lib.dptr(inv_d2)))

    def set_cpml(self, axis, coefs):
        """coefs: [a_l, a_r, b_K_l, b_K_r, a_hl, a_hr, b_K_hl, b_K_hr]."""
        arrs = []
        for c, v in enumerate(coefs):
            v = np.ascontiguousarray(v, dtype=np.float64)
            need = self.halo + (1 if c >= 4 else 0)
            assert v.size == need, (c, v.size, need)
            arrs.append(v if v.size else np.zeros(1))
        _lib.check(_lib.load().awcpml_set_cpml(self._h, axis, *[_lib.dptr(a) for a in arrs]))

    def set_sources(self, shot, pos, dt2srctf):
        pos = np.asfortranarray(np.asarray(pos, dtype=np.int64).reshape(-1, self.ndim))
        wav = np.asfortranarray(np.asarray(dt2srctf, dtype=np.float64))
        nt = wav.shape[0]
        assert wav.shape[1] == pos.shape[0]
        _lib.check(_lib.load().awcpml_set_sources(self._h, shot, pos.shape[0],
                                                  pos.ctypes.data_as(_lib._i64p), nt,
                                                  wav.ctypes.data_as(_lib._dp)))
        self.nt = nt

    def set_receivers(self, shot, pos):
        pos = np.asfortranarray(np.asarray(pos, dtype=np.int64).reshape(-1, self.ndim))
        _lib.check(_lib.load().awcpml_set_receivers(self._h, shot, pos.shape[0],
                                                    pos.ctypes.data_as(_lib._i64p)))
        if not hasattr(self, "_nrecs"):
            self._nrecs = {}
        self._nrecs[shot] = pos.shape[0]

    def set_graph(self, enable=True):
        """Capture the time loop in HIP graphs (replayed on later solves)."""
        _lib.check(_lib.load().awcpml_set_graph(self._h, int(bool(enable))))

    def set_wide_offsets(self, enable=True):
        """Use the 64-bit-offset kernels (otherwise only for fields of
        4 GiB and more per shot); bit-identical results."""
        _lib.check(_lib.load().awcpml_set_wide_offsets(self._h, int(bool(enable))))

    def set_tiling(self, block_y=0, march_len=0, prefetch=0):
        _lib.check(_lib.load().awcpml_set_tiling(self._h, block_y, march_len, prefetch))

    def tiling(self):
        """(block_y, march_len, prefetch) of the stencil launches."""
        by, ml, pf = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.load().awcpml_get_tiling(self._h, ctypes.byref(by), ctypes.byref(ml),
                                                 ctypes.byref(pf)))
        return by.value, ml.value, pf.value

    def set_shot_group(self, shot_group):
        _lib.check(_lib.load().awcpml_set_shot_group(self._h, int(shot_group)))

    def shot_group(self):
        g = ctypes.c_int()
        _lib.check(_lib.load().awcpml_get_shot_group(self._h, ctypes.byref(g)))
        return g.value

    def set_time_block(self, steps, block_rows=0, march_len=0, prefetch=0):
        """steps = 2: two time steps per stencil launch where that applies;
        results are bit-identical to one step per launch."""
        _lib.check(_lib.load().awcpml_set_time_block(self._h, int(steps), int(block_rows),
                                                     int(march_len), int(prefetch)))

    def time_block(self):
        """(steps a launch advances now, block_rows, march_len, prefetch)."""
        v = [ctypes.c_int() for _ in range(4)]
        _lib.check(_lib.load().awcpml_get_time_block(self._h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_two_step_align(self, enable=True):
        """3D two-step launches: aligned 64-node x tiles for the blocks no
        CPML slab reaches (default on); results are bit-identical either way."""
        _lib.check(_lib.load().awcpml_set_two_step_align(self._h, int(bool(enable))))

    def two_step_align(self):
        v = ctypes.c_int()
        _lib.check(_lib.load().awcpml_get_two_step_align(self._h, ctypes.byref(v)))
        return bool(v.value)

    # -- execution ----------------------------------------------------
    def reset(self):
        _lib.check(_lib.load().awcpml_reset(self._h))

    def advance(self, nsteps):
        _lib.check(_lib.load().awcpml_advance(self._h, int(nsteps)))

    def forward(self, it):
        _lib.check(_lib.load().awcpml_forward(self._h, int(it)))

    def steps_done(self):
        v = ctypes.c_int64()
        _lib.check(_lib.load().awcpml_steps_done(self._h, ctypes.byref(v)))
        return v.value

    def traces(self, shot=0):
        nrec = self._nrecs[shot]
        out = np.zeros((self.nt, nrec), order="F")
        _lib.check(_lib.load().awcpml_get_traces(self._h, shot, out.ctypes.data_as(_lib._dp)))
        return out

    # -- differentiation ----------------------------------------------
    def checkpoint_run(self, every=0):
        """Reset, run all nt steps and keep a forward checkpoint every
        ``every`` steps (0: round(sqrt(nt))) for :meth:`adjoint`."""
        _lib.check(_lib.load().awcpml_checkpoint_run(self._h, int(every)))

    def set_gradient_fold(self, steps=0):
        """Reverse steps summed in the plan's precision before they are
        folded into the FP64 gradient total (0: one checkpoint segment, the
        default; -1: never)."""
        _lib.check(_lib.load().awcpml_set_gradient_fold(self._h, int(steps)))

    def set_adjoint_group(self, shots=0):
        """Shots per thread of the reverse step (0: automatic); the shots
        of a group share one gradient accumulator."""
        _lib.check(_lib.load().awcpml_set_adjoint_group(self._h, int(shots)))

    def adjoint_group(self):
        """Shots per thread the reverse step uses."""
        v = ctypes.c_int()
        _lib.check(_lib.load().awcpml_get_adjoint_group(self._h, ctypes.byref(v)))
        return v.value

    def adjoint(self, trace_bar):
        """dJ/dfact summed over the shots, for the trace cotangents
        ``trace_bar`` = dJ/dtraces: one (nt, nrecs) array per shot (a list,
        or a single array for a one-shot plan).  Needs a preceding
        :meth:`checkpoint_run` on the same model and acquisition."""
        if isinstance(trace_bar, np.ndarray) and trace_bar.ndim == 2:
            trace_bar = [trace_bar]
        if len(trace_bar) != self.nshots:
            raise ValueError(f"need {self.nshots} cotangent arrays, got {len(trace_bar)}")
        parts = []
        for s, t in enumerate(trace_bar):
            t = np.asarray(t, dtype=np.float64)
            if t.shape != (self.nt, self._nrecs[s]):
                raise ValueError(f"shot {s}: cotangent shape {t.shape} != {(self.nt, self._nrecs[s])}")
            parts.append(t.ravel(order="F"))
        flat = np.ascontiguousarray(np.concatenate(parts)) if parts else np.zeros(1)
        out = np.zeros(self.n, order="F")
        _lib.check(_lib.load().awcpml_adjoint(self._h, _lib.dptr(flat), flat.size if parts else 0,
                                              out.ctypes.data_as(_lib._dp)))
        return out

    def field(self, shot=0, which=1):
        out = np.zeros(self.n, order="F")
        _lib.check(_lib.load().awcpml_get_field(self._h, shot, which, out.ctypes.data_as(_lib._dp)))
        return out

    def time_stencil(self, nlaunches):
        ms = ctypes.c_double()
        _lib.check(_lib.load().awcpml_time_stencil(self._h, int(nlaunches), ctypes.byref(ms)))
        return ms.value

    def time_stencil2(self, nlaunches):
        """Milliseconds of `nlaunches` two-step stencil launches."""
        ms = ctypes.c_double()
        _lib.check(_lib.load().awcpml_time_stencil2(self._h, int(nlaunches), ctypes.byref(ms)))
        return ms.value

    def time_advance(self, nsteps):
        ms = ctypes.c_double()
        _lib.check(_lib.load().awcpml_time_advance(self._h, int(nsteps), ctypes.byref(ms)))
        return ms.value

    @classmethod
    def from_setup(cls, S, dtype="f64", nshots=1):
        """Plan with the model and CPML of a host ``Setup``; shot 0 gets its
        sources and receivers."""
        P = cls(S.n, S.halo, dtype=dtype, nshots=nshots)
        P.set_model(S.fact, S.inv_d, S.inv_d2)
        for a in range(S.ndim):
            P.set_cpml(a, S.coef[a])
        P.set_sources(0, S.possrcs, S.dt2srctf)
        P.set_receivers(0, S.posrecs)
        return P


class Batch:
    """Several plans (different grids/models, same ndim and dtype) advanced
    by one stencil launch and one source/receiver launch per time step.
    The plans stay usable for reading results (``traces``/``field``)."""

    def __init__(self, plans):
        L = _lib.load()
        self.plans = list(plans)
        arr = (ctypes.c_void_p * len(self.plans))(*[p._h for p in self.plans])
        h = ctypes.c_void_p()
        _lib.check(L.awcpml_batch_create(arr, len(self.plans), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().awcpml_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def reset(self):
        _lib.check(_lib.load().awcpml_batch_reset(self._h))

    def advance(self, nsteps):
        _lib.check(_lib.load().awcpml_batch_advance(self._h, int(nsteps)))

    def run(self):
        """Reset and run every plan to its own nt."""
        self.reset()
        self.advance(max(p.nt for p in self.plans))

    def set_graph(self, enable=True):
        _lib.check(_lib.load().awcpml_batch_set_graph(self._h, int(bool(enable))))

    def time_advance(self, nsteps):
        ms = ctypes.c_double()
        _lib.check(_lib.load().awcpml_batch_time_advance(self._h, int(nsteps), ctypes.byref(ms)))
        return ms.value

    def time_stencil(self, nlaunches):
        ms = ctypes.c_double()
        _lib.check(_lib.load().awcpml_batch_time_stencil(self._h, int(nlaunches), ctypes.byref(ms)))
        return ms.value

    def set_time_block(self, steps, block_rows=0, prefetch=0):
        """Two time steps of every plan per stencil launch (steps=2) while
        every plan has two steps left; each plan keeps its own march_len.
        Results are bit-identical to single steps."""
        _lib.check(_lib.load().awcpml_batch_set_time_block(self._h, int(steps), int(block_rows),
                                                           int(prefetch)))

    def time_block(self):
        """(steps per launch now, block_rows, prefetch)."""
        v = [ctypes.c_int() for _ in range(3)]
        _lib.check(_lib.load().awcpml_batch_get_time_block(self._h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def time_stencil2(self, nlaunches):
        """Milliseconds of nlaunches two-step launches alone (reset afterwards)."""
        ms = ctypes.c_double()
        _lib.check(_lib.load().awcpml_batch_time_stencil2(self._h, int(nlaunches), ctypes.byref(ms)))
        return ms.value


def _effective_bytes(n, halo, esize):
    """A_eff of the reference's do_bench report (acoustic_2D_xPU.jl:288-294,
    acoustic_3D_xPU.jl:341-349), in GB."""
    if len(n) == 2:
        nx, ny = n
        elems = ((halo + 1) * ny * 2 * 3 + (halo + 1) * nx * 2 * 3 + (halo + 1) * ny * 2 * 2
                 + (halo + 1) * nx * 2 * 2 + 4 * nx * ny)
    else:
        nx, ny, nz = n
        elems = ((halo + 1) * ny * nz * 2 * 3 + (halo + 1) * nx * nz * 2 * 3
                 + (halo + 1) * nx * ny * 2 * 3 + (halo + 1) * ny * nz * 2 * 2
                 + (halo + 1) * nx * nz * 2 * 2 + (halo + 1) * nx * ny * 2 * 2 + 4 * nx * ny * nz)
    return elems * esize / 1e9


def _bench(P, n, halo, esize, trials=200):
    """do_bench: minimum over repeated single forward! calls with it = 1
    (acoustic_2D_xPU.jl:261-299), printed in the reference's format."""
    times = []
    P.forward(1)
    for _ in range(trials):
        t0 = time.perf_counter()
        P.forward(1)
        times.append(time.perf_counter() - t0)
    t_it = min(times)
    a_eff = _effective_bytes(n, halo, esize)
    size = "x".join(str(v) for v in n)
    print("size = %s, time = %1.3e sec, Teff = %1.3f GB/s" % (size, t_it, a_eff / t_it))
    return t_it


def _solve(lengths, lt, vel, srcs, recs, halo, rcoef, ppw, freetop, do_bench, do_vis,
           dtype, do_save=False, nsave=5, save_name="acoustic3D_xPU", save_dir=None):
    if do_vis:
        raise NotImplementedError("do_vis (Plots.jl animations) is not part of this library")
    S = setup(lengths, lt, vel, srcs, recs, halo=halo, rcoef=rcoef, ppw=ppw, freetop=freetop)
    P = Plan.from_setup(S, dtype=dtype)
    try:
        if do_bench:
            _bench(P, S.n, S.halo, 8 if P.dtype == _lib.F64 else 4)
            return None
        if do_save:
            save_dir = save_dir or os.path.join(os.getcwd(), "simulations", "tmp")
            os.makedirs(save_dir, exist_ok=True)
            done = 0
            while done < S.nt:
                step = min(nsave - (done % nsave), S.nt - done)
                P.advance(step)
                done += step
                if done % nsave == 0:
                    _save_snapshot(save_dir, save_name, done, lengths, S, P.field(0, 1))
        else:
            P.advance(S.nt)
        recs.seismograms = P.traces(0)
        return P.field(0, 1)
    finally:
        P.close()


def _save_snapshot(save_dir, save_name, it, lengths, S, pcur):
    """do_save snapshot of step ``it`` (acoustic_3D_xPU.jl:480-498): the
    reference's seven datasets with its names, shapes, element types and
    index order -- lx, ly, lz (Float64 scalars), halo (Int64 scalar), pcur
    (Float64, nx x ny x nz, indexed [i, j, k]), possrcs and posrecs (Int64,
    count x 3, 1-based grid positions) -- in ``<save_name>_it<it>.npz``
    instead of the reference's ``.h5`` (no HDF5 library in this Python
    environment; the Julia backend writes the HDF5 file itself).  An
    existing file of that name is replaced, as in the reference."""
    path = os.path.join(save_dir, f"{save_name}_it{it}.npz")
    if os.path.exists(path):
        os.remove(path)
    np.savez(path, lx=np.float64(lengths[0]), ly=np.float64(lengths[1]),
             lz=np.float64(lengths[2] if len(lengths) > 2 else 0.0), halo=np.int64(S.halo),
             pcur=np.asarray(pcur, dtype=np.float64), possrcs=np.asarray(S.possrcs, dtype=np.int64),
             posrecs=np.asarray(S.posrecs, dtype=np.int64))
    return path


def solve2D(lx, ly, lt, vel, srcs, recs, *, halo=20, rcoef=0.0001, ppw=10.0, freetop=True,
            do_bench=False, do_vis=False, nvis=5, gif_name="acoustic2D_xPU", plims=(-1, 1),
            threshold=0.01, dtype="f64"):
    """2D CPML acoustic solve (src/shared/acoustic_2D_xPU.jl:133-444).

    Returns the final pressure field (nx, ny) and fills ``recs.seismograms``
    with the (nt, nrecs) traces.  ``dtype="f32"`` runs the FP32 extension.
    """
    vel = np.asarray(vel)
    assert vel.ndim == 2
    return _solve((lx, ly), lt, vel, srcs, recs, halo, rcoef, ppw, freetop, do_bench, do_vis,
                  dtype)


def solve3D(lx, ly, lz, lt, vel, srcs, recs, *, halo=20, rcoef=0.0001, ppw=10.0, freetop=True,
            do_bench=False, do_vis=False, do_save=False, nvis=5, nsave=5,
            gif_name="acoustic3D_xPU_slice", save_name="acoustic3D_xPU", plims=(-1.0, 1.0),
            threshold=0.01, dtype="f64", save_dir=None):
    """3D CPML acoustic solve (src/shared/acoustic_3D_xPU.jl:169-544).

    Returns the final pressure field (nx, ny, nz) and fills
    ``recs.seismograms``.  ``do_save`` writes a snapshot of pcur every
    ``nsave`` steps to ``save_dir`` (default ./simulations/tmp).
    """
    vel = np.asarray(vel)
    assert vel.ndim == 3
    return _solve((lx, ly, lz), lt, vel, srcs, recs, halo, rcoef, ppw, freetop, do_bench,
                  do_vis, dtype, do_save=do_save, nsave=nsave, save_name=save_name,
                  save_dir=save_dir)
