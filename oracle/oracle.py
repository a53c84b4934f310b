"""ctypes front end of the CPU FP64 oracle (awcpml_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product path never
imports this module.

``OracleProblem`` mirrors the reference's solve2D/solve3D
(src/shared/acoustic_2D_xPU.jl:133-444, src/shared/acoustic_3D_xPU.jl:169-544)
for Ricker sources: the constructor does the host setup, ``run`` the time
loop, ``traces``/``pend`` the outputs.  Arrays are numpy, indexed [i, j(, k)]
and stored column-major (x fastest) like Julia's.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ERRORS = {
    -1: "invalid argument",
    -2: "Not enough points per wavelength!",
    -3: "At least one source is not inside the model!",
    -4: "At least one receiver is not inside the model!",
    -5: "Wrong onwhere parameter!",
    -6: "out of memory",
}

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_long)


def build(force=False):
    """Compile liboracle.so with its Makefile (gcc, -ffp-contract=off)."""
    if force and os.path.exists(LIB_PATH):
        os.remove(LIB_PATH)
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_create.restype = ctypes.c_void_p
        L.oracle_create.argtypes = [ctypes.c_int, _lp, _dp, ctypes.c_double, _dp,
                                    ctypes.c_int, _dp, _dp, ctypes.c_double,
                                    ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_int)]
        L.oracle_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_step.argtypes = [ctypes.c_void_p, ctypes.c_long]
        L.oracle_run.argtypes = [ctypes.c_void_p, ctypes.c_long]
        L.oracle_run.restype = ctypes.c_long
        L.oracle_nt.argtypes = [ctypes.c_void_p]
        L.oracle_nt.restype = ctypes.c_long
        L.oracle_it_done.argtypes = [ctypes.c_void_p]
        L.oracle_it_done.restype = ctypes.c_long
        L.oracle_dt.argtypes = [ctypes.c_void_p]
        L.oracle_dt.restype = ctypes.c_double
        for name in ("oracle_get_traces", "oracle_get_fact", "oracle_get_dt2srctf"):
            getattr(L, name).argtypes = [ctypes.c_void_p, _dp]
        L.oracle_get_field.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.oracle_get_coef.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        L.oracle_get_positions.argtypes = [ctypes.c_void_p, _lp, _lp]
        L.oracle_get_psi.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        L.oracle_get_xi.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        L.oracle_calc_kab_cpml.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int, _dp, _dp, _dp, _dp]
        L.oracle_calc_kab_cpml.restype = ctypes.c_int
        L.oracle_ricker.argtypes = [ctypes.c_double] * 3
        L.oracle_ricker.restype = ctypes.c_double
        L.oracle_round_ties_up.argtypes = [ctypes.c_double]
        L.oracle_round_ties_up.restype = ctypes.c_long
        L.oracle_update_psi_2d.argtypes = ([ctypes.c_long, ctypes.c_long, ctypes.c_int,
                                            ctypes.c_double, ctypes.c_double] + [_dp] * 13)
        L.oracle_update_psi_3d.argtypes = ([ctypes.c_long] * 3 + [ctypes.c_int]
                                           + [ctypes.c_double] * 3 + [_dp] * 19)
        L.oracle_reset.argtypes = [ctypes.c_void_p]
        L.oracle_set_fact.argtypes = [ctypes.c_void_p, _dp]
        L.oracle_fwi_gradient.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.oracle_fwi_gradient.restype = ctypes.c_int
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_set_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


class OracleError(AssertionError):
    """Raised for the reference's assertion failures (AssertionError)."""


def set_threads(n):
    """OpenMP threads of the oracle (n <= 0: the default); returns the
    count in effect."""
    return lib().oracle_set_threads(int(n))


def calc_kab_cpml(halo, dt, npower, d0, alpha_max, K_max, onwhere):
    """calc_Kab_CPML (src/utils.jl:19-57) -> (a_l, a_r, b_K_l, b_K_r)."""
    code = {"ongrd": 0, "halfgrd": 1}.get(onwhere, 2)
    n = halo + (1 if onwhere == "halfgrd" else 0)
    out = [np.zeros(max(n, 1)) for _ in range(4)]
    st = lib().oracle_calc_kab_cpml(halo, dt, npower, d0, alpha_max, K_max, code,
                                    *[_p(o) for o in out])
    if st != 0:
        raise ValueError(ERRORS.get(st, st))
    return tuple(o[:n] for o in out)


def ricker(t, t0, f0):
    return lib().oracle_ricker(float(t), float(t0), float(f0))


def round_ties_up(x):
    return lib().oracle_round_ties_up(float(x))


def update_psi_2d(pcur, psi, coefs, halo, _dx, _dy):
    """update_psi_x!/_y! on caller arrays (Fortran-ordered float64)."""
    nx, ny = pcur.shape
    arrs = [np.asfortranarray(pcur)] + list(psi)
    for a in psi:
        assert a.flags.f_contiguous and a.dtype == np.float64
    cf = [_f64(c) for c in coefs]
    lib().oracle_update_psi_2d(nx, ny, halo, _dx, _dy, *[_p(a) for a in arrs],
                               *[_p(c) for c in cf])


def update_psi_3d(pcur, psi, coefs, halo, _dx, _dy, _dz):
    nx, ny, nz = pcur.shape
    arrs = [np.asfortranarray(pcur)] + list(psi)
    for a in psi:
        assert a.flags.f_contiguous and a.dtype == np.float64
    cf = [_f64(c) for c in coefs]
    lib().oracle_update_psi_3d(nx, ny, nz, halo, _dx, _dy, _dz, *[_p(a) for a in arrs],
                               *[_p(c) for c in cf])


class OracleProblem:
    """solve2D/solve3D restated on the CPU (Ricker sources only).

    n: grid points per axis; l: model size per axis [m]; vel: array of shape n;
    srcpos/recpos: (count, ndim) positions [m]; t0s: activation times.
    """

    def __init__(self, n, l, lt, vel, srcpos, t0s, f0, recpos, halo=20,
                 rcoef=0.0001, ppw=10.0, freetop=True):
        self.ndim = len(n)
        self.n = tuple(int(v) for v in n)
        vel = np.asfortranarray(np.asarray(vel, dtype=np.float64))
        assert vel.shape == self.n
        srcpos = np.asarray(srcpos, dtype=np.float64).reshape(-1, self.ndim)
        recpos = np.asarray(recpos, dtype=np.float64).reshape(-1, self.ndim)
        self.nsrcs, self.nrecs = srcpos.shape[0], recpos.shape[0]
        n_arr = (ctypes.c_long * self.ndim)(*self.n)
        l_arr = _f64(l)
        src_f = np.asfortranarray(srcpos)
        rec_f = np.asfortranarray(recpos)
        t0s = _f64(np.broadcast_to(np.asarray(t0s, dtype=np.float64), (self.nsrcs,)))
        st = ctypes.c_int(0)
        self._h = lib().oracle_create(self.ndim, n_arr, _p(l_arr), float(lt),
                                      vel.ctypes.data_as(_dp), self.nsrcs,
                                      src_f.ctypes.data_as(_dp), _p(t0s), float(f0),
                                      self.nrecs, rec_f.ctypes.data_as(_dp), int(halo),
                                      float(rcoef), float(ppw), int(bool(freetop)),
                                      ctypes.byref(st))
        if not self._h:
            if st.value in (-2, -3, -4):
                raise OracleError(ERRORS[st.value])
            raise ValueError(ERRORS.get(st.value, st.value))
        self.halo = int(halo)
        self.nt = lib().oracle_nt(self._h)
        self.dt = lib().oracle_dt(self._h)

    def close(self):
        if getattr(self, "_h", None):
            lib().oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def reset(self):
        lib().oracle_reset(self._h)

    def set_fact(self, fact):
        f = np.asfortranarray(np.asarray(fact, dtype=np.float64))
        assert f.shape == self.n
        lib().oracle_set_fact(self._h, f.ctypes.data_as(_dp))

    def misfit(self, obs):
        """Reset, run all steps, return 1/2 sum (traces - obs)^2."""
        self.reset()
        self.run()
        r = self.traces() - obs
        return 0.5 * float((r * r).sum())

    def fwi_gradient(self, obs):
        """(misfit, dJ/dfact) of the L2 trace misfit by the discrete adjoint
        (dt and the CPML coefficients held fixed).  Leaves the forward run's
        traces and fields in place."""
        obs = np.asfortranarray(np.asarray(obs, dtype=np.float64))
        assert obs.shape == (self.nt, self.nrecs)
        g = np.zeros(self.n, order="F")
        J = ctypes.c_double()
        st = lib().oracle_fwi_gradient(self._h, obs.ctypes.data_as(_dp), g.ctypes.data_as(_dp),
                                       ctypes.byref(J))
        if st != 0:
            raise ValueError(ERRORS.get(st, st))
        return J.value, g

    def run(self, nsteps=None):
        if nsteps is None:
            nsteps = self.nt
        return lib().oracle_run(self._h, int(nsteps))

    @property
    def it_done(self):
        return lib().oracle_it_done(self._h)

    def traces(self):
        out = np.zeros((self.nt, self.nrecs), order="F")
        if out.size:
            lib().oracle_get_traces(self._h, out.ctypes.data_as(_dp))
        return out

    def field(self, which=1):
        out = np.zeros(self.n, order="F")
        lib().oracle_get_field(self._h, which, out.ctypes.data_as(_dp))
        return out

    def pend(self):
        """The field solve2D/solve3D returns (pcur after the last rotation)."""
        return self.field(1)

    def fact(self):
        out = np.zeros(self.n, order="F")
        lib().oracle_get_fact(self._h, out.ctypes.data_as(_dp))
        return out

    def coef(self, axis, which):
        n = self.halo + (1 if which >= 4 else 0)
        out = np.zeros(max(n, 1))
        lib().oracle_get_coef(self._h, axis, which, _p(out))
        return out[:n]

    def dt2srctf(self):
        out = np.zeros((self.nt, self.nsrcs), order="F")
        if out.size:
            lib().oracle_get_dt2srctf(self._h, out.ctypes.data_as(_dp))
        return out

    def positions(self):
        ps = np.zeros((self.nsrcs, self.ndim), dtype=np.int64, order="F")
        pr = np.zeros((self.nrecs, self.ndim), dtype=np.int64, order="F")
        lib().oracle_get_positions(self._h, ps.ctypes.data_as(_lp), pr.ctypes.data_as(_lp))
        return ps, pr

    def _slab_shape(self, axis, layers):
        shp = list(self.n)
        shp[axis] = layers
        return tuple(shp)

    def psi(self, axis, side):
        out = np.zeros(self._slab_shape(axis, self.halo + 1), order="F")
        lib().oracle_get_psi(self._h, axis, side, out.ctypes.data_as(_dp))
        return out

    def xi(self, axis, side):
        out = np.zeros(self._slab_shape(axis, max(self.halo, 0)), order="F")
        if out.size:
            lib().oracle_get_xi(self._h, axis, side, out.ctypes.data_as(_dp))
        return out
