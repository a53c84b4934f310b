"""awcpml_run_noalloc, the allocation-free one-shot solve, against
awcpml_run on a one-shot plan (the existing entry point it varies).

Inputs and outputs are device buffers (hipMalloc through the HIP runtime
the library uses); the wavefields live in the library's static scratch
array.  Two calls in a row,
the second on a smaller grid, each bit for bit the plan's result (FP64 and
FP32) and, in FP64, the CPU oracle's; a problem larger than the scratch
array is refused with a clear error.
"""
import ctypes

import numpy as np
import pytest

import awcpml
from awcpml import _lib
from cases import oracle_problem, small2d, small3d

pytestmark = pytest.mark.gpu


def _oracle(case):
    o = oracle_problem(case)
    o._spacing = [case["l"][a] / (case["n"][a] - 1) for a in range(len(case["n"]))]
    return o


def _inputs(o):
    sp = o._spacing
    return dict(n=o.n, halo=o.halo, fact=o.fact(), inv_d=[1 / d for d in sp],
                inv_d2=[1 / (d * d) for d in sp],
                coef=[[o.coef(a, w) for w in range(8)] for a in range(o.ndim)],
                pos=o.positions(), wav=o.dt2srctf(), nt=o.nt)


def _plan_run(x, dtype):
    """awcpml_run on a one-shot plan: traces and final field."""
    P = awcpml.Plan(x["n"], x["halo"], dtype=dtype)
    P.set_model(x["fact"], x["inv_d"], x["inv_d2"])
    for a, c in enumerate(x["coef"]):
        P.set_cpml(a, c)
    ps, pr = x["pos"]
    P.set_sources(0, ps, x["wav"])
    P.set_receivers(0, pr)
    tr = np.zeros((x["nt"], pr.shape[0]), order="F")
    pend = np.zeros(x["n"], order="F")
    _lib.check(_lib.load().awcpml_run(P._h, tr.ctypes.data_as(_lib._dp),
                                      pend.ctypes.data_as(_lib._dp)))
    P.close()
    return tr, pend


class _Dev:
    """Device buffers through the HIP runtime libawcpml itself uses (the
    library is loaded first, so this opens the same runtime)."""

    def __init__(self):
        _lib.load()
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipDeviceSynchronize.argtypes = []
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 256)) == 0
        self.ptrs.append(p)
        if a.nbytes:
            assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, shape, dtype):
        out = np.empty(int(np.prod(shape)), dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        assert self.hip.hipDeviceSynchronize() == 0
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _noalloc_run(x, dtype):
    """The same solve through awcpml_run_noalloc with device buffers."""
    ndt = np.float64 if dtype == "f64" else np.float32
    nd = len(x["n"])
    nx, ny = x["n"][0], x["n"][1]
    nz = x["n"][2] if nd == 3 else 1
    D = _Dev()
    # Fortran order (x fastest) flattened: the library's layout
    fact = D.put(np.ravel(x["fact"], order="F").astype(ndt))
    coef = [D.put(np.asarray(v, dtype=np.float64).astype(ndt)) for c in x["coef"] for v in c]
    wav = D.put(np.ravel(x["wav"], order="F").astype(ndt))
    ps, pr = x["pos"]
    nt, nrec = x["nt"], pr.shape[0]
    traces = D.put(np.full(nt * nrec, np.nan, dtype=ndt))
    pend = D.put(np.full(nx * ny * nz, np.nan, dtype=ndt))
    cptr = (ctypes.c_void_p * len(coef))(*[p.value for p in coef])
    ps = np.asfortranarray(ps, dtype=np.int64)
    pr = np.asfortranarray(pr, dtype=np.int64)
    inv_d = np.ascontiguousarray(x["inv_d"], dtype=np.float64)
    inv_d2 = np.ascontiguousarray(x["inv_d2"], dtype=np.float64)
    rc = _lib.load().awcpml_run_noalloc(
        nd, nx, ny, nz, x["halo"], _lib.F64 if dtype == "f64" else _lib.F32, fact,
        _lib.dptr(inv_d), _lib.dptr(inv_d2), cptr, None, ps.shape[0],
        ps.ctypes.data_as(_lib._i64p), nt, wav, pr.shape[0], pr.ctypes.data_as(_lib._i64p),
        traces, pend, None)
    _lib.check(rc)
    tr = D.get(traces, (nt * nrec,), ndt).astype(np.float64).reshape((nt, nrec), order="F")
    p = D.get(pend, (nx * ny * nz,), ndt).astype(np.float64).reshape(x["n"], order="F")
    D.free()
    return tr, p


def test_noalloc_matches_run_on_two_calls():
    """A 3D call, then a smaller 2D one (the scratch array keeps the first
    call's contents, which the second must not see), then the 3D one again
    in FP32: each bitwise the plan path, FP64 also bitwise the oracle."""
    seq = [("3d", small3d(freetop=True, nsrc=2, nt=90), "f64"),
           ("2d", small2d(n=(41, 33), halo=5, nt=80, freetop=False), "f64"),
           ("3d", small3d(freetop=False, nsrc=1, nt=70), "f32")]
    for name, case, dtype in seq:
        print(f"[gpu] noalloc {name} n={case['n']} {dtype}", flush=True)
        o = _oracle(case)
        x = _inputs(o)
        ref_tr, ref_p = _plan_run(x, dtype)
        tr, p = _noalloc_run(x, dtype)
        assert np.array_equal(tr, ref_tr), (name, dtype)
        assert np.array_equal(p, ref_p), (name, dtype)
        if dtype == "f64":
            o.run()
            assert np.array_equal(tr, o.traces())
            assert np.array_equal(p, o.pend())
        assert np.abs(tr).max() > 0


def test_noalloc_refuses_oversize():
    L = _lib.load()
    rc = L.awcpml_run_noalloc(3, 400, 400, 400, 20, _lib.F64, ctypes.c_void_p(16), None, None,
                              None, None, 0, None, 10, None, 0, None, None, None, None)
    assert rc == _lib.ERR_ARG
    inv = np.ones(3)
    cptr = (ctypes.c_void_p * 24)(*([16] * 24))
    rc = L.awcpml_run_noalloc(3, 400, 400, 400, 20, _lib.F64, ctypes.c_void_p(16), _lib.dptr(inv),
                              _lib.dptr(inv), cptr, None, 0, None, 10, None, 0, None, None, None,
                              None)
    assert rc == _lib.ERR_ARG
    assert b"scratch" in L.awcpml_last_error()
