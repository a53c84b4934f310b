"""Host-side setup of the product and the C ABI surface (no GPU needed).

* awcpml.host.setup (the product's FP64 setup in Python) produces exactly
  the oracle's dt, nt, fact, CPML coefficients, wavelets and snapped
  positions, and raises the reference's assertion errors.
* libawcpml.so loads and exports every function include/awcpml.h declares,
  and its Python binding declares a signature for each.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import awcpml
from awcpml import _lib
from awcpml.host import setup
from cases import load_fixtures, oracle_problem, plain2d, small2d, small3d

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "awcpml.h")


def _setup(case, **kw):
    srcs = awcpml.Sources(case["srcpos"], case["t0s"], [awcpml.rickersource1D] * len(case["t0s"]),
                          case["f0"])
    recs = awcpml.Receivers(case["recpos"])
    return setup(case["l"], case["lt"], case["vel"], srcs, recs, halo=case["halo"],
                 rcoef=case["rcoef"], ppw=case.get("ppw", 10.0), freetop=case["freetop"], **kw)


CASES = {name: case for name, (case, _) in load_fixtures().items()}
CASES.update(plain2d=plain2d(), small2d=small2d(), small3d=small3d(),
             small3d_open=small3d(freetop=False, nsrc=2))


@pytest.mark.parametrize("name", sorted(CASES))
def test_setup_matches_oracle(name):
    case = CASES[name]
    S = _setup(case)
    o = oracle_problem(case)
    assert S.dt == o.dt and S.nt == o.nt
    assert np.array_equal(S.fact, o.fact())
    assert np.array_equal(S.dt2srctf, o.dt2srctf())
    for a in range(S.ndim):
        for w in range(8):
            assert np.array_equal(S.coef[a][w], o.coef(a, w)), (a, w)
    ps, pr = o.positions()
    assert np.array_equal(S.possrcs, ps) and np.array_equal(S.posrecs, pr)


def test_setup_assertions():
    case = small2d()
    with pytest.raises(AssertionError, match="points per wavelength"):
        _setup(dict(case, ppw=1000.0))
    with pytest.raises(AssertionError, match="source is not inside"):
        _setup(dict(case, srcpos=np.array([[1e6, 10.0]]), t0s=[0.01]))
    with pytest.raises(AssertionError, match="receiver is not inside"):
        _setup(dict(case, recpos=np.array([[10.0, -100.0]])))
    with pytest.raises(RuntimeError, match="Wrong onwhere parameter"):
        awcpml.calc_Kab_CPML(5, 1e-3, 2.0, 1.0, 1.0, 1.0, "elsewhere")
    with pytest.raises(AssertionError, match="at least one source"):
        awcpml.Sources(np.zeros((0, 2)), [], [], 1.0)
    with pytest.raises(AssertionError, match="do not match"):
        awcpml.Sources(np.zeros((2, 2)), [0.1], [awcpml.rickersource1D] * 2, 1.0)
    with pytest.raises(AssertionError, match="at least one receiver"):
        awcpml.Receivers(np.zeros((0, 2)))


def test_ricker_vectorised_matches_scalar():
    t = np.linspace(0.0, 1.0, 257)
    v = awcpml.rickersource1D(t, 0.4, 10.0)
    for k in (0, 17, 100, 256):
        assert v[k] == awcpml.rickersource1D(float(t[k]), 0.4, 10.0)


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(awcpml_\w+)\s*\(", text)))


def test_header_symbols_exported():
    names = _declared_functions()
    assert len(names) >= 18
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    # the binding declares a signature for every exported entry point
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_version_and_error_without_gpu_work():
    L = _lib.load()
    assert b"gfx950" in L.awcpml_version()
    # argument errors are reported before any device call
    h = ctypes.c_void_p()
    rc = L.awcpml_plan_create(4, 10, 10, 10, 2, 0, 1, ctypes.byref(h))
    assert rc == _lib.ERR_ARG and b"ndim" in L.awcpml_last_error()
    # overlapping slabs are accepted (left side wins), but the right psi
    # layers must start inside the grid: n >= halo + 2
    rc = L.awcpml_plan_create(3, 10, 10, 10, 9, 0, 1, ctypes.byref(h))
    assert rc == _lib.ERR_ARG and b"needs >= 11" in L.awcpml_last_error()
    rc = L.awcpml_plan_create(2, 10, 10, 1, 1, 7, 1, ctypes.byref(h))
    assert rc == _lib.ERR_ARG and b"dtype" in L.awcpml_last_error()
    assert L.awcpml_reset(None) == _lib.ERR_ARG


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libawcpml.so"))
    with pytest.raises(RuntimeError, match="not found"):
        _lib.load()


def test_fwi_backward_after_a_later_forward_raises():
    """fwi.Problem keeps one set of checkpoints: a backward whose forward was
    followed by another forward on the same problem raises instead of
    returning a gradient that mixes two models (host logic; the plan is a
    stand-in that records calls)."""
    import types

    import torch

    from awcpml import fwi

    class StandInPlan:
        nshots = 1

        def __init__(self):
            self.models = []

        def set_model(self, fact, inv_d, inv_d2):
            self.models.append(fact.copy())

        def checkpoint_run(self, every):
            pass

        def traces(self, shot):
            return np.full((3, 2), float(len(self.models)))

        def adjoint(self, trace_bar):
            return np.full((4, 4), float(len(self.models)))

    pr = object.__new__(fwi.Problem)
    pr.S = types.SimpleNamespace(inv_d=np.ones(2), inv_d2=np.ones(2))
    pr.dt, pr.every, pr.forward_count, pr.plan = 1e-3, 0, 0, StandInPlan()
    _, run1 = pr.forward(np.ones((4, 4)))
    _, run2 = pr.forward(2 * np.ones((4, 4)))
    with pytest.raises(RuntimeError, match="stale backward"):
        pr.pullback(run1, [np.zeros((3, 2))])
    assert np.all(pr.pullback(run2, [np.zeros((3, 2))]) == 2.0)
    v = torch.ones((4, 4), dtype=torch.float64, requires_grad=True)
    t1 = fwi.solve_traces(v, pr)
    t2 = fwi.solve_traces(v, pr)
    with pytest.raises(RuntimeError, match="stale backward"):
        t1.sum().backward()
    t2.sum().backward()
    assert v.grad is not None and torch.all(v.grad != 0)


def test_snapshot_file_layout(tmp_path):
    """do_save writes the reference's seven datasets (acoustic_3D_xPU.jl:
    480-498) with their names, element types, shapes and index order."""
    import types

    from awcpml.solver import _save_snapshot

    pcur = np.asfortranarray(np.arange(4 * 3 * 2, dtype=np.float64).reshape((4, 3, 2), order="F"))
    S = types.SimpleNamespace(halo=5, possrcs=np.array([[2, 3, 1]]),
                              posrecs=np.array([[1, 1, 1], [4, 3, 2]]))
    (tmp_path / "snap_it10.npz").write_bytes(b"stale")  # replaced, as the reference rm's it
    path = _save_snapshot(str(tmp_path), "snap", 10, (30.0, 20.0, 10.0), S, pcur)
    assert path.endswith("snap_it10.npz")
    with np.load(path, allow_pickle=False) as z:
        assert sorted(z.files) == sorted(["lx", "ly", "lz", "halo", "pcur", "possrcs", "posrecs"])
        assert z["lx"].dtype == np.float64 and z["lx"].shape == () and float(z["lz"]) == 10.0
        assert z["halo"].dtype == np.int64 and int(z["halo"]) == 5
        assert z["pcur"].shape == (4, 3, 2) and z["pcur"][3, 2, 1] == pcur[3, 2, 1]
        assert z["possrcs"].dtype == np.int64 and z["posrecs"].shape == (2, 3)
