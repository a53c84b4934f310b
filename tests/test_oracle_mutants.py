"""Planted mistakes: every plausible slip in the oracle must fail a check.

Each mutant is a copy of oracle/awcpml_oracle.c with one edit (a flipped
sign, a shifted index, a dropped term, a wrong scale or coefficient array),
compiled on the CPU into a temporary library.  The pinning checks -- the
reference's golden fields, the psi known-answer tests, the independent numpy
restatement and the analytic Green's function -- are run against it; at
least one must fail.  The unmodified oracle passes all of them (the other
test files), so a surviving mutant would mean the checks leave that part of
the oracle unpinned.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
import numpy_step
from cases import load_fixtures, oracle_problem, rel_l2, small2d, small3d

ORACLE_SRC = os.path.join(os.path.dirname(oracle.__file__), "awcpml_oracle.c")
SQRT_EPS = 1.4901161193847656e-08

# (name, old text, new text); old must occur in the source
MUTANTS = [
    ("psi_x_sign",
     "C1(a_x_hl, i) * (A2(pcur, nx, i + 1, j) - A2(pcur, nx, i, j)) * _dx",
     "C1(a_x_hl, i) * (A2(pcur, nx, i, j) - A2(pcur, nx, i + 1, j)) * _dx"),
    ("xi_right_coef_array",
     "A2(xi_x_r, halo, ii - 1, j) = C1(cx[3], ii - 1)",
     "A2(xi_x_r, halo, ii - 1, j) = C1(cx[7], ii - 1)"),
    ("drop_y_bottom_damping",
     "damp += A2(fact, nx, i, j) * (dpsi + A2(xi_y_r, nx, i, jj - 1));",
     "damp += 0.0;"),
    ("no_halfgrid_override",
     "if (halfgrd) dist[0] = 0.0; /* utils.jl:36-38 */",
     "/* override removed */"),
    ("source_cell_scale",
     "double cell = dx * dy;\n",
     "double cell = dx;\n"),
    ("free_top_wrong_side",
     "for (int i = 0; i < halo; ++i) { o->coef[1][0][i] = 0.0; o->coef[1][2][i] = 1.0; }",
     "for (int i = 0; i < halo; ++i) { o->coef[1][1][i] = 0.0; o->coef[1][3][i] = 1.0; }"),
    ("z_right_psi_index_3d",
     "(A3(psi_z_r, nx, ny, i, j, kk) - A3(psi_z_r, nx, ny, i, j, kk - 1)) * _dz",
     "(A3(psi_z_r, nx, ny, i, j, kk) - A3(psi_z_r, nx, ny, i, j, kk)) * _dz"),
    ("z_left_d2_in_xi_3d",
     "C1(cz[0], k - 1) * (d2p_dz2 + dpsi)",
     "C1(cz[0], k - 1) * (d2p_dy2 + dpsi)"),
    ("laplacian_drops_z_3d",
     "f * (d2p_dx2 + d2p_dy2 + d2p_dz2) + damp",
     "f * (d2p_dx2 + d2p_dy2) + damp"),
    ("receiver_before_injection",
     "        pnew[idx] += o->dt2srctf[(size_t)s * o->nt + (it - 1)];",
     "        pnew[idx] += 0.0 * o->dt2srctf[(size_t)s * o->nt + (it - 1)];"),
]


def _build(tmp_path, name, old, new):
    src = open(ORACLE_SRC).read()
    assert src.count(old) >= 1, name
    mutated = src.replace(old, new, 1)
    c = tmp_path / f"{name}.c"
    so = tmp_path / f"lib{name}.so"
    c.write_text(mutated)
    subprocess.run(["gcc", "-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-std=c99", "-shared",
                    "-o", str(so), str(c), "-lm"], check=True)
    return str(so)


def _golden_ok():
    for name, (case, ref) in load_fixtures().items():
        o = oracle_problem(case)
        o.run()
        if not rel_l2(o.pend(), ref) <= SQRT_EPS:
            return False
    return True


def _numpy_ok():
    for case in (small2d(freetop=True), small2d(freetop=False, n=(31, 29), halo=4),
                 small3d(freetop=True), small3d(freetop=False, nsrc=2)):
        o = oracle_problem(case)
        sp = [case["l"][a] / (case["n"][a] - 1) for a in range(len(case["n"]))]
        st = numpy_step.from_oracle(o, sp)
        for it in range(1, o.nt + 1):
            numpy_step.step(st, it)
        o.run()
        if not (np.array_equal(st["pcur"], o.pend()) and np.array_equal(st["traces"], o.traces())):
            return False
    return True


def _symmetry_ok():
    import test_oracle_invariants as ti
    try:
        ti.test_swap_xy_3d_exact()
        ti.test_cyclic_axes_3d()
        ti.test_transpose_2d_exact()
    except AssertionError:
        return False
    return True


def _analytic3d_ok():
    import test_oracle_analytic as ta
    try:
        ta.test_green_3d(20, 50.0)
    except AssertionError:
        return False
    return True


@pytest.mark.parametrize("name,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_mutant_is_caught(tmp_path, monkeypatch, name, old, new):
    so = _build(tmp_path, name, old, new)
    monkeypatch.setattr(oracle, "LIB_PATH", so)
    monkeypatch.setattr(oracle, "_lib", None)
    # the restatement-independent checks first: golden data, symmetry,
    # closed-form physics; the numpy restatement last
    caught = []
    for check in (_golden_ok, _symmetry_ok, _analytic3d_ok, _numpy_ok):
        if not check():
            caught.append(check.__name__)
            break
    monkeypatch.setattr(oracle, "_lib", None)
    assert caught, f"mutant {name} survived every pinning check"


@pytest.mark.parametrize("name,old,new", [m for m in MUTANTS if m[0].endswith("_3d")],
                         ids=[m[0] for m in MUTANTS if m[0].endswith("_3d")])
def test_3d_mutant_caught_without_numpy_restatement(tmp_path, monkeypatch, name, old, new):
    """3D has no golden field; its mutants must fail an invariant or the
    closed-form physics, not only the second restatement."""
    so = _build(tmp_path, name, old, new)
    monkeypatch.setattr(oracle, "LIB_PATH", so)
    monkeypatch.setattr(oracle, "_lib", None)
    ok = _symmetry_ok() and _analytic3d_ok()
    monkeypatch.setattr(oracle, "_lib", None)
    assert not ok, f"3D mutant {name} is caught only by the numpy restatement"
