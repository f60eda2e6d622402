"""The reference's matrix-argument API on the device: create_A / create_b /
create_linsys / h_from_e / e_from_h from the assembled (Pε, Pμ, Cₑ, Cₘ)
(model.jl:209-284), and an operator built from the assembled A.  The
matrices come from the oracle, so each result is checked against the
oracle's own A and b (CSC structure bit-exact, values to 1e-13 of the
largest entry), against the closed-form device assembly, and against scipy
products of the same matrices."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from conftest import spec_of
from test_gpu_parity import CASES, HHCASES, make, rand_params

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def julia_csc(A):
    """(colptr, rowval, nzval), 1-based, as a SparseMatrixCSC holds them."""
    A = sp.csc_matrix(A)
    A.sort_indices()
    return A.indptr + 1, A.indices + 1, A.data


def assert_same_csc(got, ref, what, rtol=1e-13):
    gc, gr, gv = julia_csc(got)
    rc, rr, rv = julia_csc(ref)
    assert np.array_equal(gc, rc), f"{what}: colptr differs"
    assert np.array_equal(gr, rr), f"{what}: rowval differs"
    scale = max(np.abs(rv).max() if len(rv) else 0.0, 1e-300)
    assert (np.abs(gv - rv).max() if len(rv) else 0.0) <= rtol * scale, what


def factors(O, m, ε, μ):
    s = spec_of(m)
    Pe = O.to_scipy(O.paramop(s, 0, ε))
    Pm = O.to_scipy(O.paramop(s, 1, μ))
    Ce = O.to_scipy(O.curl(s, 0))
    Cm = O.to_scipy(O.curl(s, 1))
    return Pe, Pm, Ce, Cm


@pytest.mark.parametrize("case", list(CASES) + ["hh_pml"])
@pytest.mark.parametrize("ft", [0, 1])
@pytest.mark.parametrize("full", [False, True])
def test_create_A_from_matrices(case, ft, full, pkg, O):
    m = make(pkg, case)
    ε, μ = rand_params(m, 31, full=full and ft == 0, mu=True)
    if ft == 1 and full:
        pytest.skip("HH divides by Pε, which must be diagonal")
    om = 0.6 + 0.02j
    Pe, Pm, Ce, Cm = factors(O, m, ε, μ)
    say("create_A from matrices", case, ft, full)
    A = pkg.create_A(ft, om, Pe, Pm, Ce, Cm)
    ref = O.to_scipy(O.create_A(spec_of(m), om, ε, μ, ft=ft))
    assert_same_csc(A, ref, "A vs oracle")
    assert_same_csc(pkg.create_A(ft, om, m, εarr=ε, μarr=μ), A, "A vs closed-form device assembly")
    assert_same_csc(pkg.create_A(ft, om, (Pe, Pm), (Ce, Cm)), A, "tuple form")
    # ω = 0: no subtraction, the product's structural entries are all kept
    A0 = pkg.create_A(ft, 0.0, Pe, Pm, Ce, Cm)
    assert_same_csc(A0, O.to_scipy(O.create_A(spec_of(m), 0.0, ε, μ, ft=ft)), "A(ω=0) vs oracle")


@pytest.mark.parametrize("case", ["uniform_pml", "mixed_pml_bloch", "cmplast", "boundft_HH", "hh_pml"])
@pytest.mark.parametrize("ft", [0, 1])
def test_create_b_linsys_fields_from_matrices(case, ft, pkg, O):
    m = make(pkg, case)
    ε, μ = rand_params(m, 32, mu=True)
    om = 0.7 - 0.01j
    Pe, Pm, Ce, Cm = factors(O, m, ε, μ)
    n = Pe.shape[0]
    rng = np.random.default_rng(5)
    je = rng.normal(size=n) + 1j * rng.normal(size=n)
    jm = rng.normal(size=n) + 1j * rng.normal(size=n)
    say("create_b from matrices", case, ft)
    b = pkg.create_b(ft, om, Pe, Pm, Ce, Cm, je, jm)
    ref = O.create_b(spec_of(m), om, μ, je, jm, ft=ft, eps_arr=ε)
    assert np.abs(b - ref).max() <= 1e-14 * np.abs(ref).max()
    assert np.array_equal(pkg.create_b(ft, om, (Pe, Pm), (Ce, Cm), (je, jm)), b)
    A, b2 = pkg.create_linsys(ft, om, Pe, Pm, Ce, Cm, je, jm)
    assert np.array_equal(b2, b)
    assert_same_csc(A, O.to_scipy(O.create_A(spec_of(m), om, ε, μ, ft=ft)), "create_linsys A")
    # h_from_e / e_from_h against scipy products of the same matrices
    e = rng.normal(size=n) + 1j * rng.normal(size=n)
    h = pkg.h_from_e(e, om, Pm, Ce, jm)
    href = (1j / om) * ((Ce @ e + jm) / Pm.diagonal())
    assert np.abs(h - href).max() <= 1e-13 * np.abs(href).max()
    e2 = pkg.e_from_h(h, om, Pe, Cm, je)
    eref = (-1j / om) * ((Cm @ h - je) / Pe.diagonal())
    assert np.abs(e2 - eref).max() <= 1e-13 * np.abs(eref).max()
    hz = pkg.h_from_e(e, om, (Pe, Pm), (Ce, Cm), (je, np.zeros(n, complex)))
    assert np.abs(hz - (1j / om) * ((Ce @ e) / Pm.diagonal())).max() <= 1e-13 * np.abs(hz).max()


def test_matrix_api_errors(pkg, O):
    m = make(pkg, "uniform_pml")
    ε, μ = rand_params(m, 33, full=True, mu=True)
    Pe, Pm, Ce, Cm = factors(O, m, ε, μ)
    # HH divides by Pε, here not diagonal (as model.jl:239 notes, unsupported)
    with pytest.raises(pkg.MfdUnsupportedError):
        pkg.create_A(1, 0.5, Pe, Pm, Ce, Cm)
    with pytest.raises(pkg.MfdArgumentError):   # sizes disagree
        pkg.create_A(0, 0.5, Pe, Pm, Ce[:-3, :-3], Cm)
    with pytest.raises(pkg.MfdArgumentError):   # singular divisor
        Z = Pm.tolil()
        Z[4, 4] = 0
        pkg.create_A(0, 0.5, Pe, Z.tocsc(), Ce, Cm)


@pytest.mark.parametrize("case,sym", [("uniform_pml", True), ("cmplast", True), ("mixed_pml_bloch", False),
                                      ("nonuniform_bloch", False)])
def test_operator_from_matrix(case, sym, pkg, O):
    """An operator built from the assembled A applies A exactly and finds the
    diagonal S with S A symmetric when there is one (PML grids: S = the
    E-cell volumes, not known to it otherwise), so SQMR and the adjoint serve
    it; a complex Bloch phase leaves A without one (BiCGSTAB, SQMR refused)."""
    m = make(pkg, case)
    ε, μ = rand_params(m, 12)
    om = 0.55 + 0.02j
    A = O.to_scipy(O.create_A(spec_of(m), om, ε, μ))
    say("operator from A", case)
    op = pkg.MaxwellOperator.from_csc(A)
    rng = np.random.default_rng(6)
    n = A.shape[0]
    X = rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n))
    Y = op.apply(X)
    for r in range(2):
        assert np.linalg.norm(Y[r] - A @ X[r]) <= 1e-13 * np.linalg.norm(A @ X[r])
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    xd = sla.spsolve(A.tocsc(), b)
    for prec in ("fp64", "mixed"):
        e, st = op.solve(b, tol=1e-10, precision=prec)
        rel = np.linalg.norm(b - A @ e) / np.linalg.norm(b)
        say("  ", prec, "iterations", st.iterations, "residual", rel)
        assert st.converged and rel <= 1.05e-10
        assert np.linalg.norm(e - xd) / np.linalg.norm(xd) <= 1e-7
    if sym:
        e2, _ = op.solve(b, tol=1e-10, method="sqmr")
        assert np.linalg.norm(b - A @ e2) / np.linalg.norm(b) <= 1.05e-10
        g = rng.normal(size=n) + 1j * rng.normal(size=n)
        lam, _ = op.solve_adjoint(g, tol=1e-11)
        ld = sla.spsolve(A.T.tocsc(), g)
        assert np.linalg.norm(lam - ld) / np.linalg.norm(ld) <= 1e-8
    else:
        with pytest.raises(pkg.MfdUnsupportedError):
            op.solve(b, method="sqmr")
    e3, st3 = pkg.solve_linsys(A, b, tol=1e-10)
    op.close()
    assert np.linalg.norm(b - A @ e3) / np.linalg.norm(b) <= 1.05e-10


def test_operator_from_matrix_matches_matrix_free(pkg, O):
    """On the smoothed-ball A of a periodic grid the sparse operator and the
    matrix-free operator give the same solution, in both precisions."""
    from maxwellfdfd import workloads as W
    m = pkg.ModelFull(pkg.Grid([np.arange(11.0)] * 3, (True, True, True)))
    pkg.set_ωpml(m, 1.0)
    pkg.add_obj(m, "bg", pkg.Cuboid((5, 5, 5), (10, 10, 10)), ε=1.0)
    pkg.add_obj(m, "ball", pkg.Ball((5, 5, 5), 2.5), ε=(4.0 + 0.3j))
    pkg.calc_matparams(m)
    om = 0.9
    A = pkg.create_A(pkg.EE, om, m)
    pkg.add_srce(m, pkg.PointSrc((5, 5, 5), (0, 0, 1)))
    je, jm = pkg.create_srcs(m)
    b = pkg.create_b(pkg.EE, om, m, je, jm)
    mf = pkg.MaxwellOperator(m, om)
    sp_op = pkg.MaxwellOperator.from_csc(A)
    for precision in ("fp64", "mixed"):
        say("sparse vs matrix-free", precision)
        e1, s1 = mf.solve(b, tol=1e-9, precision=precision)
        e2, s2 = sp_op.solve(b, tol=1e-9, precision=precision)
        assert s1.converged and s2.converged
        assert np.linalg.norm(e1 - e2) / np.linalg.norm(e1) <= 1e-7
    mf.close()
    sp_op.close()
