"""The oracle's CPU FP64 solver (same scaling and Krylov method as the device
solver) against a direct sparse LU solve (scipy SuperLU, standing in for the
reference workflow's `A \\ b`)."""
import numpy as np
import scipy.sparse.linalg as sla

from conftest import shapes_of, spec_of


def test_cpu_sqmr_matches_direct(O, pkg):
    from maxwellfdfd import workloads as W
    m, om = W.sphere_dipole(N=12, npml=3, r=2.5, lam=8.0)
    s = spec_of(m)
    shapes, et, mt = shapes_of(m)
    ea, ma = O.matparams(s, shapes, et, mt)
    A = O.create_A(s, om, ea, ma)
    j = np.zeros((12, 12, 12, 3), dtype=np.complex128, order="F")
    O.add_pointsrc(s, 0, (6, 6, 6), (0, 0, 1), 1.0, j)
    je = O.field_arr2vec(s, j)
    b = O.create_b(s, om, ma, je, np.zeros_like(je))
    x, it, rr = O.solve_sqmr(s, A, b, tol=1e-10, maxit=20000, nthreads=4)
    assert rr <= 1e-10
    xd = sla.spsolve(O.to_scipy(A).tocsc(), b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8
