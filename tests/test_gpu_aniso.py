"""The large anisotropic workload (seeded smoothed spheres: full 3×3 ε with
off-diagonals at every sphere surface, several dipoles batched) at a size the
oracle and a direct solve handle: device calc_matparams against the oracle
(off-diagonals included), the anisotropic tiled stencil against the oracle's
CSC A, and the batched FP64 and mixed-precision solves against the oracle's
residual and a direct solve."""
import sys

import numpy as np
import pytest
import scipy.sparse.linalg as sla

from conftest import shapes_of, spec_of

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def spheres(pkg, O):
    """24×24×20 grid, 4-cell PML, 6 spheres of radius 3–6 and ε in [2, 12],
    2 dipoles, λ = 16 cells (the 384×384×256 generator at reduced size; at
    λ = 10, under 3 cells per wavelength inside ε = 12, Jacobi-SQMR stalls,
    on the oracle's CPU solver too)."""
    from maxwellfdfd import workloads as W
    m, om, dips = W.random_spheres(N=(24, 24, 20), npml=4, nsph=6, rmin=3.0, rmax=6.0, lam=16.0, seed=5, ndip=2)
    say("spheres: matparams")
    pkg.calc_matparams(m)
    B = W.dipole_rhs(m, om, dips)
    s = spec_of(m)
    shapes, et, mt = shapes_of(m)
    ea, ma = O.matparams(s, shapes, et, mt)
    A = O.to_scipy(O.create_A(s, om, ea, ma))
    say("spheres: direct solves")
    X = sla.splu(A.tocsc()).solve(B.T.copy()).T
    return m, om, B, ea, ma, A, X


def test_spheres_matparams(spheres):
    m, om, B, ea, ma, A, X = spheres
    scale = np.abs(ea).max()
    for v in range(3):
        for w in range(3):
            err = np.abs(m.εarr[..., v, w] - ea[..., v, w]).max()
            assert err <= 1e-12 * scale, (v, w, err)
    assert np.abs(m.μarr - ma).max() <= 1e-12
    # the workload's point: off-diagonals at the sphere surfaces
    off = np.abs(ea[..., 0, 1]) + np.abs(ea[..., 0, 2]) + np.abs(ea[..., 1, 2])
    assert (off > 1e-3).sum() > 200 and off.max() > 0.5


@pytest.mark.parametrize("tile", [None, (8, 32, 1), (2, 4, 0), (0, 1, 0)])
def test_spheres_apply(tile, spheres, pkg):
    """The anisotropic tiled stencil (default, the autotuner's large-grid
    choice 8×32, a small tile in dispatch order) and the direct kernel give
    the oracle's A x."""
    m, om, B, ea, ma, A, X = spheres
    op = pkg.MaxwellOperator(m, om)
    if tile is not None:
        op.set_tile("fp64", *tile)
    rng = np.random.default_rng(2)
    V = rng.normal(size=(2, A.shape[0])) + 1j * rng.normal(size=(2, A.shape[0]))
    say("spheres apply", tile)
    Y = op.apply(V)
    op.close()
    for r in range(2):
        ref = A @ V[r]
        assert np.linalg.norm(Y[r] - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_spheres_batched_solve(precision, spheres, pkg):
    """Both dipoles in one batched solve to 1e-8: each column meets the
    tolerance on the oracle's A, and its field agrees with a direct solve."""
    m, om, B, ea, ma, A, X = spheres
    op = pkg.MaxwellOperator(m, om)
    say("spheres batched solve", precision)
    E, st = op.solve(B, tol=1e-8, precision=precision)
    op.close()
    say("  iterations", st.iterations, "fp32", st.iterations_fp32, "relres", st.relres)
    assert st.converged
    for r in range(B.shape[0]):
        rel = np.linalg.norm(B[r] - A @ E[r]) / np.linalg.norm(B[r])
        xd = X[r]
        err = np.linalg.norm(E[r] - xd) / np.linalg.norm(xd)
        say("  rhs", r, "oracle residual", rel, "field error vs direct", err)
        assert rel <= 1.05e-8
        assert err <= 1e-6
