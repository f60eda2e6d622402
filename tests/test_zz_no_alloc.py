"""mfd_solve_dev_noalloc (the solve entry point that allocates nothing: its
workspace is the library's fixed per-GPU scratch array) against
mfd_solve_dev, through the C ABI on device buffers: identical fields and
statistics on two calls in a row (the second smaller, so it finds the first
call's contents in the array), the residual on the oracle's A, and the size
limit."""
import sys

import numpy as np
import pytest

from conftest import spec_of

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def problem(pkg, N, npml, nrhs, seed):
    from maxwellfdfd import workloads as W
    m, om = W.random_boxes(N=N, npml=npml, nbox=4, lam=7.0, seed=seed, ndip=0)
    pkg.calc_matparams(m)
    B = W.dipole_rhs(m, om, W.random_dipoles(N, npml, nrhs, seed + 1))
    return m, om, B


def run(devmem, op, B, noalloc, **kw):
    b = devmem.put(B)
    e = devmem.put(np.zeros_like(B))
    nrhs = B.shape[0] if B.ndim == 2 else 1
    f = op.solve_dev_noalloc if noalloc else op.solve_dev
    if not noalloc:
        kw = dict(kw, use_graph=False)
    st = f(b, e, nrhs, **kw)
    return devmem.get(e, B), st


@pytest.mark.parametrize("method,precision", [("sqmr", "fp64"), ("sqmr", "mixed"), ("cocg", "fp64"),
                                              ("bicgstab", "fp64")])
def test_noalloc_matches_solve_dev(method, precision, pkg, O, devmem):
    cases = [(16, 3, 2, 11), (12, 2, 1, 21)]   # the second call is smaller (grid and right-hand sides)
    for N, npml, nrhs, seed in cases:
        m, om, B = problem(pkg, N, npml, nrhs, seed)
        A = O.to_scipy(O.create_A(spec_of(m), om, m.εarr, m.μarr))
        op = pkg.MaxwellOperator(m, om)
        say("noalloc", method, precision, N, nrhs)
        # BiCGSTAB does not converge on these random-box problems: it is
        # capped, and the two entry points must then agree on the best iterate
        kw = dict(maxit=3000, raise_on_noconv=False) if method == "bicgstab" else {}
        E0, st0 = run(devmem, op, B, False, tol=1e-8, precision=precision, method=method, **kw)
        E1, st1 = run(devmem, op, B, True, tol=1e-8, precision=precision, method=method, **kw)
        op.close()
        say("  iterations", st0.iterations, st1.iterations, "relres", st0.relres, st1.relres)
        # the same launches in the same order: bitwise the same result
        assert np.array_equal(E0, E1)
        assert st1.converged == st0.converged and st1.iterations == st0.iterations and st1.relres == st0.relres
        if not st1.converged:
            continue
        for r in range(nrhs):
            rel = np.linalg.norm(B[r] - A @ E1[r]) / np.linalg.norm(B[r])
            assert rel <= 1.05e-8


def test_noalloc_rejects_oversize(pkg, devmem):
    """A workspace above the 64 MiB array is refused before any GPU work."""
    m, om, B = problem(pkg, 48, 4, 2, 31)   # 2 × 331,776 unknowns: about 100 MB of workspace
    op = pkg.MaxwellOperator(m, om)
    b = devmem.put(B)
    e = devmem.put(np.zeros_like(B))
    with pytest.raises(pkg.MfdArgumentError):
        op.solve_dev_noalloc(b, e, 2)
    op.close()
    assert not np.any(devmem.get(e, B))
