"""Solver edge cases through the C ABI: per-column best iterate on a solve that
ends without converging, zero right-hand sides with a nonzero initial guess,
and NaN right-hand sides."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import spec_of

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def periodic_case(pkg, O, seed=21):
    """A small periodic, PML-free problem on a uniform grid (so the solver's
    scaling is P = Q A Q with Q = diag(A)^-1/2) whose Jacobi-scaled BiCGSTAB
    residual rises above 3x its start at every one of its first 16 steps from
    the seeded white-noise right-hand side c (checked on the CPU with a numpy
    BiCGSTAB: 4.8x, 11.7x, ..., 47x at step 8)."""
    rng = np.random.default_rng(seed)
    N = tuple(int(v) for v in rng.integers(3, 9, 3))
    m = pkg.ModelFull(pkg.Grid([np.arange(n + 1.0) for n in N], (True, True, True)))
    pkg.set_ωpml(m, 0.7)
    pkg.set_Npml(m, ((0, 0, 0), (0, 0, 0)))
    eps = np.zeros(N + (3, 3), complex, order="F")
    mu = np.zeros_like(eps)
    for c in range(3):
        eps[..., c, c] = 1 + 11 * rng.random(N) + 0.02j * rng.random(N)
        mu[..., c, c] = 1
    om = float(rng.uniform(0.8, 1.6))
    A = O.to_scipy(O.create_A(spec_of(m), om, eps, mu)).tocsr()
    c = rng.normal(size=A.shape[0]) + 1j * rng.normal(size=A.shape[0])
    return m, om, eps, mu, A, c


def bicgstab_residuals(P, c, steps):
    """Relative residuals of `steps` BiCGSTAB steps on P x = c from x = 0."""
    x = np.zeros_like(c)
    r = c.copy()
    rh = r.copy()
    p = r.copy()
    rho = np.vdot(rh, r)
    out = []
    for _ in range(steps):
        v = P @ p
        a = rho / np.vdot(rh, v)
        s = r - a * v
        t = P @ s
        om = np.vdot(t, s) / np.vdot(t, t)
        x = x + a * p + om * s
        r = s - om * t
        rho1 = np.vdot(rh, r)
        p = r + (rho1 / rho) * (a / om) * (p - om * v)
        rho = rho1
        out.append(np.linalg.norm(c - P @ x) / np.linalg.norm(c))
    return np.array(out)


def test_noconv_restores_per_column(pkg, O):
    """Two right-hand sides in one BiCGSTAB solve capped at 8 steps: column 0
    is Q⁻¹u for an eigenvector u of P (it converges in the first step),
    column 1 is Q⁻¹c, whose residual rises far above its start.  Column 0 must
    keep its converged iterate; column 1 goes back to its best iterate (the
    initial guess 0, residual 1)."""
    m, om, eps, mu, A, c = periodic_case(pkg, O)
    q = 1 / np.sqrt(A.diagonal())
    P = sp.diags(q) @ A @ sp.diags(q)
    hist = bicgstab_residuals(P, c, 16)
    assert hist.min() > 3.0, hist   # the premise: every step of this cycle is worse than its start
    lam, V = np.linalg.eig(P.toarray())
    k = int(np.argmax(np.abs(lam)))
    u = V[:, k]
    B = np.stack([u / q, c / q])
    op = pkg.MaxwellOperator(m, om, εarr=eps, μarr=mu)
    say("noconv per-column restore")
    E, st = op.solve(B, tol=1e-8, maxit=8, check_every=8, method="bicgstab", raise_on_noconv=False)
    op.close()
    rel = [np.linalg.norm(B[r] - A @ E[r]) / np.linalg.norm(B[r]) for r in range(2)]
    say("  column residuals", rel, "relres", st.relres)
    assert not st.converged
    assert rel[0] <= 1e-8, "the converged column lost its iterate"
    assert np.array_equal(E[1], np.zeros_like(E[1])), "column 1 was not restored to its best iterate"
    assert abs(st.relres - 1.0) <= 1e-12


@pytest.mark.parametrize("method", ["sqmr", "bicgstab"])
def test_noconv_single_restore(method, pkg, O):
    """One right-hand side whose only cycle makes the residual worse: the
    solve returns the initial guess and reports its residual."""
    m, om, eps, mu, A, c = periodic_case(pkg, O)
    q = 1 / np.sqrt(A.diagonal())
    b = c / q
    x0 = np.linalg.solve(A.toarray(), b) * (1 + 1e-3)   # residual 1e-3 of ‖b‖, up to rounding
    r0 = np.linalg.norm(b - A @ x0) / np.linalg.norm(b)
    op = pkg.MaxwellOperator(m, om, εarr=eps, μarr=mu)
    say("noconv single restore", method)
    e, st = op.solve(b, tol=1e-30, x0=x0, maxit=8, check_every=8, method=method, raise_on_noconv=False)
    op.close()
    rel = np.linalg.norm(b - A @ e) / np.linalg.norm(b)
    say("  start", r0, "returned", rel, "relres", st.relres)
    assert not st.converged
    # never worse than the initial guess, and relres is the returned iterate's
    assert rel <= r0 * (1 + 1e-9)
    assert abs(rel - st.relres) <= 1e-6 * st.relres


@pytest.mark.parametrize("method", ["sqmr", "cocg", "bicgstab"])
def test_zero_rhs_with_initial_guess(method, pkg, O):
    """b = 0 in one column with a nonzero initial guess: that column returns
    the exact solution 0 (as A \\ b does) and the other column converges."""
    from maxwellfdfd import workloads as W
    m, om = W.sphere_dipole(N=12, npml=3, r=2.5, lam=8.0)
    pkg.calc_matparams(m)
    je, jm = pkg.create_srcs(m)
    b = pkg.create_b(pkg.EE, om, m, je, jm)
    A = O.to_scipy(O.create_A(spec_of(m), om, m.εarr, m.μarr))
    rng = np.random.default_rng(4)
    B = np.stack([np.zeros_like(b), b])
    X0 = np.zeros_like(B)
    X0[0] = rng.normal(size=b.shape) + 1j * rng.normal(size=b.shape)
    op = pkg.MaxwellOperator(m, om)
    say("zero rhs", method)
    E, st = op.solve(B, tol=1e-8, x0=X0, method=method)
    assert st.converged and st.relres <= 1e-8
    assert np.array_equal(E[0], np.zeros_like(E[0]))
    assert np.linalg.norm(b - A @ E[1]) / np.linalg.norm(b) <= 1.05e-8
    e1, st1 = op.solve(np.zeros_like(b), tol=1e-8, x0=X0[0], method=method)
    op.close()
    assert st1.converged and st1.iterations == 0 and np.array_equal(e1, np.zeros_like(e1))


def test_zero_rhs_in_uneven_batch(pkg, O):
    """A problem with b = 0 in a batched solve returns 0, converged."""
    from maxwellfdfd import workloads as W
    probs = W.uneven_problems(count=2, nmin=8, nmax=12, npml=2, nbox=3, seed=9)
    ops, bs = [], []
    for m, om in probs:
        pkg.calc_matparams(m)
        je, jm = pkg.create_srcs(m)
        bs.append(pkg.create_b(pkg.EE, om, m, je, jm))
        ops.append(pkg.MaxwellOperator(m, om))
    bs[0] = np.zeros_like(bs[0])
    say("zero rhs in batch")
    es, st = pkg.solve_batch(ops, bs, tol=1e-8)
    for op in ops:
        op.close()
    assert st[0].converged and np.array_equal(es[0], np.zeros_like(es[0]))
    assert st[1].converged and st[1].relres <= 1e-8


@pytest.mark.parametrize("method", ["sqmr", "cocg"])
def test_nan_rhs_reports_noconv(method, pkg, O):
    """A NaN in one right-hand side: that column stops, the solve reports
    MFD_ERR_NOCONV with relres NaN, and the other column still converges."""
    from maxwellfdfd import workloads as W
    m, om = W.sphere_dipole(N=10, npml=2, r=2.0, lam=7.0)
    pkg.calc_matparams(m)
    je, jm = pkg.create_srcs(m)
    b = pkg.create_b(pkg.EE, om, m, je, jm)
    A = O.to_scipy(O.create_A(spec_of(m), om, m.εarr, m.μarr))
    bad = b.copy()
    bad[5] = np.nan
    B = np.stack([b, bad])
    op = pkg.MaxwellOperator(m, om)
    say("nan rhs", method)
    with pytest.raises(pkg.MfdNoConvergence):
        op.solve(B, tol=1e-8, method=method)
    E, st = op.solve(B, tol=1e-8, method=method, raise_on_noconv=False)
    op.close()
    assert not st.converged and np.isnan(st.relres)
    assert np.linalg.norm(b - A @ E[0]) / np.linalg.norm(b) <= 1.05e-8


def test_bicgstab_probes(pkg, O, capfd, monkeypatch):
    """BiCGSTAB cycles probe the true FP64 residual every 2,048 steps: on the
    64³ random-box system (thousands of BiCGSTAB steps) the probes run, the
    cycle ends on a probe that meets the tolerance, and the returned iterate is
    no worse than any probe and meets the tolerance on the oracle's A."""
    import re
    from maxwellfdfd import workloads as W
    m, om = W.random_boxes(N=64, npml=10, nbox=40, lam=24.0, seed=1, ndip=1)
    pkg.calc_matparams(m)
    je, jm = pkg.create_srcs(m)
    b = pkg.create_b(pkg.EE, om, m, je, jm)
    A = O.to_scipy(O.create_A(spec_of(m), om, m.εarr, m.μarr))
    monkeypatch.setenv("MFD_SOLVE_TRACE", "1")
    op = pkg.MaxwellOperator(m, om)
    say("bicgstab probes")
    e, st = op.solve(b, tol=1e-8, maxit=60000, method="bicgstab", raise_on_noconv=False)
    op.close()
    err = capfd.readouterr().err
    probes = [float(v) for v in re.findall(r"\[mfd probe\] it \d+ true residual ([0-9.e+-]+)", err)]
    rel = np.linalg.norm(b - A @ e) / np.linalg.norm(b)
    say("  iterations", st.iterations, "probes", len(probes), probes[-3:], "returned", rel, "relres", st.relres)
    assert st.iterations > 4096, "premise: a solve long enough for probes"
    assert probes, "no probe ran"
    assert st.converged and rel <= 1.05e-8
    assert st.relres <= min(probes) * (1 + 1e-9)
