"""Device path (libmfdfd through its C ABI) against the oracle.

CSC index structure (colptr/rowval) must be bit-exact; values agree to 1e-13
relative to the matrix's largest entry; the matrix-free operator to 1e-12;
solved fields to 1e-7 relative at FP64 residual 1e-8 (the drop-in contract
for operator values and solved fields).
"""
import sys

import numpy as np
import pytest
import scipy.sparse.linalg as sla

from conftest import shapes_of, spec_of

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def nonuniform(N, seed):
    rng = np.random.default_rng(seed)
    return [np.cumsum(np.concatenate(([0.0], rng.uniform(0.6, 1.4, n)))) for n in N]


CASES = {
    "uniform_pml": dict(N=(6, 6, 6), npml=((2, 2, 2), (2, 2, 2)), om=0.7, bloch=(False,) * 3),
    "nonuniform_bloch": dict(N=(5, 4, 6), npml=((0,) * 3, (0,) * 3), om=0.9, bloch=(True,) * 3,
                             kb=(0.3, -0.2, 0.5), nu=1),
    "mixed_pml_bloch": dict(N=(4, 5, 3), npml=((1, 0, 1), (1, 1, 0)), om=0.6 + 0.05j, bloch=(True, False, True),
                            kb=(0.1, 0.0, 0.4), nu=2),
    "tiny_bloch_axes": dict(N=(1, 2, 5), npml=((0, 0, 1), (0, 0, 1)), om=0.5, bloch=(True, True, False),
                            kb=(0.0, 0.3, 0.0)),
    "cmplast": dict(N=(4, 3, 5), npml=((1, 1, 1), (1, 1, 1)), om=0.8, bloch=(False,) * 3, cmpfirst=False),
    "boundft_HH": dict(N=(4, 4, 4), npml=((1, 1, 1), (1, 1, 1)), om=0.8, bloch=(False, True, False),
                       bft=(0, 1, 0)),
    "wide": dict(N=(70, 5, 7), npml=((3, 1, 1), (2, 1, 2)), om=0.45, bloch=(False,) * 3),
    "wide_bloch": dict(N=(67, 3, 5), npml=((0, 0, 1), (0, 0, 1)), om=0.5, bloch=(True, True, False),
                       kb=(0.02, 0.3, 0.0)),
    # one-cell z axes (the extruded 2D models), with a z phase and without
    "slab_bloch": dict(N=(9, 7, 1), npml=((2, 1, 0), (1, 2, 0)), om=0.6, bloch=(False, False, True),
                       kb=(0.0, 0.0, 0.35)),
    "slab_2d": dict(N=(66, 5, 1), npml=((0, 1, 0), (0, 1, 0)), om=0.5, bloch=(True, False, True),
                    kb=(0.1, 0.0, 0.0)),
    "slab_pec": dict(N=(8, 6, 1), npml=((1, 0, 0), (1, 0, 0)), om=0.6, bloch=(False, True, False),
                     kb=(0.0, 0.25, 0.0), nu=7),
}


def make(pkg, case):
    c = CASES[case] if case in CASES else HHCASES[case]
    N = c["N"]
    lp = nonuniform(N, c["nu"]) if "nu" in c else [np.arange(n + 1.0) for n in N]
    g = pkg.Grid(lp, c["bloch"])
    m = pkg.ModelFull(g)
    pkg.set_ωpml(m, c["om"])
    pkg.set_Npml(m, c["npml"])
    pkg.set_kbloch(m, c.get("kb", (0, 0, 0)))
    pkg.set_boundft(m, c.get("bft", (0, 0, 0)))
    m.order_cmpfirst = c.get("cmpfirst", True)
    return m


def rand_params(m, seed, full=False, mu=False):
    rng = np.random.default_rng(seed)
    N = m.grid.N
    eps = np.zeros(N + (3, 3), dtype=np.complex128, order="F")
    muu = np.zeros_like(eps)
    for c in range(3):
        eps[..., c, c] = 1 + 11 * rng.random(N) + 0.1j * rng.random(N)
        muu[..., c, c] = (1 + 0.5 * rng.random(N)) if mu else 1.0
    if full:
        for (v, w) in ((0, 1), (0, 2), (1, 2)):
            x = rng.normal(size=N) + 0.1j * rng.normal(size=N)
            x[rng.random(N) < 0.5] = 0
            eps[..., v, w] = x
            eps[..., w, v] = x
    return eps, muu


def assert_csc_equal(got, ref, rtol, what):
    gc, gr, gv = got
    rc, rr, rv = ref
    assert np.array_equal(gc, rc), f"{what}: colptr differs"
    assert np.array_equal(gr, rr), f"{what}: rowval differs"
    scale = max(np.abs(rv).max(), 1e-300)
    err = np.abs(gv - rv).max() if len(rv) else 0.0
    assert err <= rtol * scale, f"{what}: values differ by {err / scale:.2e} (relative to max)"


@pytest.mark.parametrize("case", list(CASES))
def test_stretched_dl(case, pkg, O):
    m = make(pkg, case)
    say("stretched_dl", case)
    got = pkg.create_stretched_dls(m)
    T = O.grid_tables(spec_of(m))
    for gt in range(2):
        for w in range(3):
            np.testing.assert_allclose(got[gt][w], T[("sdl", gt, w)], rtol=1e-14, atol=0)


@pytest.mark.parametrize("case", list(CASES))
def test_curl_csc(case, pkg, O):
    m = make(pkg, case)
    s = spec_of(m)
    for which in (0, 1):
        say("curl_csc", case, which)
        assert_csc_equal(pkg.curl_csc(m, which), O.curl(s, which), 1e-14, f"curl{which}")


@pytest.mark.parametrize("case", list(CASES))
def test_Mcs_csc(case, pkg, O):
    """create_Mcs on the device (per-column enumeration) against the oracle's
    per-corner restatement: identical structure, values to 1e-14."""
    m = make(pkg, case)
    s = spec_of(m)
    for which in (0, 1):
        say("Mcs_csc", case, which)
        assert_csc_equal(pkg.Mcs_csc(m, which), O.Mcs(s, which), 1e-14, f"Mcs{which}")
    Mce, Mcm = pkg.create_Mcs(m)
    assert Mce.shape == Mcm.shape == (3 * m.grid.M,) * 2


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("full", [False, True])
def test_paramop_csc(case, full, pkg, O):
    m = make(pkg, case)
    s = spec_of(m)
    eps, mu = rand_params(m, 3, full=full, mu=True)
    for which, arr in ((0, eps), (1, mu)):
        say("paramop_csc", case, full, which)
        assert_csc_equal(pkg.paramop_csc(m, which, arr), O.paramop(s, which, arr), 1e-13, f"param{which}")


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("full", [False, True])
def test_A_csc(case, full, pkg, O):
    m = make(pkg, case)
    s = spec_of(m)
    eps, mu = rand_params(m, 4, full=full, mu=True)
    for om in (0.55, 0.0):
        say("A_csc", case, full, om)
        assert_csc_equal(pkg.A_csc(m, om, eps, mu), O.create_A(s, om, eps, mu), 1e-13, "A")


@pytest.mark.parametrize("case", list(CASES))
def test_A_csc_hh(case, pkg, O):
    """HH formulation: A = Cₑ (Pε \\ Cₘ) − ω² Pμ (full μ allowed), bit-exact structure."""
    m = make(pkg, case)
    s = spec_of(m)
    mu, eps = rand_params(m, 14, full=True, mu=True)   # ε diagonal, μ with off-diagonals
    for om in (0.55, 0.0):
        say("A_csc HH", case, om)
        assert_csc_equal(pkg.A_csc(m, om, eps, mu, ft=pkg.HH), O.create_A(s, om, eps, mu, ft=1), 1e-13, "A_HH")


def test_A_rejects_nondiagonal_divisor(pkg):
    m = make(pkg, "uniform_pml")
    eps, mu = rand_params(m, 5)
    mu2 = mu.copy()
    mu2[..., 0, 1] = 0.1
    with pytest.raises(pkg.MfdUnsupportedError):
        pkg.A_csc(m, 0.5, eps, mu2)
    eps2 = eps.copy()
    eps2[..., 0, 2] = 0.1
    with pytest.raises(pkg.MfdUnsupportedError):
        pkg.A_csc(m, 0.5, eps2, mu, ft=pkg.HH)
    with pytest.raises(pkg.MfdUnsupportedError):
        pkg.create_b(pkg.HH, 0.5, m, np.zeros(3 * m.grid.M, complex), np.zeros(3 * m.grid.M, complex), εarr=eps2)
    m2 = make(pkg, "uniform_pml")
    pkg.set_ωpml(m2, 0.0)
    with pytest.raises(pkg.MfdArgumentError):
        pkg.create_stretched_dls(m2)


@pytest.mark.parametrize("case", ["uniform_pml", "mixed_pml_bloch", "cmplast", "boundft_HH", "hh_pml", "hh_all_bloch"])
def test_create_b(case, pkg, O):
    m = make(pkg, case)
    s = spec_of(m)
    eps, mu = rand_params(m, 6, mu=True)
    rng = np.random.default_rng(1)
    n = 3 * m.grid.M
    je = rng.normal(size=n) + 1j * rng.normal(size=n)
    jm = rng.normal(size=n) + 1j * rng.normal(size=n)
    for om in (0.7, 0.0):
        say("create_b", case, om)
        got = pkg.create_b(pkg.EE, om, m, je, jm, μarr=mu)
        ref = O.create_b(s, om, mu, je, jm)
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
        got = pkg.create_b(pkg.HH, om, m, je, jm, εarr=eps)
        ref = O.create_b(s, om, mu, je, jm, ft=1, eps_arr=eps)
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


def geom_model(pkg, N=(14, 12, 13), nonuni=True, seed=0):
    lp = nonuniform(N, 30 + seed) if nonuni else [np.arange(n + 1.0) for n in N]
    g = pkg.Grid(lp)
    m = pkg.ModelFull(g)
    pkg.set_ωpml(m, 0.6)
    pkg.set_Npml(m, ((2, 2, 2), (2, 2, 2)))
    L = [a[-1] - a[0] for a in lp]
    c0 = [0.5 * (a[-1] + a[0]) for a in lp]
    pkg.add_obj(m, "bg", pkg.Cuboid(tuple(c0), tuple(L)), ε=1.0)
    rng = np.random.default_rng(seed)
    for k in range(4):
        pkg.add_obj(m, f"box{k}", pkg.Cuboid(tuple(rng.uniform(3, 9, 3)), tuple(rng.uniform(1.5, 5, 3))),
                    ε=rng.uniform(2, 12))
    pkg.add_obj(m, "ball", pkg.Ball(tuple(c0), 3.7), ε=12.0 + 0.5j)
    pkg.add_obj(m, "rod", pkg.Cylinder((c0[0] + 2.3, c0[1] - 1.9, c0[2]), 2.2, 6.5), ε=5.0)
    pkg.add_obj(m, "aniso", pkg.Ball((c0[0] + 2, c0[1] - 1, c0[2]), 1.9), ε=np.diag([3.0, 4.0, 5.0]), μ=1.0)
    return m


@pytest.mark.parametrize("nonuni", [False, True])
def test_matparams(nonuni, pkg, O):
    m = geom_model(pkg, nonuni=nonuni)
    say("matparams", nonuni)
    pkg.calc_matparams(m)
    shapes, et, mt = shapes_of(m)
    ea, ma = O.matparams(spec_of(m), shapes, et, mt)
    assert np.abs(m.εarr - ea).max() <= 1e-12 * np.abs(ea).max()
    assert np.abs(m.μarr - ma).max() <= 1e-12 * np.abs(ma).max()
    assert np.abs(ea[..., 0, 1]).max() > 0.1


OPCASES = ["uniform_pml", "nonuniform_bloch", "mixed_pml_bloch", "tiny_bloch_axes", "cmplast", "wide"]


@pytest.mark.parametrize("case", ["wide", "wide_bloch", "nonuniform_bloch", "tiny_bloch_axes", "mixed_pml_bloch",
                                  "slab_bloch", "slab_2d", "slab_pec"])
@pytest.mark.parametrize("ty,kz,xo", [(1, 1, 1), (1, 4, 0), (2, 3, 1), (2, 16, 0), (4, 2, 1), (4, 16, 0),
                                      (4, 16, 1), (8, 3, 1), (8, 16, 0), (0, 1, 0),
                                      (1, 3, 3), (2, 5, 2), (4, 16, 3), (8, 2, 3)])
@pytest.mark.parametrize("params", ["diag", "full", "mu"])
def test_apply_tile_shapes(case, ty, kz, xo, params, pkg, O):
    """Every tile shape and tile order of the stencil (and the direct kernel,
    TY = 0) gives the CSC product, with diagonal ε, with full ε (the corner
    off-diagonal terms read planes k−1, k, k+1 and wrap with the Bloch phase)
    and with μ ≠ 1 (1/μ weights on the H plane).  Orders 2 and 3 ask for
    32-wide tiles, which the diagonal-ε, μ = 1 stencil uses and the others
    ignore."""
    m = make(pkg, case)
    eps, muu = rand_params(m, 8, full=params == "full", mu=params == "mu")
    om = 0.55
    say("tiles", case, ty, kz, xo, params)
    A = O.to_scipy(O.create_A(spec_of(m), om, eps, muu))
    op = pkg.MaxwellOperator(m, om, εarr=eps, μarr=muu)
    op.set_tile("fp64", ty, kz, xo)
    rng = np.random.default_rng(4)
    n = 3 * m.grid.M
    X = rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n))
    Y = op.apply(X)
    for r in range(2):
        ref = A @ X[r]
        assert np.linalg.norm(Y[r] - ref) <= 1e-12 * np.linalg.norm(ref)
    op.close()


# boundft = HH on some axes (an EE model so configured, or the HH formulation,
# which flips every axis): the tiled stencil on the grid reflected along them
REVCASES = {
    "rev_pml": dict(N=(6, 7, 5), npml=((2, 1, 2), (1, 2, 1)), om=0.7, bloch=(False,) * 3, nu=5),
    "rev_bloch": dict(N=(5, 4, 6), npml=((0,) * 3, (0,) * 3), om=0.9, bloch=(True,) * 3,
                      kb=(0.3, -0.2, 0.5), nu=1),
    "rev_mixed": dict(N=(4, 5, 3), npml=((1, 0, 1), (1, 1, 0)), om=0.6 + 0.05j, bloch=(True, False, True),
                      kb=(0.1, 0.0, 0.4), nu=2),
    "rev_wide": dict(N=(70, 5, 7), npml=((3, 1, 1), (2, 1, 2)), om=0.45, bloch=(False,) * 3, nu=6),
    "rev_cmplast": dict(N=(4, 3, 5), npml=((1, 1, 1), (1, 1, 1)), om=0.8, bloch=(False,) * 3, cmpfirst=False),
    "rev_slab": dict(N=(9, 7, 1), npml=((2, 1, 0), (1, 2, 0)), om=0.6, bloch=(False, True, True),
                     kb=(0.0, 0.2, 0.35)),
    "rev_tiny": dict(N=(1, 2, 5), npml=((0, 0, 1), (0, 0, 1)), om=0.5, bloch=(True, True, False),
                     kb=(0.0, 0.3, 0.0)),
}


def make_rev(pkg, case, bft):
    """The model of REVCASES[case] with the given boundft per axis."""
    c = dict(REVCASES[case])
    c["bft"] = bft
    CASES["_rev"] = c
    try:
        return make(pkg, "_rev")
    finally:
        del CASES["_rev"]


@pytest.mark.parametrize("case", list(REVCASES))
@pytest.mark.parametrize("ty,kz,xo", [(1, 1, 1), (1, 4, 0), (2, 3, 1), (4, 2, 1), (4, 16, 0), (8, 3, 1),
                                      (8, 16, 0), (0, 1, 0)])
@pytest.mark.parametrize("ft,bft", [(0, (1, 1, 1)), (0, (0, 1, 0)), (0, (1, 0, 1)), (0, (0, 0, 1)),
                                    (1, (0, 0, 0)), (1, (1, 0, 0))])
@pytest.mark.parametrize("mu", [False, True])
def test_apply_tile_shapes_reflected(case, ty, kz, xo, ft, bft, mu, pkg, O):
    """boundft = HH on some or all axes, in the EE formulation and in the HH
    formulation (which flips every axis): every tile shape of the stencil on
    the reflected grid (and the direct kernel, TY = 0) against the oracle's
    CSC of that formulation, with and without μ ≠ 1 (in the HH formulation ε
    and μ trade places, so ε ≠ 1 there is the stencil's 1/μ weight)."""
    m = make_rev(pkg, case, bft)
    eps, muu = rand_params(m, 18, mu=mu)
    om = 0.55 + 0.02j
    say("reflected tiles", case, ty, kz, xo, ft, bft, mu)
    A = O.to_scipy(O.create_A(spec_of(m), om, eps, muu, ft=ft))
    op = pkg.MaxwellOperator(m, om, ft=pkg.HH if ft else pkg.EE, εarr=eps, μarr=muu)
    op.set_tile("fp64", ty, kz, xo)
    rng = np.random.default_rng(11)
    n = 3 * m.grid.M
    X = rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n))
    Y = op.apply(X)
    op.close()
    for r in range(2):
        ref = A @ X[r]
        assert np.linalg.norm(Y[r] - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("case,bft", [("rev_pml", (0, 0, 0)), ("rev_bloch", (0, 0, 0)), ("rev_slab", (0, 0, 0)),
                                      ("rev_pml", (0, 1, 1)), ("rev_mixed", (1, 0, 0))])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_reflected(case, bft, precision, pkg, O):
    """The HH formulation's solve on the reflected stencil (every axis, or
    the axes whose boundft is EE, flipped) against the direct solve, and
    against the same solve on the direct kernel."""
    m = make_rev(pkg, case, bft)
    muu, eps = rand_params(m, 23, mu=True)
    om = 0.5 + 0.03j
    A = O.to_scipy(O.create_A(spec_of(m), om, eps, muu, ft=1))
    b = np.random.default_rng(12).normal(size=A.shape[0]) + 0j
    say("solve reflected", case, bft, precision)
    xd = sla.spsolve(A.tocsc(), b)
    out = []
    for ty in (4, 0):
        op = pkg.MaxwellOperator(m, om, ft=pkg.HH, εarr=eps, μarr=muu)
        op.set_tile("fp64", ty, 4, 1)
        op.set_tile("fp32", ty, 4, 1)
        h, st = op.solve(b, tol=1e-9, precision=precision)
        op.close()
        assert st.converged and np.linalg.norm(b - A @ h) / np.linalg.norm(b) <= 1.05e-9
        assert np.linalg.norm(h - xd) / np.linalg.norm(xd) <= 1e-6
        out.append(h)
    assert np.linalg.norm(out[0] - out[1]) / np.linalg.norm(out[1]) <= 1e-6


@pytest.mark.parametrize("case", OPCASES)
@pytest.mark.parametrize("full,mu", [(False, False), (True, False), (False, True), (True, True)])
def test_apply_matches_csc(case, full, mu, pkg, O):
    m = make(pkg, case)
    s = spec_of(m)
    eps, muu = rand_params(m, 7, full=full, mu=mu)
    om = 0.55
    say("apply", case, full, mu)
    A = O.to_scipy(O.create_A(s, om, eps, muu))
    op = pkg.MaxwellOperator(m, om, εarr=eps, μarr=muu)
    rng = np.random.default_rng(2)
    n = 3 * m.grid.M
    X = rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))
    Y = op.apply(X)
    for r in range(3):
        ref = A @ X[r]
        assert np.linalg.norm(Y[r] - ref) <= 1e-12 * np.linalg.norm(ref)
    op.close()


@pytest.mark.parametrize("seed", range(24))
def test_apply_random_configs(seed, pkg, O):
    """Seeded random configurations against the oracle's CSC A: grid sizes
    1–21 (one axis sometimes 60–79, past a 64-wide tile), uniform or
    nonuniform spacing, Bloch axes with real phases, PML on random ends,
    boundft and ft per case, μ ≠ 1, full ε where allowed, and a random tile
    shape (TY = 0: the direct kernel)."""
    rng = np.random.default_rng(1000 + seed)
    N = tuple(int(v) for v in rng.integers(1, 22, 3))
    if rng.random() < 0.2:
        N = (int(rng.integers(60, 80)),) + N[1:]
    bloch = tuple(bool(N[a] == 1 or rng.random() < 0.4) for a in range(3))
    npml = [[0, 0, 0], [0, 0, 0]]
    for a in range(3):
        for e in range(2):
            if N[a] >= 6 and rng.random() < 0.6:
                npml[e][a] = int(rng.integers(1, 3))
    kb = tuple(float(rng.uniform(-0.4, 0.4)) if bloch[a] and rng.random() < 0.5 else 0.0 for a in range(3))
    bft = tuple(int(rng.random() < 0.3) for a in range(3))
    ft = int(rng.random() < 0.3)
    lp = ([np.cumsum(np.concatenate(([0.0], rng.uniform(0.6, 1.4, n)))) for n in N] if rng.random() < 0.5
          else [np.arange(n + 1.0) for n in N])
    m = pkg.ModelFull(pkg.Grid(lp, bloch))
    pkg.set_ωpml(m, 0.7)
    pkg.set_Npml(m, (tuple(npml[0]), tuple(npml[1])))
    pkg.set_kbloch(m, kb)
    pkg.set_boundft(m, bft)
    mu_on = rng.random() < 0.3
    full = ft == 0 and bft == (0, 0, 0) and rng.random() < 0.3
    eps = np.zeros(N + (3, 3), dtype=np.complex128, order="F")
    muu = np.zeros_like(eps)
    for c in range(3):
        eps[..., c, c] = 1 + 11 * rng.random(N) + 0.1j * rng.random(N)
        muu[..., c, c] = (1 + 0.5 * rng.random(N)) if mu_on else 1.0
    if full:
        for (v, w) in ((0, 1), (0, 2), (1, 2)):
            x = rng.normal(size=N) + 0.1j * rng.normal(size=N)
            x[rng.random(N) < 0.5] = 0
            eps[..., v, w] = x
            eps[..., w, v] = x
    om = 0.55 + 0.03j * rng.random()
    ty, kz, xo = int(rng.choice([0, 1, 2, 4, 8])), int(rng.integers(1, 17)), int(rng.integers(0, 2))
    say("random config", seed, N, bloch, npml, bft, ft, mu_on, full, (ty, kz, xo))
    A = O.to_scipy(O.create_A(spec_of(m), om, eps, muu, ft=ft))
    op = pkg.MaxwellOperator(m, om, ft=pkg.HH if ft else pkg.EE, εarr=eps, μarr=muu)
    op.set_tile("fp64", ty, kz, xo)
    n = 3 * m.grid.M
    X = rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n))
    Y = op.apply(X)
    op.close()
    for r in range(2):
        ref = A @ X[r]
        assert np.linalg.norm(Y[r] - ref) <= 1e-12 * np.linalg.norm(ref)


HHCASES = {
    "boundft_HH": CASES["boundft_HH"],
    "hh_pml": dict(N=(5, 6, 4), npml=((1, 1, 0), (1, 0, 1)), om=0.7, bloch=(False,) * 3, bft=(1, 0, 1), nu=3),
    "hh_all_bloch": dict(N=(4, 5, 3), npml=((0, 0, 1), (0, 0, 1)), om=0.6, bloch=(True, True, False),
                         kb=(0.2, 0.0, 0.0), bft=(1, 1, 1), nu=4),
}


@pytest.mark.parametrize("case", list(HHCASES))
@pytest.mark.parametrize("mu", [False, True])
def test_apply_boundft_hh(case, mu, pkg, O):
    """boundft = HH on some axes: the matrix-free EE operator (direct kernel,
    difference directions per axis) against the oracle's CSC; off-diagonal ε
    there is refused."""
    m = make(pkg, case)
    eps, muu = rand_params(m, 17, mu=mu)
    om = 0.55
    say("apply boundft HH", case, mu)
    A = O.to_scipy(O.create_A(spec_of(m), om, eps, muu))
    op = pkg.MaxwellOperator(m, om, εarr=eps, μarr=muu)
    n = 3 * m.grid.M
    X = np.random.default_rng(3).normal(size=(2, n)) + 1j * np.random.default_rng(4).normal(size=(2, n))
    Y = op.apply(X)
    for r in range(2):
        ref = A @ X[r]
        assert np.linalg.norm(Y[r] - ref) <= 1e-12 * np.linalg.norm(ref)
    op.close()
    epsf, _ = rand_params(m, 17, full=True)
    if any(b == 1 for b in m.boundft):
        with pytest.raises(pkg.MfdUnsupportedError):
            pkg.MaxwellOperator(m, om, εarr=epsf, μarr=muu)


@pytest.mark.parametrize("case", ["uniform_pml", "hh_pml", "boundft_HH"])
def test_operator_ft_hh(case, pkg, O):
    """The HH formulation Cₑ (Pε \\ Cₘ) − ω² Pμ as a device operator (the dual
    EE operator of the flipped model) against the oracle's HH CSC, and its
    solve against the direct solve; the EE-only field conversions refuse it."""
    m = make(pkg, case)
    muu, eps = rand_params(m, 19, mu=True)   # ε diagonal (the divisor), μ ≠ 1
    om = 0.5 + 0.03j
    say("operator ft=HH", case)
    A = O.to_scipy(O.create_A(spec_of(m), om, eps, muu, ft=1))
    op = pkg.MaxwellOperator(m, om, ft=pkg.HH, εarr=eps, μarr=muu)
    n = 3 * m.grid.M
    X = np.random.default_rng(5).normal(size=n) + 1j * np.random.default_rng(6).normal(size=n)
    ref = A @ X
    assert np.linalg.norm(op.apply(X) - ref) <= 1e-12 * np.linalg.norm(ref)
    b = np.random.default_rng(7).normal(size=n) + 1j * np.random.default_rng(8).normal(size=n)
    h, st = op.solve(b, tol=1e-9)
    xd = sla.spsolve(A.tocsc(), b)
    assert st.converged and np.linalg.norm(b - A @ h) / np.linalg.norm(b) <= 1.05e-9
    assert np.linalg.norm(h - xd) / np.linalg.norm(xd) <= 1e-6
    with pytest.raises(pkg.MfdUnsupportedError):
        op.h_from_e(h)
    op.close()


@pytest.mark.parametrize("case", list(HHCASES))
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_boundft_hh(case, precision, pkg, O):
    """EE solve with boundft = HH on some axes against the direct solve."""
    m = make(pkg, case)
    eps, muu = rand_params(m, 21)
    om = 0.45 + 0.05j
    A = O.to_scipy(O.create_A(spec_of(m), om, eps, muu))
    op = pkg.MaxwellOperator(m, om, εarr=eps, μarr=muu)
    b = np.random.default_rng(9).normal(size=A.shape[0]) + 0j
    say("solve boundft HH", case, precision)
    e, st = op.solve(b, tol=1e-9, precision=precision)
    xd = sla.spsolve(A.tocsc(), b)
    assert st.converged and np.linalg.norm(b - A @ e) / np.linalg.norm(b) <= 1.05e-9
    assert np.linalg.norm(e - xd) / np.linalg.norm(xd) <= 1e-6
    op.close()


@pytest.mark.parametrize("case", ["uniform_pml", "nonuniform_bloch", "mixed_pml_bloch", "cmplast", "wide",
                                  "boundft_HH", "hh_pml", "hh_all_bloch"])
@pytest.mark.parametrize("mu", [False, True])
def test_h_from_e_e_from_h(case, mu, pkg, O):
    """h_from_e / e_from_h (model.jl:276-284) against the oracle's Cₑ, Cₘ, Pε, Pμ
    CSC products; and the identity e_from_h(h_from_e(e)) = e + ω⁻² Pε⁻¹ A e."""
    import scipy.sparse.linalg as spla
    m = make(pkg, case)
    s = spec_of(m)
    eps, muu = rand_params(m, 9, mu=mu)
    om = 0.55 - 0.02j
    say("field convert", case, mu)
    Ce, Cm = (O.to_scipy(O.curl(s, w)).tocsc() for w in (0, 1))
    Pe, Pm = (O.to_scipy(O.paramop(s, w, a)).tocsc() for w, a in ((0, eps), (1, muu)))
    A = O.to_scipy(O.create_A(s, om, eps, muu))
    op = pkg.MaxwellOperator(m, om, εarr=eps, μarr=muu)
    rng = np.random.default_rng(4)
    n = 3 * m.grid.M
    cr = lambda: rng.normal(size=n) + 1j * rng.normal(size=n)
    e, jm, je = cr(), cr(), cr()
    href = (1j / om) * spla.spsolve(Pm, Ce @ e + jm)
    h = op.h_from_e(e, jm)
    assert np.linalg.norm(h - href) <= 1e-12 * np.linalg.norm(href)
    eref = (-1j / om) * spla.spsolve(Pe, Cm @ h - je)
    e2 = op.e_from_h(h, je)
    assert np.linalg.norm(e2 - eref) <= 1e-12 * np.linalg.norm(eref)
    rt = op.e_from_h(op.h_from_e(e))
    ident = e + spla.spsolve(Pe, A @ e) / om**2
    assert np.linalg.norm(rt - ident) <= 1e-11 * np.linalg.norm(ident)
    op.close()


def test_e_from_h_rejects_full_eps(pkg):
    m = make(pkg, "uniform_pml")
    eps, muu = rand_params(m, 9, full=True)
    op = pkg.MaxwellOperator(m, 0.5, εarr=eps, μarr=muu)
    with pytest.raises(pkg.MfdUnsupportedError):
        op.e_from_h(np.zeros(3 * m.grid.M, dtype=complex))
    op.close()


def c1_like(pkg, N=12, npml=3, r=2.5, lam=8.0):
    from maxwellfdfd import workloads as W
    m, om = W.sphere_dipole(N=N, npml=npml, r=r, lam=lam)
    pkg.calc_matparams(m)
    je, jm = pkg.create_srcs(m)
    b = pkg.create_b(pkg.EE, om, m, je, jm)
    return m, om, b


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_vs_direct(precision, pkg, O):
    m, om, b = c1_like(pkg)
    say("solve vs direct", precision)
    A = O.to_scipy(O.create_A(spec_of(m), om, m.εarr, m.μarr))
    op = pkg.MaxwellOperator(m, om)
    e, st = op.solve(b, tol=1e-8, precision=precision)
    rel = np.linalg.norm(b - A @ e) / np.linalg.norm(b)
    say("  iterations", st.iterations, "cycles", st.outer_cycles, "relres", st.relres, "oracle residual", rel)
    assert st.converged and rel <= 1.05e-8
    xd = sla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(e - xd) / np.linalg.norm(xd) <= 1e-6
    op.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_narrow_tiles(precision, pkg, O):
    """Solves with 32-wide stencil tiles (x size 36: one full tile and one of
    4 columns) against the oracle's A and the same solve with 64-wide tiles;
    two right-hand sides, so the RHS-interleaved tile order runs too."""
    m, om, b = c1_like(pkg, N=36, npml=4, r=7.0, lam=16.0)
    A = O.to_scipy(O.create_A(spec_of(m), om, m.εarr, m.μarr))
    B = np.array([b, 1j * b[::-1].copy()])
    out = {}
    for xo in (1, 3):
        op = pkg.MaxwellOperator(m, om)
        op.set_tile("fp64", 4, 8, xo)
        op.set_tile("mixed", 2, 4, xo)
        say("narrow tiles", precision, xo)
        E, st = op.solve(B, tol=1e-9, precision=precision)
        op.close()
        for r in range(2):
            rel = np.linalg.norm(B[r] - A @ E[r]) / np.linalg.norm(B[r])
            say("  order", xo, "rhs", r, "iterations", st.iterations, "oracle residual", rel)
            assert st.converged and rel <= 1.05e-9
        out[xo] = E
    for r in range(2):
        assert np.linalg.norm(out[3][r] - out[1][r]) <= 1e-7 * np.linalg.norm(out[1][r])


@pytest.mark.parametrize("kw", [dict(), dict(use_graph=False, check_every=5), dict(check_every=7),
                                dict(precision="mixed")], ids=["graph", "nograph_odd", "graph_odd", "mixed"])
def test_solve_cocg(kw, pkg, O):
    """COCG (fused step: p formed in the stencil, deferred iterate update) against
    the direct solve, single and batched RHS."""
    m, om, b = c1_like(pkg)
    A = O.to_scipy(O.create_A(spec_of(m), om, m.εarr, m.μarr))
    op = pkg.MaxwellOperator(m, om)
    say("cocg", kw)
    e, st = op.solve(b, tol=1e-8, method="cocg", **kw)
    rel = np.linalg.norm(b - A @ e) / np.linalg.norm(b)
    say("  iterations", st.iterations, "updates", st.outer_cycles, "oracle residual", rel)
    assert st.converged and rel <= 1.05e-8
    xd = sla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(e - xd) / np.linalg.norm(xd) <= 1e-6
    B = np.array([b, 1j * b[::-1].copy()])
    E, st = op.solve(B, tol=1e-8, method="cocg", **kw)
    for r in range(2):
        assert np.linalg.norm(B[r] - A @ E[r]) / np.linalg.norm(B[r]) <= 1.05e-8
    op.close()


@pytest.mark.parametrize("case", ["uniform_pml", "cmplast"])
def test_solve_cocg_direct_kernel(case, pkg, O):
    """COCG on operators served by the direct stencil (off-diagonal ε, μ ≠ 1),
    where p and the deferred update run as their own kernel."""
    m = make(pkg, case)
    eps, muu = rand_params(m, 5, full=True, mu=True)
    om = 0.45
    A = O.to_scipy(O.create_A(spec_of(m), om, eps, muu))
    op = pkg.MaxwellOperator(m, om, εarr=eps, μarr=muu)
    rng = np.random.default_rng(6)
    n = 3 * m.grid.M
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    for meth in ("cocg", "sqmr"):
        say("direct kernel", case, meth)
        e, st = op.solve(b, tol=1e-8, method=meth)
        assert np.linalg.norm(b - A @ e) / np.linalg.norm(b) <= 1.05e-8
    op.close()


@pytest.mark.parametrize("case", ["nonuniform_bloch", "mixed_pml_bloch"])
@pytest.mark.parametrize("full,mu", [(False, False), (True, False), (True, True)])
def test_solve_bloch_bicgstab(case, full, mu, pkg, O):
    """A complex Bloch phase makes A non-symmetric: the default solve uses
    BiCGSTAB; SQMR/COCG and the transpose-based adjoint refuse."""
    m = make(pkg, case)
    eps, muu = rand_params(m, 12, full=full, mu=mu)
    om = 0.55 + 0.02j
    A = O.to_scipy(O.create_A(spec_of(m), om, eps, muu))
    op = pkg.MaxwellOperator(m, om, εarr=eps, μarr=muu)
    rng = np.random.default_rng(13)
    n = 3 * m.grid.M
    B = rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n))
    say("bloch bicgstab", case, full)
    for prec in ("fp64", "mixed"):
        E, st = op.solve(B, tol=1e-8, precision=prec)
        for r in range(2):
            rel = np.linalg.norm(B[r] - A @ E[r]) / np.linalg.norm(B[r])
            assert st.converged and rel <= 1.05e-8
        assert st.iterations_fp32 == 0
    xd = sla.spsolve(A.tocsc(), B[0])
    assert np.linalg.norm(E[0] - xd) / np.linalg.norm(xd) <= 1e-6
    for meth in ("sqmr", "cocg"):
        with pytest.raises(pkg.MfdUnsupportedError):
            op.solve(B[0], method=meth)
    with pytest.raises(pkg.MfdUnsupportedError):
        op.solve_adjoint(B[0])
    op.close()


@pytest.mark.parametrize("case", ["uniform_pml", "cmplast", "wide"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_full_eps_tiled(case, precision, pkg, O):
    """Anisotropic ε (off-diagonals, μ = 1) through the tiled stencil: SQMR and
    COCG reach the oracle residual, and the tiled and direct kernels give the
    same field.  (The random 70-cell-wide medium needs loss to converge in a
    few thousand steps.)"""
    m = make(pkg, case)
    eps, muu = rand_params(m, 15, full=True)
    om = 0.45 + 0.15j if case == "wide" else 0.45
    A = O.to_scipy(O.create_A(spec_of(m), om, eps, muu))
    op = pkg.MaxwellOperator(m, om, εarr=eps, μarr=muu)
    rng = np.random.default_rng(16)
    n = 3 * m.grid.M
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    xd = sla.spsolve(A.tocsc(), b)
    for meth in ("sqmr", "cocg"):
        say("full eps tiled", case, precision, meth)
        e, st = op.solve(b, tol=1e-9, method=meth, precision=precision)
        assert st.converged and np.linalg.norm(b - A @ e) / np.linalg.norm(b) <= 1.05e-9
        assert np.linalg.norm(e - xd) / np.linalg.norm(xd) <= 1e-6
    op.set_tile("fp64", 0, 1, 0)
    op.set_tile("mixed", 0, 1, 0)
    e0, st0 = op.solve(b, tol=1e-9, precision=precision)
    assert st0.converged and np.linalg.norm(e0 - xd) / np.linalg.norm(xd) <= 1e-6
    op.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_streamed(precision, pkg, O):
    """A 100³ problem (3 M unknowns: its step exceeds the 256 MB Infinity
    Cache, so the vector updates stream non-temporally) converges, and its
    field meets the tolerance against the oracle's assembled A."""
    from maxwellfdfd import workloads as W
    m, om = W.random_boxes(N=100, npml=10, nbox=20, lam=16.0, seed=3, ndip=1)
    pkg.calc_matparams(m)
    je, jm = pkg.create_srcs(m)
    b = pkg.create_b(pkg.EE, om, m, je, jm)
    op = pkg.MaxwellOperator(m, om)
    say("streamed solve", precision)
    e, st = op.solve(b, tol=1e-8, precision=precision)
    say("  iterations", st.iterations, "fp32", st.iterations_fp32, "relres", st.relres)
    op.close()
    A = O.to_scipy(O.create_A(spec_of(m), om, m.εarr, m.μarr))
    rel = np.linalg.norm(b - A @ e) / np.linalg.norm(b)
    assert st.converged and rel <= 1.05e-8


def test_solve_bicgstab_symmetric(pkg, O):
    """BiCGSTAB chosen explicitly on a symmetric operator, against the direct solve."""
    m, om, b = c1_like(pkg)
    A = O.to_scipy(O.create_A(spec_of(m), om, m.εarr, m.μarr))
    op = pkg.MaxwellOperator(m, om)
    e, st = op.solve(b, tol=1e-8, method="bicgstab", use_graph=False, check_every=7)
    assert st.converged and np.linalg.norm(b - A @ e) / np.linalg.norm(b) <= 1.05e-8
    op.close()


def test_solve_c1_against_oracle(pkg, O):
    """Plain method, small: 32³, 8-cell PML, ε=12 sphere r=6, z-dipole, ω=2π/20;
    GPU FP64 solve vs the oracle CPU FP64 solve, both to residual 1e-10."""
    from maxwellfdfd import workloads as W
    m, om = W.sphere_dipole()
    pkg.calc_matparams(m)
    je, jm = pkg.create_srcs(m)
    b = pkg.create_b(pkg.EE, om, m, je, jm)
    s = spec_of(m)
    shapes, et, mt = shapes_of(m)
    ea, ma = O.matparams(s, shapes, et, mt)
    assert np.abs(m.εarr - ea).max() <= 1e-12 * 12
    A = O.create_A(s, om, ea, ma)
    say("c1: oracle solve")
    xo, it_o, rr_o = O.solve_sqmr(s, A, b, tol=1e-10, maxit=50000, nthreads=0)
    op = pkg.MaxwellOperator(m, om)
    say("c1: gpu solve")
    e, st = op.solve(b, tol=1e-10)
    As = O.to_scipy(A)
    rel = np.linalg.norm(b - As @ e) / np.linalg.norm(b)
    say("  gpu its", st.iterations, "res", rel, "cpu its", it_o, "res", rr_o)
    assert rel <= 1.05e-10 and rr_o <= 1e-10
    assert np.linalg.norm(e - xo) / np.linalg.norm(xo) <= 1e-7
    e8, st8 = op.solve(b, tol=1e-8, precision="mixed")
    rel8 = np.linalg.norm(b - As @ e8) / np.linalg.norm(b)
    say("  mixed its", st8.iterations, "cycles", st8.outer_cycles, "res", rel8,
        "field err", np.linalg.norm(e8 - xo) / np.linalg.norm(xo))
    assert rel8 <= 1.05e-8
    # fields to 1e-7 at residual 1e-8 (the drop-in contract), mixed precision included
    assert np.linalg.norm(e8 - xo) / np.linalg.norm(xo) <= 1e-7
    op.close()


def test_solve_batched(pkg, O):
    from maxwellfdfd import workloads as W
    m, om = W.random_boxes(N=16, npml=3, nbox=6, lam=8.0, seed=3, ndip=0)
    pkg.calc_matparams(m)
    dips = W.random_dipoles(16, 3, 3, 11)
    B = W.dipole_rhs(m, om, dips)
    A = O.to_scipy(O.create_A(spec_of(m), om, m.εarr, m.μarr))
    op = pkg.MaxwellOperator(m, om)
    for precision in ("fp64", "mixed"):
        say("batched", precision)
        E, st = op.solve(B, tol=1e-8, precision=precision)
        for r in range(3):
            rel = np.linalg.norm(B[r] - A @ E[r]) / np.linalg.norm(B[r])
            assert rel <= 1.05e-8
            e1, _ = op.solve(B[r], tol=1e-8, precision=precision)
            assert np.linalg.norm(E[r] - e1) / np.linalg.norm(e1) <= 1e-6
    op.close()


def test_solve_options(pkg, O):
    m, om, b = c1_like(pkg, N=10, npml=2, r=2.0, lam=7.0)
    A = O.to_scipy(O.create_A(spec_of(m), om, m.εarr, m.μarr))
    op = pkg.MaxwellOperator(m, om)
    for kw in (dict(use_graph=False), dict(check_every=5), dict(precision="mixed", inner_tol=1e-3)):
        say("solve options", kw)
        e, st = op.solve(b, tol=1e-8, **kw)
        assert np.linalg.norm(b - A @ e) / np.linalg.norm(b) <= 1.05e-8
    op.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_batch_uneven(precision, pkg, O):
    """Independent problems of different sizes, ε and ω in one lockstep batched
    solve: each field against the oracle CSC residual and the single solve."""
    from maxwellfdfd import workloads as W
    probs = W.uneven_problems(count=5, nmin=8, nmax=16, npml=2, nbox=4, seed=5)
    ops, bs, As = [], [], []
    for m, om in probs:
        pkg.calc_matparams(m)
        je, jm = pkg.create_srcs(m)
        bs.append(pkg.create_b(pkg.EE, om, m, je, jm))
        As.append(O.to_scipy(O.create_A(spec_of(m), om, m.εarr, m.μarr)))
        ops.append(pkg.MaxwellOperator(m, om))
    say("batch uneven", precision, [m.grid.N[0] for m, _ in probs])
    es, st = pkg.solve_batch(ops, bs, tol=1e-8, precision=precision)
    for p in range(len(ops)):
        rel = np.linalg.norm(bs[p] - As[p] @ es[p]) / np.linalg.norm(bs[p])
        say("  problem", p, "iterations", st[p].iterations, "fp32", st[p].iterations_fp32, "relres", rel)
        assert st[p].converged and rel <= 1.05e-8
        e1, _ = ops[p].solve(bs[p], tol=1e-8, precision="fp64")
        assert np.linalg.norm(es[p] - e1) / np.linalg.norm(e1) <= 1e-6
    with pytest.raises(pkg.MfdArgumentError):
        pkg.solve_batch([ops[0], ops[0]], [bs[0], bs[0]])
    with pytest.raises(pkg.MfdArgumentError):
        pkg.solve_batch([ops[0], ops[1]], [bs[1], bs[0]])
    for op in ops:
        op.close()


def test_bench_iteration_and_mixed_stats(pkg):
    """Per-kernel timing entry point, and the FP32/FP64 iteration split of a mixed solve."""
    m, om, b = c1_like(pkg, N=10, npml=2, r=2.0, lam=7.0)
    op = pkg.MaxwellOperator(m, om)
    for prec in ("fp64", "mixed"):
        say("bench iteration", prec)
        ms = op.bench_iteration(1, prec, 3)
        assert len(ms) == 3 and all(t > 0 for t in ms)
    e, st = op.solve(b, tol=1e-8, precision="mixed")
    assert st.converged and 0 < st.iterations_fp32 <= st.iterations
    e, st = op.solve(b, tol=1e-8, precision="fp64")
    assert st.converged and st.iterations_fp32 == 0
    op.close()


def test_solve_iteration_cap_reports(pkg):
    m, om, b = c1_like(pkg, N=10, npml=2, r=2.0, lam=7.0)
    op = pkg.MaxwellOperator(m, om)
    say("iteration cap")
    with pytest.raises(pkg.MfdNoConvergence):
        op.solve(b, tol=1e-8, maxit=5, check_every=5)
    e, st = op.solve(b, tol=1e-8, maxit=5, check_every=5, raise_on_noconv=False)
    assert not st.converged and st.iterations <= 10
    op.close()


@pytest.mark.parametrize("method", ["sqmr", "bicgstab"])
def test_noconv_returns_best_iterate(method, pkg, O):
    """A solve that ends without converging leaves the iterate with the
    smallest FP64 residual at a cycle start (never worse than the initial
    guess) and reports that iterate's residual."""
    m, om, b = c1_like(pkg, N=10, npml=2, r=2.0, lam=7.0)
    A = O.to_scipy(O.create_A(spec_of(m), om, m.εarr, m.μarr))
    op = pkg.MaxwellOperator(m, om)
    say("noconv best iterate", method)
    e, st = op.solve(b, tol=1e-15, maxit=600, check_every=8, method=method, raise_on_noconv=False)
    op.close()
    rel = np.linalg.norm(b - A @ e) / np.linalg.norm(b)
    assert not st.converged and st.relres <= 1.0
    assert abs(rel - st.relres) <= 1e-6 * st.relres + 1e-14


def test_adjoint(pkg, O):
    """Aᵀλ = g against a direct solve of the transpose; ∂f/∂ε against central
    finite differences (h = 1e-4) for f = Σ|e_z|² over a monitor."""
    m, om, b = c1_like(pkg, N=10, npml=2, r=2.0, lam=7.0)
    A = O.to_scipy(O.create_A(spec_of(m), om, m.εarr, m.μarr))
    op = pkg.MaxwellOperator(m, om)
    e, _ = op.solve(b, tol=1e-12)
    N = m.grid.N
    mon = [2 + 3 * (i + N[0] * (j + N[1] * 7)) for i in range(4, 6) for j in range(4, 6)]
    g = np.zeros_like(e)
    g[mon] = np.conj(e[mon])
    say("adjoint solve")
    lam, st = op.solve_adjoint(g, tol=1e-12)
    ref = sla.spsolve(A.T.tocsc(), g)
    assert np.linalg.norm(lam - ref) / np.linalg.norm(ref) <= 1e-9
    grad = op.adjoint_grad(e, lam)
    op.close()

    def f_of(eps_arr):
        o = pkg.MaxwellOperator(m, om, εarr=eps_arr)
        x, _ = o.solve(b, tol=1e-13)
        o.close()
        return np.sum(np.abs(x[mon]) ** 2)

    rng = np.random.default_rng(0)
    for t in range(4):
        c = int(rng.integers(3))
        i, j, k = (int(v) for v in rng.integers(3, 7, 3))
        h = 1e-4
        ep, em = m.εarr.copy(), m.εarr.copy()
        ep[i, j, k, c, c] += h
        em[i, j, k, c, c] -= h
        fd = (f_of(ep) - f_of(em)) / (2 * h)
        dof = c + 3 * (i + N[0] * (j + N[1] * k))
        say("  fd", fd, "adjoint", grad[dof])
        assert grad[dof] == pytest.approx(fd, rel=1e-5, abs=1e-9 * np.abs(grad).max())


def test_adjoint_design_region(pkg, O):
    """Adjoint case (design region ε = 1 + 11·sigmoid(G∗ξ), dipole, E_z monitor)
    at a reduced size: the pullback against the CPU FP64 adjoint (direct solve of
    the oracle's Aᵀ) and against central finite differences (h = 1e-4)."""
    from maxwellfdfd import workloads as W
    m, om, b, mon, sl = W.adjoint_design(N=(32, 32, 16), npml=4, design=(16, 16, 4), lam=16.0, seed=3,
                                         monitor=(2, 2, 1))
    A = O.to_scipy(O.create_A(spec_of(m), om, m.εarr, m.μarr)).tocsc()
    op = pkg.MaxwellOperator(m, om)
    say("adjoint design: forward")
    e, pullback = op.solve_with_pullback(b, tol=1e-12)
    g = np.zeros_like(e)
    g[mon] = np.conj(e[mon])
    say("adjoint design: pullback")
    grad = pullback(g)
    lam_ref = sla.spsolve(A.T.tocsc(), g)
    x_ref = sla.spsolve(A, b)
    grad_ref = 2 * np.real(om ** 2 * lam_ref * x_ref)
    assert np.linalg.norm(grad - grad_ref) <= 1e-8 * np.linalg.norm(grad_ref)
    op.close()
    N = m.grid.N

    def f_of(eps_arr):
        o = pkg.MaxwellOperator(m, om, εarr=eps_arr)
        x, _ = o.solve(b, tol=1e-13, x0=e)
        o.close()
        return np.sum(np.abs(x[mon]) ** 2)

    rng = np.random.default_rng(1)
    h = 1e-4
    for t in range(4):
        c = int(rng.integers(3))
        i, j, k = (int(rng.integers(s.start, s.stop)) for s in sl)
        ep, em = m.εarr.copy(), m.εarr.copy()
        ep[i, j, k, c, c] += h
        em[i, j, k, c, c] -= h
        fd = (f_of(ep) - f_of(em)) / (2 * h)
        dof = c + 3 * (i + N[0] * (j + N[1] * k))
        say("  voxel", (i, j, k, c), "fd", fd, "adjoint", grad[dof])
        assert grad[dof] == pytest.approx(fd, rel=1e-5, abs=1e-9 * np.abs(grad).max())


def model2d(pkg, kind, N=(14, 11), npml=2, bloch=(False, False), om=0.6, mu=1.0):
    T = pkg.twod
    lp = nonuniform(N, 21)
    m = (T.ModelTE if kind == "te" else T.ModelTM)(T.Grid2(lp, bloch))
    T.set_ωpml(m, om)
    T.set_Npml(m, ((npml, npml), (npml, npml)))
    c = [0.5 * (a[0] + a[-1]) for a in lp]
    T.add_obj(m, "bg", T.Rect(tuple(c), (lp[0][-1] - lp[0][0], lp[1][-1] - lp[1][0])), ε=1.0)
    T.add_obj(m, "rod", T.Disk((c[0] + 0.7, c[1] - 0.4), 2.6), ε=9.0 + 0.2j)
    T.add_obj(m, "bar", T.Rect((c[0] - 2.1, c[1] + 1.3), (1.7, 3.1)), ε=4.0, μ=mu)
    T.calc_matparams(m)
    return m


@pytest.mark.parametrize("kind", ["te", "tm"])
def test_model2d_operator_and_solve(kind, pkg, O):
    """2D TE/TM (te.jl, tm.jl) as the single-z-cell extruded 3D model: the
    polarisations decouple exactly, the 2D operator equals the oracle's
    extruded operator restricted to the 2D DOFs, and the solve matches the
    direct solve of the 2D CSC."""
    T = pkg.twod
    om = 0.6
    m = model2d(pkg, kind, om=om)
    s3 = spec_of(m.m3)
    shapes, et, mt = shapes_of(m.m3)
    ea, ma = O.matparams(s3, shapes, et, mt)
    assert np.abs(m.m3.εarr - ea).max() <= 1e-12 * np.abs(ea).max()
    A3 = O.to_scipy(O.create_A(s3, om, ea, ma)).tocsr()
    te, tm = m.esel(), (T.ModelTM if kind == "te" else T.ModelTE)(m.grid).esel()
    assert abs(A3[te][:, tm]).max() == 0 and abs(A3[tm][:, te]).max() == 0   # exact split
    A2 = T.create_A(pkg.EE, om, m)
    ref = A3[te][:, te]
    assert abs(A2 - ref).max() <= 1e-13 * abs(ref).max()
    T.add_srce(m, T.PointSrc2((7.1, 5.3), (1.0, 0.4) if kind == "te" else (1.0,)))
    je, jm = T.create_srcs(m)
    b = T.create_b(pkg.EE, om, m, je, jm)
    op = T.MaxwellOperator2D(m, om)
    say("model2d", kind)
    e, st = op.solve(b, tol=1e-10)
    xd = sla.spsolve(A2.tocsc(), b)
    assert np.linalg.norm(e - xd) / np.linalg.norm(xd) <= 1e-7
    assert np.linalg.norm(b - A2 @ e) / np.linalg.norm(b) <= 1.05e-10
    op.close()


def _oracle_2d(pkg, O, m, kind, om):
    """The oracle's extruded 3D operator restricted to the 2D model's E
    components, in the model's component-first order."""
    s3 = spec_of(m.m3)
    shapes, et, mt = shapes_of(m.m3)
    ea, ma = O.matparams(s3, shapes, et, mt)
    A3 = O.to_scipy(O.create_A(s3, om, ea, ma)).tocsr()
    te = m.esel()
    return A3[te][:, te].tocsc()


@pytest.mark.parametrize("kind", ["te", "tm"])
@pytest.mark.parametrize("bloch", [(False, False), (True, False)])
@pytest.mark.parametrize("ty,kz,xo", [(1, 1, 1), (2, 4, 0), (4, 16, 1), (8, 2, 1)])
def test_model2d_window_apply(kind, bloch, ty, kz, xo, pkg, O):
    """The 2D operator with the component window (K = 2 unknowns per cell
    for TE, 1 for TM, component-last on the device) against the oracle's
    extruded operator restricted to the 2D DOFs, for several tile shapes."""
    T = pkg.twod
    om = 0.6
    m = model2d(pkg, kind, bloch=bloch, om=om)
    A2 = _oracle_2d(pkg, O, m, kind, om)
    op = T.MaxwellOperator2D(m, om)
    assert op.n == A2.shape[0] == (2 if kind == "te" else 1) * m.grid.M
    op.op.set_tile("fp64", ty, kz, xo)
    say("model2d window apply", kind, bloch, ty, kz, xo)
    rng = np.random.default_rng(41)
    X = rng.normal(size=(2, op.n)) + 1j * rng.normal(size=(2, op.n))
    Y = op.apply(X)
    op.close()
    for r in range(2):
        ref = A2 @ X[r]
        assert np.linalg.norm(Y[r] - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("kind", ["te", "tm"])
@pytest.mark.parametrize("precision,method", [("fp64", None), ("mixed", None), ("fp64", "cocg"),
                                              ("mixed", "cocg"), ("fp64", "bicgstab")])
def test_model2d_window_solve(kind, precision, method, pkg, O):
    """Two right-hand sides through the windowed 2D solve (SQMR, COCG, whose
    step runs unfused with a window, and BiCGSTAB): each to 1e-9 in the
    oracle's residual, against the direct solve, and against the solve of the
    full single-cell 3D operator (three components per cell)."""
    T = pkg.twod
    om = 0.6
    m = model2d(pkg, kind, om=om)
    A2 = _oracle_2d(pkg, O, m, kind, om)
    rng = np.random.default_rng(43)
    B = rng.normal(size=(2, A2.shape[0])) + 1j * rng.normal(size=(2, A2.shape[0]))
    op = T.MaxwellOperator2D(m, om)
    say("model2d window solve", kind, precision, method)
    E, st = op.solve(B, tol=1e-9, precision=precision, method=method)
    op.close()
    full = pkg.MaxwellOperator(m.m3, om)
    sel = m.esel()
    B3 = np.zeros((2, full.n), dtype=complex)
    B3[:, sel] = B
    E3, _ = full.solve(B3, tol=1e-9, precision=precision, method=method)
    full.close()
    for r in range(2):
        assert np.linalg.norm(B[r] - A2 @ E[r]) / np.linalg.norm(B[r]) <= 1.05e-9
        xd = sla.spsolve(A2, B[r])
        assert np.linalg.norm(E[r] - xd) / np.linalg.norm(xd) <= 1e-6
        assert np.linalg.norm(E[r] - E3[r][sel]) / np.linalg.norm(xd) <= 1e-6


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_model2d_window_batch(precision, pkg, O):
    """Windowed 2D operators of different sizes and polarisations (TE: two
    unknowns per cell, TM: one) in one uneven-problem batch, against the
    oracle's residual and the direct solve of each."""
    T = pkg.twod
    om = 0.6
    cases = [("te", (14, 11)), ("tm", (9, 13)), ("te", (6, 7))]
    ws, bs, As = [], [], []
    for q, (kind, N) in enumerate(cases):
        m = model2d(pkg, kind, N=N, om=om)
        As.append(_oracle_2d(pkg, O, m, kind, om))
        w = T.MaxwellOperator2D(m, om)
        assert w.windowed
        ws.append(w)
        rng = np.random.default_rng(60 + q)
        bs.append(rng.normal(size=w.n) + 1j * rng.normal(size=w.n))
    say("model2d window batch", precision)
    es, st = pkg.solve_batch([w.op for w in ws], [w._to_dev(b)[0] for w, b in zip(ws, bs)], tol=1e-9,
                             precision=precision)
    for w, A, b, e in zip(ws, As, bs, es):
        x = w._from_dev(e, b.shape)
        w.close()
        assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1.05e-9
        xd = sla.spsolve(A, b)
        assert np.linalg.norm(x - xd) / np.linalg.norm(xd) <= 1e-6


@pytest.mark.parametrize("kind", ["te", "tm"])
def test_model2d_window_noalloc(kind, pkg, O, devmem):
    """A windowed 2D operator through mfd_solve_dev_noalloc (its workspace
    sized for K unknowns per cell in the fixed scratch array) against
    mfd_solve_dev: bitwise the same field and statistics, to 1e-9 in the
    oracle's residual."""
    T = pkg.twod
    om = 0.6
    m = model2d(pkg, kind, om=om)
    A2 = _oracle_2d(pkg, O, m, kind, om)
    w = T.MaxwellOperator2D(m, om)
    rng = np.random.default_rng(71)
    b = rng.normal(size=w.n) + 1j * rng.normal(size=w.n)
    bd = w._to_dev(b)[0]
    say("model2d window noalloc", kind)
    out = []
    for noalloc in (False, True):
        db, de = devmem.put(bd), devmem.put(np.zeros_like(bd))
        if noalloc:
            st = w.op.solve_dev_noalloc(db, de, 1, tol=1e-9)
        else:
            st = w.op.solve_dev(db, de, 1, tol=1e-9, use_graph=False)
        out.append((devmem.get(de, bd), st))
    w.close()
    (e0, s0), (e1, s1) = out
    assert np.array_equal(e0, e1) and s0.iterations == s1.iterations and s0.relres == s1.relres
    x = w._from_dev(e1, b.shape)
    assert s1.converged and np.linalg.norm(b - A2 @ x) / np.linalg.norm(b) <= 1.05e-9


def test_model2d_window_fallback(pkg, O):
    """μ ≠ 1: the window is not built for it, so the 2D operator falls back
    to the full single-cell operator (three components per cell) and still
    matches the oracle."""
    T = pkg.twod
    om = 0.6
    m = model2d(pkg, "te", om=om, mu=2.0)
    A2 = _oracle_2d(pkg, O, m, "te", om)
    op = T.MaxwellOperator2D(m, om)
    assert not op.windowed and op.n == A2.shape[0]
    X = np.random.default_rng(44).normal(size=op.n) + 0j
    Y = op.apply(X)
    op.close()
    ref = A2 @ X
    assert np.linalg.norm(Y - ref) <= 1e-12 * np.linalg.norm(ref)


def test_component_window_refusals(pkg):
    """A window needs the component-last order; a windowed operator refuses
    the direct kernel and the three-component field conversions."""
    m = model2d(pkg, "te")
    op = pkg.MaxwellOperator(m.m3, 0.6)   # component-first
    with pytest.raises(pkg.MfdArgumentError):
        op.set_components(0, 2)
    op.close()
    w = pkg.twod.MaxwellOperator2D(m, 0.6)
    assert w.windowed
    with pytest.raises(pkg.MfdUnsupportedError):
        w.op.set_tile("fp64", 0, 1, 0)
    with pytest.raises(pkg.MfdUnsupportedError):
        w.op.h_from_e(np.zeros(w.n, dtype=complex))
    w.close()


def test_model2d_tm_is_five_point_laplacian(pkg):
    """Uniform grid, no PML: the TM operator is −Δ₅ − ω² ε with the textbook
    5-point stencil (diagonal 4 − ω²ε, four −1 neighbours in the interior)."""
    T = pkg.twod
    N = (9, 7)
    m = T.ModelTM(T.Grid2([np.arange(n + 1.0) for n in N]))
    T.add_obj(m, "bg", T.Rect((N[0] / 2, N[1] / 2), (float(N[0]), float(N[1]))), ε=2.5)
    T.calc_matparams(m)
    om = 0.7
    A = T.create_A(pkg.EE, om, m).tocsr()
    for i in range(1, N[0] - 1):
        for j in range(1, N[1] - 1):
            p = i + N[0] * j
            row = A.getrow(p)
            assert row[0, p] == pytest.approx(4 - om ** 2 * 2.5, abs=1e-13)
            for q in (p - 1, p + 1, p - N[0], p + N[0]):
                assert row[0, q] == pytest.approx(-1.0, abs=1e-13)
            assert row.nnz == 5


def model1d(pkg, N=40, npml=5, om=0.5):
    D = pkg.oned
    lp = nonuniform((N,), 31)[0]
    m = D.ModelTEM(D.Grid1(lp))
    D.set_ωpml(m, om)
    D.set_Npml(m, (npml, npml))
    D.add_obj(m, "bg", D.Segment(0.5 * (lp[0] + lp[-1]), lp[-1] - lp[0]), ε=1.0)
    D.add_obj(m, "slab", D.Segment(0.5 * (lp[0] + lp[-1]) + 1.3, 6.4), ε=6.0 + 0.1j)
    D.calc_matparams(m)
    return m


def test_model1d_tem_operator_and_solve(pkg, O):
    """1D TEM (tem.jl) as the single-x-y-cell extruded 3D model: the Ex block
    decouples exactly, the 1D operator equals the oracle's extruded operator
    restricted to Ex, and the solve matches the direct solve of the 1D CSC."""
    D = pkg.oned
    om = 0.5
    m = model1d(pkg, om=om)
    s3 = spec_of(m.m3)
    shapes, et, mt = shapes_of(m.m3)
    ea, ma = O.matparams(s3, shapes, et, mt)
    assert np.abs(m.m3.εarr - ea).max() <= 1e-12 * np.abs(ea).max()
    A3 = O.to_scipy(O.create_A(s3, om, ea, ma)).tocsr()
    ex = m.esel()
    other = np.setdiff1d(np.arange(3 * m.grid.M), ex)
    assert abs(A3[ex][:, other]).max() == 0 and abs(A3[other][:, ex]).max() == 0   # exact split
    A1 = D.create_A(pkg.EE, om, m)
    ref = A3[ex][:, ex]
    assert abs(A1 - ref).max() <= 1e-13 * abs(ref).max()
    assert A1.nnz <= 3 * m.grid.M   # tridiagonal
    D.add_srce(m, D.PointSrc1(17.3))
    je, jm = D.create_srcs(m)
    b = D.create_b(pkg.EE, om, m, je, jm)
    op = D.MaxwellOperator1D(m, om)
    say("model1d tem")
    e, st = op.solve(b, tol=1e-10)
    xd = sla.spsolve(A1.tocsc(), b)
    assert np.linalg.norm(e - xd) / np.linalg.norm(xd) <= 1e-7
    assert np.linalg.norm(b - A1 @ e) / np.linalg.norm(b) <= 1.05e-10
    h = op.h_from_e(e)
    assert h.shape == (m.grid.M,)
    op.close()


def test_model1d_tem_is_three_point_laplacian(pkg):
    """Uniform grid, no PML: the TEM operator is −d²/dz² − ω² ε with the
    textbook 3-point stencil (diagonal 2 − ω²ε, two −1 neighbours)."""
    D = pkg.oned
    N = 12
    m = D.ModelTEM(D.Grid1(np.arange(N + 1.0)))
    D.add_obj(m, "bg", D.Segment(N / 2, float(N)), ε=3.5)
    D.calc_matparams(m)
    om = 0.8
    A = D.create_A(pkg.EE, om, m).tocsr()
    for k in range(1, N - 1):
        row = A.getrow(k)
        assert row[0, k] == pytest.approx(2 - om ** 2 * 3.5, abs=1e-13)
        for q in (k - 1, k + 1):
            assert row[0, q] == pytest.approx(-1.0, abs=1e-13)
        assert row.nnz == 3
