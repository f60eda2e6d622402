"""Source discretisation: the reference's own known answers (test/source.jl).

distweights (source.jl:208-306) is checked against the golden vectors of
test/source.jl:17-133 in both the oracle and the library's host entry point;
PointSrc (pointsrc.jl:64-103) against the invariants of test/source.jl:184-242.
"""
import json
import os
import types

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "distweights_kat.json")))


def _grid_for(gt):
    g = KAT["grid"]
    return (g["lprim"], g["dlprim"]) if gt == 0 else (g["ldual"], g["dldual"])


def _impls(O, pkg):
    return [("oracle", O.distweights), ("capi", pkg.distweights)]


@pytest.mark.parametrize("case", KAT["cases"], ids=lambda c: f"line{c['line']}")
def test_distweights_kat(case, O, pkg):
    l, dl = _grid_for(case["gt"])
    dom = KAT["grid"]["domain"]
    for name, f in _impls(O, pkg):
        if case["throws"]:
            with pytest.raises(ValueError):
                f(case["c"], case["gt"], dom, l, dl, case["isbloch"])
            continue
        ind, wt = f(case["c"], case["gt"], dom, l, dl, case["isbloch"])
        assert list(ind) == case["ind"], name
        np.testing.assert_allclose(wt, case["wt"], rtol=1e-14, atol=1e-15, err_msg=name)


def test_distweights_n1(O, pkg):
    n1 = KAT["n1"]
    for name, f in _impls(O, pkg):
        for c in n1["c"]:
            ind, wt = f(c, 0, n1["domain"], n1["lprim"], n1["dlprim"], True)
            assert list(ind) == [1, 1]
            assert sum(wt) == pytest.approx(n1["sum_prim"], rel=1e-14)
            ind, wt = f(c, 1, n1["domain"], n1["ldual"], n1["dldual"], True)
            assert list(ind) == [1, 1]
            assert sum(wt) == pytest.approx(n1["sum_dual"], rel=1e-14)
        # N = 1 with a symmetry boundary is rejected (source.jl:230)
        with pytest.raises(ValueError):
            f(0.5, 0, n1["domain"], n1["lprim"], n1["dlprim"], False)


def _spec(lprim, bloch=(True, True, True)):
    return types.SimpleNamespace(lprim=lprim, isbloch=bloch, boundft=(0, 0, 0), npml=((0,) * 3, (0,) * 3),
                                 omega_pml=1.0, kbloch=(0, 0, 0), pml_m=4.0, pml_lnR=-16.0, order_cmpfirst=True)


def _pointsrc_oracle(O, lprim, c, p):
    N = tuple(len(a) - 1 for a in lprim)
    j = np.zeros(N + (3,), dtype=np.complex128, order="F")
    O.add_pointsrc(_spec(lprim), 0, c, p, 1.0, j)
    return j


def _pointsrc_capi(pkg, lprim, c, p):
    g = pkg.Grid(lprim, (True, True, True))
    m = pkg.ModelFull(g)
    pkg.add_srce(m, pkg.PointSrc(c, p, 1.0))
    return m.jₑarr


@pytest.mark.parametrize("impl", ["oracle", "capi"])
def test_pointsrc_invariants(impl, O, pkg):
    """test/source.jl:184-242: 8 nonzeros per component, dipole moment conserved."""
    make = (lambda lp, c, p: _pointsrc_oracle(O, lp, c, p)) if impl == "oracle" else \
        (lambda lp, c, p: _pointsrc_capi(pkg, lp, c, p))
    c = [0.7, 0.7, 0.7]
    p = np.array([1.0, 1.0, 1.0]) / np.sqrt(3.0)
    for step, dv in ((1.0, 1.0), (0.5, 0.125)):
        lp = [np.arange(-10, 10 + step / 2, step)] * 3
        j = make(lp, c, [1, 1, 1])
        for w in range(3):
            assert np.count_nonzero(j[..., w]) == 8
            # the reference asserts exact equality (:201-203, :214-216); the sum of
            # 8 products is order-dependent, so allow 4 ulp
            assert abs(np.sum(j[..., w]) * dv - p[w]) <= 4 * np.spacing(p[w])
    # non-uniform grid (:219-234), seeded
    rng = np.random.default_rng(7)
    lp = []
    for _ in range(3):
        a = np.sort(rng.random(21)) * 20
        a -= a.mean() - 0.7
        lp.append(a)
    j_nu = make(lp, c, [1, 1, 1])
    g = types.SimpleNamespace(**{})
    N = tuple(len(a) - 1 for a in lp)
    ldg = []
    for a in lp:
        ld = 0.5 * (a[:-1] + a[1:])
        ldg.append(np.concatenate(([ld[-1] - (a[-1] - a[0])], ld)))
    dxp, dyp = np.diff(ldg[0]), np.diff(ldg[1])
    dzd = np.diff(lp[2])
    vol = dxp[:, None, None] * dyp[None, :, None] * dzd[None, None, :]
    j1 = make([np.arange(-10, 11.0)] * 3, c, [1, 1, 1])
    assert np.sum(j1[..., 2]) == pytest.approx(np.sum(j_nu[..., 2] * vol), rel=1e-12)


def test_pointsrc_capi_matches_oracle(O, pkg):
    rng = np.random.default_rng(3)
    lp = [np.cumsum(np.concatenate(([0.0], rng.uniform(0.5, 1.5, n)))) for n in (9, 7, 8)]
    for bloch in ((False, False, False), (True, False, True)):
        N = tuple(len(a) - 1 for a in lp)
        for _ in range(5):
            c = [rng.uniform(a[0], a[-1]) for a in lp]
            p = rng.normal(size=3)
            jo = np.zeros(N + (3,), dtype=np.complex128, order="F")
            O.add_pointsrc(_spec(lp, bloch), 0, c, p, 1.5 - 0.5j, jo)
            g = pkg.Grid(lp, bloch)
            m = pkg.ModelFull(g)
            pkg.add_srce(m, pkg.PointSrc(c, p, 1.5 - 0.5j))
            np.testing.assert_allclose(m.jₑarr, jo, rtol=1e-14, atol=1e-15)


def test_c1_dipole_closed_form(O):
    """A z-dipole at (16,16,16) on 0:32 lands on primal x,y and splits 50/50
    between the dual z planes 15.5 and 16.5."""
    lp = [np.arange(33.0)] * 3
    spec = _spec(lp, (False, False, False))
    j = np.zeros((32, 32, 32, 3), dtype=np.complex128, order="F")
    O.add_pointsrc(spec, 0, [16, 16, 16], [0, 0, 1], 1.0, j)
    nz = np.argwhere(j != 0)
    assert sorted(map(tuple, nz)) == [(16, 16, 15, 2), (16, 16, 16, 2)]
    assert j[16, 16, 15, 2] == 0.5 and j[16, 16, 16, 2] == 0.5


def _planesrc_oracle(O, lprim, n, c, p, Jdn=1.0, bloch=(True, True, True)):
    N = tuple(len(a) - 1 for a in lprim)
    j = np.zeros(N + (3,), dtype=np.complex128, order="F")
    O.add_planesrc(_spec(lprim, bloch), 0, n, c, p, Jdn, j)
    return j


def _planesrc_capi(pkg, lprim, n, c, p, Jdn=1.0, bloch=(True, True, True)):
    m = pkg.ModelFull(pkg.Grid(lprim, bloch))
    pkg.add_srce(m, pkg.PlaneSrc(n, c, p, Jdn))
    return m.jₑarr


@pytest.mark.parametrize("impl", ["oracle", "capi"])
def test_planesrc_reference_testset(impl, O, pkg):
    """test/source.jl:123-170: x-polarised source in the z = 0 plane, Bloch grid."""
    make = (lambda lp: _planesrc_oracle(O, lp, [0, 0, 1], 0.0, [1, 0, 0])) if impl == "oracle" else \
        (lambda lp: _planesrc_capi(pkg, lp, [0, 0, 1], 0.0, [1, 0, 0]))
    lp = [np.arange(-10, 11.0)] * 3
    j = make(lp)
    assert np.max(np.abs(j)) == 1.0           # J = K / ∆z (:139)
    assert np.all(j[..., 1] == 0) and np.all(j[..., 2] == 0)   # (:140-141)
    lpf = [np.arange(-10, 10.25, 0.5)] * 3
    jf = make(lpf)
    # total current through an x-normal cross section is resolution independent (:151)
    assert np.sum(j[0, :, :, 0]) * 1.0 == pytest.approx(np.sum(jf[0, :, :, 0]) * 0.25, rel=1e-14)
    assert np.max(np.abs(jf)) == 2.0          # (:152)
    # non-uniform z (:155-168), seeded
    rng = np.random.default_rng(11)
    z = np.sort(rng.random(21)) * 20
    z -= z.mean()
    lpn = [np.arange(-10, 11.0), np.arange(-10, 11.0), z]
    jn = make(lpn)
    # ∆l at primal points = diff of the dual vertices including the Bloch ghost
    def dprim(a):
        ld = 0.5 * (a[:-1] + a[1:])
        return np.diff(np.concatenate(([ld[-1] - (a[-1] - a[0])], ld)))
    area = dprim(lpn[1])[:, None] * dprim(lpn[2])[None, :]
    assert np.sum(j[0, :, :, 0]) == pytest.approx(np.sum(jn[0, :, :, 0] * area), rel=1e-12)


def test_planesrc_capi_matches_oracle(O, pkg):
    rng = np.random.default_rng(5)
    lp = [np.cumsum(np.concatenate(([0.0], rng.uniform(0.5, 1.5, n)))) for n in (9, 7, 8)]
    N = tuple(len(a) - 1 for a in lp)
    for bloch in ((False, False, False), (True, False, True)):
        for nn in range(3):
            n = [0, 0, 0]
            n[nn] = 2.5
            for _ in range(3):
                c = rng.uniform(lp[nn][0], lp[nn][-1])
                p = rng.normal(size=3)
                jo = _planesrc_oracle(O, lp, n, c, p, 0.5 + 2j, bloch)
                jc = _planesrc_capi(pkg, lp, n, c, p, 0.5 + 2j, bloch)
                np.testing.assert_allclose(jc, jo, rtol=1e-14, atol=1e-15)
                # each component lands on at most two planes normal to nn
                for w in range(3):
                    planes = np.unique(np.argwhere(jo[..., w] != 0)[:, nn])
                    assert len(planes) <= 2


def test_planesrc_rejects(pkg):
    with pytest.raises(pkg.MfdArgumentError):
        pkg.PlaneSrc([1, 1, 0], 0.0, [0, 0, 1])
    m = pkg.ModelFull(pkg.Grid([np.arange(5.0)] * 3, (False, False, False)))
    with pytest.raises(pkg.MfdArgumentError):
        pkg.add_srce(m, pkg.PlaneSrc([0, 0, -1], 1.0, [1, 0, 0]))   # no +1 entry (planesrc.jl:42)
    with pytest.raises(pkg.MfdArgumentError):
        pkg.add_srce(m, pkg.PlaneSrc([0, 0, 1], 9.0, [1, 0, 0]))    # outside the domain (source.jl:214)
