"""Pins the oracle's material assignment and Kottke smoothing (MaxwellBase's
assign_param!/smooth_param!, called at full.jl:58-67, are absent) with brute
force and closed forms: the planar-cut volume fraction against sampling, the
Kottke average against its isotropic closed form and tensor invariants, and
smoothed ε of a ball and of a planar interface against direct integration."""
import types

import numpy as np
import pytest


def spec(N, lprim=None):
    lp = lprim if lprim is not None else [np.arange(n + 1.0) for n in N]
    return types.SimpleNamespace(lprim=lp, isbloch=(False,) * 3, boundft=(0, 0, 0), npml=((0,) * 3, (0,) * 3),
                                 omega_pml=1.0, kbloch=(0, 0, 0), pml_m=4.0, pml_lnR=-16.0, order_cmpfirst=True)


def test_volfrac_vs_sampling(O):
    rng = np.random.default_rng(1)
    g = (np.arange(40) + 0.5) / 40
    T = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    for _ in range(40):
        vmin = rng.uniform(-1, 0, 3)
        h = rng.uniform(0.3, 2.0, 3)
        n = rng.normal(size=3)
        if rng.random() < 0.3:
            n[rng.integers(3)] = 0.0
        n /= np.linalg.norm(n)
        r0 = vmin + rng.uniform(0, 1, 3) * h
        f = O.volfrac(vmin, vmin + h, n, r0)
        X = vmin + T * h
        ref = np.mean((X - r0) @ n <= 0)
        assert abs(f - ref) < 0.02
    # exact special cases
    assert O.volfrac([0, 0, 0], [1, 1, 1], [1, 0, 0], [0.3, 0, 0]) == pytest.approx(0.3, abs=1e-15)
    s = np.array([1, 1, 1]) / np.sqrt(3)
    assert O.volfrac([0, 0, 0], [1, 1, 1], s, [0.5, 0.5, 0.5]) == pytest.approx(0.5, abs=1e-14)
    assert O.volfrac([0, 0, 0], [1, 1, 1], s, s / np.sqrt(3) * 1.0 / 1.0 * 1 / np.sqrt(3) * np.sqrt(3)) == \
        pytest.approx(1 / 6, abs=1e-14)   # corner tetrahedron x+y+z <= 1


def test_kottke_isotropic_closed_form(O):
    rng = np.random.default_rng(2)
    for _ in range(20):
        e1, e2 = complex(rng.uniform(1, 12), rng.uniform(0, 1)), complex(rng.uniform(1, 12), 0)
        r = rng.uniform(0, 1)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        out = O.kottke(e1 * np.eye(3), e2 * np.eye(3), r, n)
        epar = r * e1 + (1 - r) * e2
        eperp = 1 / (r / e1 + (1 - r) / e2)
        ref = epar * np.eye(3) + (eperp - epar) * np.outer(n, n)
        np.testing.assert_allclose(out, ref, rtol=1e-13, atol=1e-13)


def test_kottke_tensor_invariants(O):
    rng = np.random.default_rng(3)
    for _ in range(20):
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        E1 = Q @ np.diag(rng.uniform(1, 12, 3)) @ Q.T
        E2 = rng.uniform(1, 12) * np.eye(3)
        r = rng.uniform(0, 1)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        out = O.kottke(E1, E2, r, n)
        assert np.allclose(out, out.T, atol=1e-12)
        ev = np.linalg.eigvalsh(out.real)
        lo = min(np.linalg.eigvalsh(E1).min(), E2[0, 0])
        hi = max(np.linalg.eigvalsh(E1).max(), E2[0, 0])
        assert ev.min() >= lo - 1e-10 and ev.max() <= hi + 1e-10
        # r = 1 and r = 0 return the inputs
        np.testing.assert_allclose(O.kottke(E1, E2, 1.0, n), E1, atol=1e-12)
        np.testing.assert_allclose(O.kottke(E1, E2, 0.0, n), E2, atol=1e-12)


def test_planar_interface_exact(O):
    """A half-space of ε2 cutting x at a fractional plane position: the voxels
    on the interface get the exact harmonic (normal) / arithmetic (tangential)
    averages with the true fill fraction."""
    N = (8, 6, 6)
    s = spec(N)
    xcut = 3.3
    e1, e2 = 2.0, 9.0
    shapes = [("cuboid", (4, 3, 3), (8.0, 6.0, 6.0)), ("cuboid", (xcut + 50, 3, 3), (100.0, 100.0, 100.0))]
    I3 = np.eye(3)
    ea, ma = O.matparams(s, shapes, [e1 * I3, e2 * I3], [I3, I3])
    # εxx at dual x (i+0.5); its voxel spans primal [i, i+1] in x
    for i in range(N[0]):
        lo, hi = i, i + 1
        r = np.clip((hi - xcut) / (hi - lo), 0, 1)
        got = ea[i, 2, 2, 0, 0]
        if 0 < r < 1:
            assert got == pytest.approx(1 / (r / e2 + (1 - r) / e1), rel=1e-13)
        # εyy at primal x = i: voxel spans dual [i-0.5, i+0.5]
        lo2, hi2 = i - 0.5, i + 0.5
        r2 = np.clip((hi2 - xcut) / (hi2 - lo2), 0, 1)
        if 0 < r2 < 1:
            assert ea[i, 2, 2, 1, 1] == pytest.approx(r2 * e2 + (1 - r2) * e1, rel=1e-13)
    assert np.all(ea[..., 0, 1] == 0)   # axis-aligned interface: no off-diagonal
    assert np.all(ma[..., 0, 0] == 1)


def test_ball_smoothing_bruteforce(O):
    """Each smoothed ε_xx of a ball matches Kottke's closed form evaluated with a
    supersampled fill fraction of its voxel and the radial normal; interior
    and exterior voxels keep the assigned values."""
    N = (16, 16, 16)
    s = spec(N)
    R, e_in = 4.3, 12.0
    c = np.array([8.2, 7.9, 8.1])
    shapes = [("cuboid", (8, 8, 8), (16.0, 16.0, 16.0)), ("ball", tuple(c), R)]
    I3 = np.eye(3)
    ea, ma = O.matparams(s, shapes, [I3, e_in * I3], [I3, I3])
    g = (np.arange(12) + 0.5) / 12
    sub = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    checked = 0
    for i in range(N[0]):
        for j in range(N[1]):
            for k in range(N[2]):
                got = ea[i, j, k, 0, 0].real
                x0 = np.array([i + 0.5, j, k])          # E_x location
                vmin = np.array([i, j - 0.5, k - 0.5])   # its voxel: primal x, dual y/z
                X = vmin + sub
                r = np.mean(np.sum((X - c) ** 2, axis=1) <= R * R)
                if got in (1.0, e_in):
                    assert r < 0.35 or r > 0.65            # unsmoothed voxels are (nearly) uniform
                    continue
                if not 0.05 < r < 0.95:
                    continue
                n = (x0 - c) / np.linalg.norm(x0 - c)
                epar = r * e_in + (1 - r)
                eperp = 1 / (r / e_in + (1 - r))
                ref = epar + (eperp - epar) * n[0] ** 2
                assert abs(got - ref) <= 0.06 * (e_in - 1), (i, j, k, got, ref, r)
                checked += 1
    assert checked > 100
    assert ea[8, 8, 8, 0, 0] == e_in and ea[0, 0, 0, 2, 2] == 1.0
    assert np.max(np.abs(ea[..., 0, 1])) > 0.5           # anisotropy at the surface
    np.testing.assert_allclose(ea[..., 0, 1], ea[..., 1, 0], atol=1e-14)


def test_priority_and_mu(O):
    N = (6, 6, 6)
    s = spec(N)
    I3 = np.eye(3)
    shapes = [("cuboid", (3, 3, 3), (6.0, 6.0, 6.0)), ("cuboid", (3, 3, 3), (4.0, 4.0, 4.0)),
              ("cuboid", (3, 3, 3), (2.0, 2.0, 2.0))]
    ea, ma = O.matparams(s, shapes, [1 * I3, 2 * I3, 3 * I3], [I3, 1.5 * I3, I3])
    assert ea[3, 3, 3, 0, 0] == 3.0 or ea[2, 3, 3, 0, 0] == 3.0   # the last-added object wins
    assert ma[0, 0, 0, 0, 0] == 1.0
    assert np.any(ma[..., 0, 0] == 1.5)


def test_cylinder_smoothing_bruteforce(O):
    """A z-cylinder (2D disk extruded): smoothed ε_xx matches Kottke's closed
    form with a supersampled in-plane fill fraction and the radial normal, is
    the same on every z plane, and has no xz/yz coupling."""
    N = (16, 16, 3)
    s = spec(N)
    R, e_in = 4.6, 10.0
    c = np.array([8.3, 7.8, 1.5])
    shapes = [("cuboid", (8, 8, 1.5), (16.0, 16.0, 3.0)), ("cylinder", tuple(c), R, 1e30)]
    I3 = np.eye(3)
    ea, ma = O.matparams(s, shapes, [I3, e_in * I3], [I3, I3])
    for k in range(1, N[2]):
        np.testing.assert_array_equal(ea[..., k, :, :], ea[..., 0, :, :])
    assert np.all(ea[..., 0, 2] == 0) and np.all(ea[..., 1, 2] == 0)
    g = (np.arange(24) + 0.5) / 24
    sub = np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2)
    checked = 0
    for i in range(N[0]):
        for j in range(N[1]):
            got = ea[i, j, 0, 0, 0].real
            x0 = np.array([i + 0.5, j])           # E_x location in the plane
            vmin = np.array([i, j - 0.5])         # its voxel: primal x, dual y
            X = vmin + sub
            r = np.mean(np.sum((X - c[:2]) ** 2, axis=1) <= R * R)
            if got in (1.0, e_in):
                assert r < 0.35 or r > 0.65
                continue
            if not 0.05 < r < 0.95:
                continue
            n = (x0 - c[:2]) / np.linalg.norm(x0 - c[:2])
            epar = r * e_in + (1 - r)
            eperp = 1 / (r / e_in + (1 - r))
            ref = epar + (eperp - epar) * n[0] ** 2
            assert abs(got - ref) <= 0.06 * (e_in - 1), (i, j, got, ref, r)
            checked += 1
    assert checked > 20
    assert ea[8, 8, 1, 2, 2] == e_in and ea[0, 0, 0, 2, 2] == 1.0


def test_ball_offdiagonal_bruteforce(O):
    """Each smoothed off-diagonal ε_vw of a ball, stored at the primal voxel
    corners (the corner grid g₀ of boundft = EE) and averaged over the voxel
    [i−½, i+½]×[j−½, j+½]×[k−½, k+½] whose corners are dual points, matches
    Kottke's isotropic closed form (ε∥ − ε⊥ with ε∥ arithmetic and ε⊥ harmonic
    in the fill fraction) times n_v n_w, with the fill fraction supersampled
    and n the radial normal at the corner.  The tolerance is chosen so that a
    flipped sign, or the normal components of another (v, w) pair, fail."""
    N = (16, 16, 16)
    s = spec(N)
    R, e_in = 4.6, 12.0
    c = np.array([8.3, 7.6, 8.15])
    shapes = [("cuboid", (8, 8, 8), (16.0, 16.0, 16.0)), ("ball", tuple(c), R)]
    I3 = np.eye(3)
    ea, _ = O.matparams(s, shapes, [I3, e_in * I3], [I3, I3])
    g = (np.arange(12) + 0.5) / 12
    sub = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pairs = [(0, 1), (0, 2), (1, 2)]
    tol = 0.025 * (e_in - 1)
    worst = {"right": 0.0, "sign": 0.0, "pair": 0.0}
    checked = 0
    for i in range(1, N[0] - 1):
        for j in range(1, N[1] - 1):
            for k in range(1, N[2] - 1):
                x0 = np.array([i, j, k], float)            # primal corner point
                X = x0 - 0.5 + sub                         # its dual-cornered voxel
                r = np.mean(np.sum((X - c) ** 2, axis=1) <= R * R)
                if not 0.05 < r < 0.95:
                    continue
                n = (x0 - c) / np.linalg.norm(x0 - c)
                epar = r * e_in + (1 - r)
                eperp = 1 / (r / e_in + (1 - r))
                for q, (v, w) in enumerate(pairs):
                    got = ea[i, j, k, v, w].real
                    ref = (eperp - epar) * n[v] * n[w]
                    ov, ow = pairs[(q + 1) % 3]
                    alt = (eperp - epar) * n[ov] * n[ow]   # the normal components of another pair
                    worst["right"] = max(worst["right"], abs(got - ref))
                    worst["sign"] = max(worst["sign"], abs(got + ref))
                    worst["pair"] = max(worst["pair"], abs(got - alt))
                    assert abs(got - ref) <= tol, (i, j, k, v, w, got, ref, r)
                    assert abs(ea[i, j, k, w, v] - ea[i, j, k, v, w]) <= 1e-14
                checked += 1
    assert checked > 100
    # the check discriminates: the plausible mistakes are far outside it
    assert worst["sign"] > 4 * tol and worst["pair"] > 4 * tol, worst
    assert worst["right"] < 0.5 * max(worst["sign"], worst["pair"])
