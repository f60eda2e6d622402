"""Pins the oracle's curl, PML, material-operator and A restatement with
identities the mathematics fixes (the reference's MaxwellBase, which holds
this arithmetic, is absent): div∘curl = 0, Cₘ = S_E⁻¹ Cₑᵀ S_H, the discrete
plane-wave eigenvalue, the diagonal-ε operator, S-symmetry, A as the product
of its factors, the SC-PML closed form and its absorption."""
import types

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla


def spec(N, npml=0, om=1.0, bloch=(False,) * 3, kb=(0.0, 0.0, 0.0), lprim=None, cmpfirst=True, bft=(0, 0, 0)):
    N = (N,) * 3 if np.isscalar(N) else tuple(N)
    lp = lprim if lprim is not None else [np.arange(n + 1.0) for n in N]
    npml = ((npml,) * 3, (npml,) * 3) if np.isscalar(npml) else npml
    return types.SimpleNamespace(lprim=lp, isbloch=bloch, boundft=bft, npml=npml, omega_pml=om, kbloch=kb,
                                 pml_m=4.0, pml_lnR=-16.0, order_cmpfirst=cmpfirst)


def nonuniform(N, seed):
    rng = np.random.default_rng(seed)
    return [np.cumsum(np.concatenate(([0.0], rng.uniform(0.6, 1.4, n)))) for n in N]


def tables(O, s):
    return O.grid_tables(s)


def axis_op(n, inv, fwd, bloch, ph):
    """1D difference operator with the ghost rule (Bloch phase or zero)."""
    D = sp.lil_matrix((n, n), dtype=complex)
    for i in range(n):
        if fwd:
            D[i, i] += -inv[i]
            if i < n - 1:
                D[i, i + 1] += inv[i]
            elif bloch:
                D[i, 0] += inv[i] * ph
        else:
            D[i, i] += inv[i]
            if i > 0:
                D[i, i - 1] += -inv[i]
            elif bloch:
                D[i, n - 1] += -inv[i] / ph
    return D.tocsr()


def kron3(ops, N):
    """operator acting along each axis on a column-major (Nx,Ny,Nz) scalar field"""
    Ix, Iy, Iz = (sp.identity(n) for n in N)
    return [sp.kron(Iz, sp.kron(Iy, ops[0])), sp.kron(Iz, sp.kron(ops[1], Ix)), sp.kron(ops[2], sp.kron(Iy, Ix))]


def comp(M, c, N):
    """select component c of a component-first DOF vector"""
    n = int(np.prod(N))
    return sp.csr_matrix((np.ones(n), (np.arange(n), 3 * np.arange(n) + c)), shape=(n, 3 * n))


@pytest.mark.parametrize("case", ["uniform_pml", "nonuniform_bloch", "mixed"])
def test_div_curl_zero(case, O):
    if case == "uniform_pml":
        s = spec(6, npml=2, om=0.7)
    elif case == "nonuniform_bloch":
        N = (5, 4, 6)
        s = spec(N, bloch=(True,) * 3, kb=(0.3, -0.2, 0.5), lprim=nonuniform(N, 1))
    else:
        N = (4, 5, 3)
        s = spec(N, npml=((1, 0, 1), (1, 1, 0)), om=0.9, bloch=(True, False, True), kb=(0.1, 0, 0.4),
                 lprim=nonuniform(N, 2))
    N = tuple(len(a) - 1 for a in s.lprim)
    T = tables(O, s)
    Ce = O.to_scipy(O.curl(s, 0))
    Cm = O.to_scipy(O.curl(s, 1))
    L = [a[-1] - a[0] for a in s.lprim]
    ph = [np.exp(-1j * s.kbloch[w] * L[w]) for w in range(3)]
    # forward divergence of H (dual planes) and backward divergence of E (primal planes)
    dfw = kron3([axis_op(N[w], 1 / T[("sdl", 1, w)], True, s.isbloch[w], ph[w]) for w in range(3)], N)
    dbw = kron3([axis_op(N[w], 1 / T[("sdl", 0, w)], False, s.isbloch[w], ph[w]) for w in range(3)], N)
    divH = sum(dfw[c] @ comp(None, c, N) for c in range(3))
    divE = sum(dbw[c] @ comp(None, c, N) for c in range(3))
    assert abs(divH @ Ce).max() <= 1e-12 * abs(Ce).max() ** 2
    assert abs(divE @ Cm).max() <= 1e-12 * abs(Cm).max() ** 2
    # the curls themselves are the Yee combinations of the same 1D operators
    for c in range(3):
        v, w = (c + 1) % 3, (c + 2) % 3
        ref = dfw[v] @ comp(None, w, N) - dfw[w] @ comp(None, v, N)
        assert abs(comp(None, c, N) @ Ce - ref).max() <= 1e-14 * abs(Ce).max()
        ref = dbw[v] @ comp(None, w, N) - dbw[w] @ comp(None, v, N)
        assert abs(comp(None, c, N) @ Cm - ref).max() <= 1e-14 * abs(Cm).max()


def S_of(T, N, ft):
    """stretched cell volume at the E (ft=0) or H (ft=1) DOFs, component-first"""
    out = np.zeros(3 * int(np.prod(N)), dtype=complex)
    I, J, K = np.meshgrid(*[np.arange(n) for n in N], indexing="ij")
    for c in range(3):
        v = np.ones(N, dtype=complex)
        for a, idx in enumerate((I, J, K)):
            g0 = 0 if ft == 0 else 1
            gt = 1 - g0 if a == c else g0
            v = v * T[("sdl", gt, a)][idx]
        out[c::3] = v.reshape(-1, order="F")
    return out


@pytest.mark.parametrize("bloch", [False, True])
def test_cm_is_scaled_transpose(bloch, O):
    N = (5, 6, 4)
    s = spec(N, npml=((2, 1, 1), (1, 2, 1)), om=0.8, bloch=(bloch,) * 3, lprim=nonuniform(N, 4))
    T = tables(O, s)
    Ce = O.to_scipy(O.curl(s, 0))
    Cm = O.to_scipy(O.curl(s, 1))
    SE, SH = S_of(T, N, 0), S_of(T, N, 1)
    ref = sp.diags(1 / SE) @ Ce.T @ sp.diags(SH)
    assert abs(Cm - ref).max() <= 1e-13 * abs(Cm).max()


def test_plane_wave_eigenvalue(O):
    """Periodic vacuum: CₘCₑ restricted to the plane waves of wavevector k has
    eigenvalues {0, |k̃|², |k̃|²}, k̃_w = (2/∆) sin(k_w ∆/2)."""
    N = (8, 6, 10)
    dx = (1.0, 0.5, 2.0)
    lp = [np.arange(N[w] + 1) * dx[w] for w in range(3)]
    m = (1, 2, 3)
    k = [2 * np.pi * m[w] / (N[w] * dx[w]) + 0.0 for w in range(3)]
    kb = [0.05, -0.03, 0.02]
    k = [k[w] + kb[w] for w in range(3)]
    s = spec(N, bloch=(True,) * 3, kb=kb, lprim=lp)
    Ce = O.to_scipy(O.curl(s, 0))
    Cm = O.to_scipy(O.curl(s, 1))
    CC = Cm @ Ce
    T = tables(O, s)
    I, J, K = np.meshgrid(*[np.arange(n) for n in N], indexing="ij")
    basis = []
    for c in range(3):
        pos = [T[("l", 1 if a == c else 0, a)] for a in range(3)]
        phase = np.exp(-1j * (k[0] * pos[0][I] + k[1] * pos[1][J] + k[2] * pos[2][K]))
        v = np.zeros(3 * I.size, dtype=complex)
        v[c::3] = phase.reshape(-1, order="F")
        basis.append(v)
    B = np.array(basis).T
    Kmat = np.linalg.lstsq(B, CC @ B, rcond=None)[0]
    assert np.linalg.norm(CC @ B - B @ Kmat) <= 1e-10 * np.linalg.norm(CC @ B)
    ev = np.sort(np.linalg.eigvals(Kmat).real)
    kt2 = sum((2 / dx[w] * np.sin(k[w] * dx[w] / 2)) ** 2 for w in range(3))
    np.testing.assert_allclose(ev, [0, kt2, kt2], atol=1e-10 * kt2)


def test_pml_closed_form(O):
    N, npml, om = 20, 5, 0.6
    s = spec((N, 3, 3), npml=((npml, 0, 0), (npml, 0, 0)), om=om)
    T = tables(O, s)
    for gt in (0, 1):
        u = T[("l", gt, 0)]
        sdl = T[("sdl", gt, 0)]
        dl = T[("dl", gt, 0)]
        sig = np.zeros(N)
        Lp = float(npml)
        smax = -(4 + 1) * (-16.0) / (2 * Lp)
        lo, hi = npml, N - npml
        sig = np.where(u < lo, smax * ((lo - u) / Lp) ** 4, sig)
        sig = np.where(u > hi, smax * ((u - hi) / Lp) ** 4, sig)
        np.testing.assert_allclose(sdl, (1 + sig / (1j * om)) * dl, rtol=1e-14)
        inner = (u >= lo) & (u <= hi)
        assert np.all(sdl[inner] == dl[inner])
    # other axes have no PML: s = 1 exactly
    assert np.all(T[("sdl", 0, 1)] == T[("dl", 0, 1)])


def test_pml_absorbs(O):
    """Quasi-1D guided plane wave (periodic, one cell in y and z) launched by
    a sheet of current: with the PML the field decays inside it and the
    standing-wave ratio in the interior stays below 1 + 1e-3."""
    N, npml, lam = 120, 15, 20.0
    om = 2 * np.pi / lam
    s = spec((N, 1, 1), npml=((npml, 0, 0), (npml, 0, 0)), om=om, bloch=(False, True, True))
    eps = np.zeros((N, 1, 1, 3, 3), dtype=complex, order="F")
    mu = np.zeros_like(eps)
    for c in range(3):
        eps[..., c, c] = 1
        mu[..., c, c] = 1
    A = O.to_scipy(O.create_A(s, om, eps, mu))
    je = np.zeros(3 * N, dtype=complex)
    je[3 * (N // 2) + 1] = 1.0          # E_y sheet at x = N/2
    b = O.create_b(s, om, mu, je, np.zeros_like(je))
    e = sla.spsolve(A.tocsc(), b)
    ey = np.abs(e[1::3])
    interior = ey[npml + 5: N // 2 - 5]
    assert interior.max() / interior.min() < 1 + 1e-3
    # inside the PML |E| follows exp(-∫σ dx) (normal incidence, k/ω = 1)
    smax = 5 * 16.0 / (2 * npml)
    for x in (2, 6, 10):
        att = np.exp(-smax * npml / 5 * ((npml - x) / npml) ** 5)
        assert ey[x] / interior.max() == pytest.approx(att, rel=0.1)


def test_paramop_diagonal_and_symmetry(O):
    N = (4, 5, 3)
    s = spec(N, npml=((1, 1, 1), (1, 1, 0)), om=0.5, lprim=nonuniform(N, 9))
    rng = np.random.default_rng(0)
    M = int(np.prod(N))
    eps = np.zeros(N + (3, 3), dtype=complex, order="F")
    for c in range(3):
        eps[..., c, c] = 1 + rng.random(N) + 0.1j * rng.random(N)
    P = O.to_scipy(O.paramop(s, 0, eps))
    dvec = np.zeros(3 * M, dtype=complex)
    for c in range(3):
        dvec[c::3] = eps[..., c, c].reshape(-1, order="F")
    assert P.nnz == 3 * M
    assert abs(P - sp.diags(dvec)).max() == 0
    # anisotropic: S_E Pε is symmetric when ε_vw = ε_wv at the corners
    for (v, w) in ((0, 1), (0, 2), (1, 2)):
        x = rng.normal(size=N) + 1j * rng.normal(size=N)
        eps[..., v, w] = x
        eps[..., w, v] = x
    P = O.to_scipy(O.paramop(s, 0, eps))
    assert P.nnz > 3 * M
    SE = S_of(tables(O, s), N, 0)
    SP = sp.diags(SE) @ P
    assert abs(SP - SP.T).max() <= 1e-13 * abs(SP).max()


@pytest.mark.parametrize("full", [False, True])
def test_A_is_product_of_factors(full, O):
    N = (4, 4, 5)
    s = spec(N, npml=1, om=0.45, lprim=nonuniform(N, 11))
    rng = np.random.default_rng(5)
    eps = np.zeros(N + (3, 3), dtype=complex, order="F")
    mu = np.zeros_like(eps)
    for c in range(3):
        eps[..., c, c] = 1 + 11 * rng.random(N)
        mu[..., c, c] = 1 + 0.5 * rng.random(N)
    if full:
        for (v, w) in ((0, 1), (1, 2)):
            x = rng.normal(size=N)
            eps[..., v, w] = x
            eps[..., w, v] = x
    om = 0.45
    A = O.to_scipy(O.create_A(s, om, eps, mu))
    Ce = O.to_scipy(O.curl(s, 0))
    Cm = O.to_scipy(O.curl(s, 1))
    Pe = O.to_scipy(O.paramop(s, 0, eps))
    Pm = O.to_scipy(O.paramop(s, 1, mu))
    ref = Cm @ sp.diags(1 / Pm.diagonal()) @ Ce - om ** 2 * Pe
    assert abs(A - ref).max() <= 1e-13 * abs(A).max()
    # Aᵀ = S_E A S_E⁻¹ (complex symmetry used by the solver and the adjoint)
    SE = S_of(tables(O, s), N, 0)
    assert abs(A.T - sp.diags(SE) @ A @ sp.diags(1 / SE)).max() <= 1e-12 * abs(A).max()
    # b = −Cₘ μ⁻¹ jₘ − iω jₑ
    je = rng.normal(size=A.shape[0]) + 1j * rng.normal(size=A.shape[0])
    jm = rng.normal(size=A.shape[0]) + 1j * rng.normal(size=A.shape[0])
    b = O.create_b(s, om, mu, je, jm)
    np.testing.assert_allclose(b, -(Cm @ (jm / Pm.diagonal())) - 1j * om * je, rtol=1e-12, atol=1e-12)


def test_A_hh_is_product_of_factors(O):
    """HH formulation (model.jl:238-240, 266-268): A = Cₑ Pε⁻¹ Cₘ − ω² Pμ with
    Aᵀ = S_H A S_H⁻¹, and b = Cₑ Pε⁻¹ jₑ − iω jₘ."""
    N = (4, 5, 4)
    s = spec(N, npml=1, om=0.5, lprim=nonuniform(N, 13))
    rng = np.random.default_rng(8)
    eps = np.zeros(N + (3, 3), dtype=complex, order="F")
    mu = np.zeros_like(eps)
    for c in range(3):
        eps[..., c, c] = 1 + 11 * rng.random(N) + 0.2j * rng.random(N)
        mu[..., c, c] = 1 + 0.5 * rng.random(N)
    om = 0.45 + 0.01j
    A = O.to_scipy(O.create_A(s, om, eps, mu, ft=1))
    Ce = O.to_scipy(O.curl(s, 0))
    Cm = O.to_scipy(O.curl(s, 1))
    Pe = O.to_scipy(O.paramop(s, 0, eps))
    Pm = O.to_scipy(O.paramop(s, 1, mu))
    ref = Ce @ sp.diags(1 / Pe.diagonal()) @ Cm - om ** 2 * Pm
    assert abs(A - ref).max() <= 1e-13 * abs(A).max()
    SH = S_of(tables(O, s), N, 1)
    assert abs(A.T - sp.diags(SH) @ A @ sp.diags(1 / SH)).max() <= 1e-12 * abs(A).max()
    je = rng.normal(size=A.shape[0]) + 1j * rng.normal(size=A.shape[0])
    jm = rng.normal(size=A.shape[0]) + 1j * rng.normal(size=A.shape[0])
    b = O.create_b(s, om, mu, je, jm, ft=1, eps_arr=eps)
    np.testing.assert_allclose(b, Ce @ (je / Pe.diagonal()) - 1j * om * jm, rtol=1e-12, atol=1e-12)
    # the E and H solutions of one source pair are linked by h = (i/ω) μ⁻¹ (Cₑ e + jₘ)
    bE = O.create_b(s, om, mu, je, jm)
    AE = O.to_scipy(O.create_A(s, om, eps, mu))
    import scipy.sparse.linalg as spla
    e = spla.spsolve(AE.tocsc(), bE)
    h = spla.spsolve(A.tocsc(), b)
    h_from_e = (1j / om) * ((Ce @ e + jm) / Pm.diagonal())
    assert np.linalg.norm(h - h_from_e) <= 1e-9 * np.linalg.norm(h)


def test_A_omega_zero_keeps_structure(O):
    s = spec(4, npml=0)
    eps = np.zeros((4, 4, 4, 3, 3), dtype=complex, order="F")
    mu = np.zeros_like(eps)
    for c in range(3):
        eps[..., c, c] = 2.0
        mu[..., c, c] = 1.0
    A0 = O.to_scipy(O.create_A(s, 0.0, eps, mu))
    Ce = O.to_scipy(O.curl(s, 0))
    Cm = O.to_scipy(O.curl(s, 1))
    assert abs(A0 - Cm @ Ce).max() <= 1e-14


# ---------------------------------------------------------------- create_Mcs

def mcs_block(Mc, w, N):
    """the w-component block of a component-first corner-mean operator"""
    S = comp(Mc, w, N)
    return (S @ Mc @ S.T).tocsr()


@pytest.mark.parametrize("which", [0, 1])
def test_Mcs_uniform_weights(which, O):
    """Uniform real grid: every corner is the plain mean (1/2, 1/2) of the two
    field points beside it along w; the boundary corner keeps 1/2 of its own
    point (the ghost drops)."""
    N = (5, 4, 6)
    s = spec(N)
    Mc = O.to_scipy(O.Mcs(s, which)).tocsr()
    for r in range(Mc.shape[0]):
        row = Mc.getrow(r)
        assert set(np.round(row.data.real, 15)) <= {0.5} and np.all(row.data.imag == 0)
        assert row.nnz in (1, 2)
    # one boundary corner per line along w
    n1 = sum(1 for r in range(Mc.shape[0]) if Mc.getrow(r).nnz == 1)
    assert n1 == sum(int(np.prod(N)) // N[w] for w in range(3))


@pytest.mark.parametrize("which,bft", [(0, (0, 0, 0)), (1, (1, 1, 1))])
def test_Mcs_partition_of_unity(which, bft, O):
    """Nonuniform real grid, fields at the cell centres along w and corners
    at the primal planes (Mcₑ with boundft EE, Mcₘ with boundft HH): the ∆l
    weights of each corner away from the boundary (whose ghost cells take
    mirrored spacings) sum to 1 — s∆l of the two field cells over twice the
    corner cell — which a swapped grid type or a wrong index breaks."""
    N = (6, 5, 7)
    s = spec(N, lprim=nonuniform(N, 3), bft=bft)
    Mc = O.to_scipy(O.Mcs(s, which)).tocsr()
    two = [r for r in range(Mc.shape[0]) if Mc.getrow(r).nnz == 2]
    assert len(two) == sum(int(np.prod(N)) // N[w] * (N[w] - 1) for w in range(3))
    n = int(np.prod(N))
    q = np.array(np.unravel_index(np.arange(n), N, order="F"))
    inner = [r for r in two if 1 <= q[r % 3][r // 3] <= N[r % 3] - 2]
    assert len(inner) > len(two) // 2
    sums = np.array([Mc.getrow(r).sum() for r in inner])
    assert np.abs(sums - 1).max() <= 1e-14
    w = np.array([Mc.getrow(r).data.real for r in inner])
    assert np.all(w > 0.2) and np.all(w < 0.8) and not np.allclose(w, 0.5)


@pytest.mark.parametrize("which", [0, 1])
def test_Mcs_bloch_eigenrelation(which, O):
    """Uniform Bloch grid: a sampled Bloch wave along w, F_i = e^{-i k i}
    (F(N) = e^{-ikL} F(0), the phase convention of the curls), maps to itself
    times (1 + e^{+ik})/2 at every corner, the wrapped one included."""
    N = (7, 5, 6)
    kb = (0.37, -0.52, 0.81)
    s = spec(N, bloch=(True,) * 3, kb=kb)
    Mc = O.to_scipy(O.Mcs(s, which)).tocsr()
    n = int(np.prod(N))
    q = np.array(np.unravel_index(np.arange(n), N, order="F"))
    for w in range(3):
        F = np.zeros(3 * n, dtype=complex)
        F[w::3] = np.exp(-1j * kb[w] * q[w])
        G = Mc @ F
        bft_fwd = False   # boundft EE: Mcₑ averages backward, Mcₘ forward
        fwd = (which == 1) != bft_fwd
        lam = 0.5 * (1 + np.exp(-1j * kb[w])) if fwd else 0.5 * (1 + np.exp(1j * kb[w]))
        assert np.abs(G - lam * F).max() <= 1e-14


@pytest.mark.parametrize("which,bft", [(0, (0, 0, 0)), (0, (1, 0, 1)), (1, (0, 0, 0)), (1, (0, 1, 1))])
def test_Mcs_is_inner_average_of_paramop(which, bft, O):
    """Pε's (Pμ's) off-diagonal block (v, w) is the corner mean of F_w (the w
    block of Mcₑ, Mcₘ), scaled by the tensor entry at the corners, then
    averaged half-half back to F_v along the transposed pattern of the v block
    — on a nonuniform PML grid, with either boundary field type."""
    N = (5, 4, 6)
    s = spec(N, npml=((1, 0, 1), (1, 1, 0)), om=0.8, lprim=nonuniform(N, 5), bft=bft)
    n = int(np.prod(N))
    Mc = O.to_scipy(O.Mcs(s, which)).tocsr()
    rng = np.random.default_rng(9)
    for (v, w) in ((0, 1), (2, 0), (1, 2)):
        eps = np.zeros(N + (3, 3), dtype=complex, order="F")
        for c in range(3):
            eps[..., c, c] = 1.0
        x = rng.normal(size=N) + 0.2j * rng.normal(size=N)
        eps[..., v, w] = x
        eps[..., w, v] = x
        P = O.to_scipy(O.paramop(s, which, eps)).tocsr()
        Pvw = comp(P, v, N) @ P @ comp(P, w, N).T
        Mw = mcs_block(Mc, w, N)
        Hv = mcs_block(Mc, v, N).T.copy()
        Hv.data[:] = 0.5
        ref = Hv @ sp.diags(x.reshape(-1, order="F")) @ Mw
        assert abs(Pvw - ref).max() <= 1e-14 * abs(ref).max()


@pytest.mark.parametrize("bft", [(0, 0, 0), (1, 0, 1)])
@pytest.mark.parametrize("bloch", [(False,) * 3, (True, False, True)])
def test_A_hh_is_ee_of_flipped_model(bft, bloch, O):
    """E↔H duality: the HH operator Cₑ (Pε \\ Cₘ) − ω² Pμ of a model equals,
    entry for entry, the EE operator of the same grid with every boundft
    flipped and ε, μ swapped."""
    N = (5, 4, 6)
    kw = dict(npml=((1, 0, 1), (1, 1, 0)), om=0.8, lprim=nonuniform(N, 5), bloch=bloch, kb=(0.2, 0.0, -0.3))
    s = spec(N, bft=bft, **kw)
    sf = spec(N, bft=tuple(1 - b for b in bft), **kw)
    rng = np.random.default_rng(1)
    eps = np.zeros(N + (3, 3), dtype=complex, order="F")
    mu = np.zeros_like(eps)
    for c in range(3):
        eps[..., c, c] = 1 + 5 * rng.random(N) + 0.1j
        mu[..., c, c] = 1 + rng.random(N)
    Ahh = O.to_scipy(O.create_A(s, 0.7, eps, mu, ft=1))
    Aee = O.to_scipy(O.create_A(sf, 0.7, mu, eps, ft=0))
    assert (Ahh != Aee).nnz == 0 and Ahh.nnz == Aee.nnz
