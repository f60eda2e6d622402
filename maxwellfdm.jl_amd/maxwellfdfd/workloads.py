"""Seeded synthetic problems of the published benchmark configurations.

Every geometry draw uses numpy's PCG64 with an explicit seed, so the CPU
checker and the device path see identical inputs.
"""
import numpy as np

from . import core as M


def uniform_grid(N):
    N = (N, N, N) if np.isscalar(N) else tuple(N)
    return M.Grid([np.arange(n + 1.0) for n in N])


def base_model(N, npml, omega):
    g = uniform_grid(N)
    m = M.ModelFull(g)
    M.set_ωpml(m, omega)
    M.set_Npml(m, ((npml,) * 3, (npml,) * 3))
    Ns = g.N
    M.add_obj(m, "vacuum", M.Cuboid(tuple(n / 2 for n in Ns), tuple(float(n) for n in Ns)), ε=1.0, μ=1.0)
    return m


def sphere_dipole(N=32, npml=8, r=6.0, eps=12.0, lam=20.0):
    """Plain-method small case: centred sphere ε=12 (r=6), z-dipole at the centre, ω=2π/20."""
    om = 2 * np.pi / lam
    m = base_model(N, npml, om)
    c = (N / 2,) * 3
    M.add_obj(m, "dielectric", M.Ball(c, r), ε=eps)
    M.add_srce(m, M.PointSrc(c, (0, 0, 1), 1.0))
    return m, om


def random_dipoles(N, npml, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        pos = rng.uniform(npml + 2, N - npml - 2, size=3)
        out.append((pos, v))
    return out


def random_boxes(N=128, npml=10, nbox=40, lam=24.0, seed=1, ndip=1):
    """Mixed-precision case: background ε=1 with nbox axis-aligned boxes, corners
    and sizes uniform inside the non-PML interior, ε ~ U[1,12]; dipoles of
    uniform random orientation at seeded interior positions."""
    om = 2 * np.pi / lam
    m = base_model(N, npml, om)
    rng = np.random.default_rng(seed)
    lo_i, hi_i = float(npml), float(N - npml)
    for b in range(nbox):
        span = hi_i - lo_i
        s = rng.uniform(min(4.0, span / 6), span / 3, size=3)
        lo = rng.uniform(lo_i, hi_i - s)
        eps = rng.uniform(1.0, 12.0)
        M.add_obj(m, f"box{b}", M.Cuboid(tuple(lo + s / 2), tuple(s)), ε=eps)
    for pos, v in random_dipoles(N, npml, ndip, seed + 1000):
        M.add_srce(m, M.PointSrc(pos, v, 1.0))
    return m, om


def dipole_rhs(mdl, omega, dipoles):
    """One b vector per dipole (nrhs x n)."""
    bs = []
    for pos, v in dipoles:
        M.clear_srcs(mdl)
        M.add_srce(mdl, M.PointSrc(pos, v, 1.0))
        je, jm = M.create_srcs(mdl)
        bs.append(M.create_b(M.EE, omega, mdl, je, jm))
    return np.array(bs)


def uneven_problems(count=32, nmin=48, nmax=128, npml=10, nbox=40, seed=7):
    """Batch of independent problems: cubic grids N ~ U{nmin..nmax}, each with
    its own seeded random-box ε (as `random_boxes`), one random dipole and
    ω ~ U[2π/30, 2π/16].  Returns [(model, ω)]."""
    rng = np.random.default_rng(seed)
    out = []
    for p in range(count):
        N = int(rng.integers(nmin, nmax + 1))
        om = float(rng.uniform(2 * np.pi / 30, 2 * np.pi / 16))
        m, _ = random_boxes(N=N, npml=npml, nbox=nbox, lam=2 * np.pi / om, seed=seed * 1000 + p, ndip=1)
        out.append((m, om))
    return out


def adjoint_design(N=(128, 128, 64), npml=10, design=(64, 64, 16), lam=24.0, seed=3, monitor=(4, 4, 1)):
    """Adjoint case: vacuum with a centred design region whose ε per voxel is
    1 + 11·sigmoid(G∗ξ), ξ ~ N(0,1) seeded and G a Gaussian of σ = 3 cells
    (applied to all three diagonal components); a z-dipole below the region and
    a monitor patch of E_z DOFs above it.  Returns (model with εarr assigned,
    ω, b, monitor DOF indices, design-region cell ranges)."""
    from scipy.ndimage import gaussian_filter
    om = 2 * np.pi / lam
    m = base_model(N, npml, om)
    M.calc_matparams(m)
    rng = np.random.default_rng(seed)
    lo = [(n - d) // 2 for n, d in zip(N, design)]
    xi = gaussian_filter(rng.normal(size=design), sigma=3.0, mode="nearest")
    xi /= xi.std()
    eps = 1.0 + 11.0 / (1.0 + np.exp(-xi))
    sl = tuple(slice(l, l + d) for l, d in zip(lo, design))
    for c in range(3):
        m.εarr[sl + (c, c)] = eps
    zsrc = max(npml + 2, lo[2] - 4)
    M.add_srce(m, M.PointSrc((N[0] / 2, N[1] / 2, zsrc), (0, 0, 1), 1.0))
    je, jm = M.create_srcs(m)
    b = M.create_b(M.EE, om, m, je, jm, μarr=m.μarr)
    zmon = min(N[2] - npml - 2, lo[2] + design[2] + 4)
    i0, j0 = N[0] // 2 - monitor[0] // 2, N[1] // 2 - monitor[1] // 2
    mon = [2 + 3 * (i + N[0] * (j + N[1] * k)) for k in range(zmon, zmon + monitor[2])
           for j in range(j0, j0 + monitor[1]) for i in range(i0, i0 + monitor[0])]
    return m, om, b, np.array(mon), sl


def random_spheres(N=(384, 384, 256), npml=12, nsph=200, rmin=4.0, rmax=20.0, emin=2.0, emax=12.0, lam=32.0,
                   seed=5, ndip=8):
    """Large anisotropic case: background ε=1 with nsph seeded spheres
    (radius ~ U[rmin, rmax], ε ~ U[emin, emax], centres inside the non-PML
    interior); subpixel smoothing then yields full 3×3 ε tensors at the
    sphere surfaces.  Returns (model, ω, dipoles)."""
    om = 2 * np.pi / lam
    m = base_model(N, npml, om)
    rng = np.random.default_rng(seed)
    for k in range(nsph):
        r = rng.uniform(rmin, rmax)
        c = tuple(rng.uniform(npml + r, n - npml - r) for n in N)
        M.add_obj(m, f"sphere{k}", M.Ball(c, r), ε=rng.uniform(emin, emax))
    dips = []
    for _ in range(ndip):
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        dips.append((np.array([rng.uniform(npml + 2, n - npml - 2) for n in N]), v))
    return m, om, dips
