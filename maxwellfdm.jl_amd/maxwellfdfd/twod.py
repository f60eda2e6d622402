"""2D models: ModelTE (E = (Ex, Ey), H = Hz) and ModelTM (E = Ez, H = (Hx, Hy))
of MaxwellFDFD (src/model/te.jl, src/model/tm.jl).

A 2D model is the 3D model of a single z cell with a Bloch (k = 0) z boundary
and shapes extruded along z: every z difference then vanishes, the 3D system
splits into the TE block (Ex, Ey, Hz) and the TM block (Ez, Hx, Hy), and the
3D device path (assembly, subpixel smoothing, matrix-free solve) serves both.
2D DOF vectors are component-first, c + K·(i + Nx·j) with K = 2 for the
two-component field and K = 1 for the other; 2D parameter arrays are
(Nx, Ny, K, K) for the field's components (te.jl:4-13, tm.jl:4-13).
"""
from dataclasses import dataclass
from typing import Tuple

import numpy as np

from . import core as M
from ._lib import MfdArgumentError, MfdUnsupportedError
from . import _lib as L


class Grid2:
    """2D staggered grid from the primal x and y plane locations."""

    def __init__(self, lprim, isbloch=(False, False)):
        if len(lprim) != 2 or len(isbloch) != 2:
            raise ValueError("Grid2 needs 2 axes")
        self.grid3 = M.Grid([lprim[0], lprim[1], np.array([0.0, 1.0])], (isbloch[0], isbloch[1], True))
        self.N = self.grid3.N[:2]
        self.lprim = self.grid3.lprim[:2]
        self.isbloch = tuple(bool(b) for b in isbloch)

    @property
    def M(self):
        return self.N[0] * self.N[1]


@dataclass
class Disk:
    """2D disk (GeometryPrimitives Ball in 2D): centre c, radius r."""
    c: Tuple[float, float]
    r: float


@dataclass
class Rect:
    """2D axis-aligned rectangle (GeometryPrimitives Cuboid in 2D): centre c, sides s."""
    c: Tuple[float, float]
    s: Tuple[float, float]


def _extrude(shape):
    if isinstance(shape, Disk):
        return M.Cylinder((shape.c[0], shape.c[1], 0.5), shape.r, 1e30)
    if isinstance(shape, Rect):
        return M.Cuboid((shape.c[0], shape.c[1], 0.5), (shape.s[0], shape.s[1], 1e30))
    raise TypeError("2D shapes must be Disk or Rect")


class _Model2D:
    ECMP = ()   # 3D components of E
    HCMP = ()   # 3D components of H

    def __init__(self, grid: Grid2):
        self.grid = grid
        self.m3 = M.ModelFull(grid.grid3)

    # reference-style accessors (2D views of the 3D state)
    @property
    def εarr(self):
        c = self.ECMP
        return self.m3.εarr[:, :, 0][:, :, c][:, :, :, c]

    @property
    def μarr(self):
        c = self.HCMP
        return self.m3.μarr[:, :, 0][:, :, c][:, :, :, c]

    def _sel(self, cmps):
        """3D DOF index of every 2D DOF (component-first 2D order)."""
        Mn = self.grid.M
        p = np.arange(Mn)
        K = len(cmps)
        out = np.empty(K * Mn, dtype=np.int64)
        for k, c in enumerate(cmps):
            out[k::K] = c + 3 * p
        return out

    def esel(self):
        return self._sel(self.ECMP)

    def hsel(self):
        return self._sel(self.HCMP)


class ModelTE(_Model2D):
    """2D TE model (te.jl:4-13): E = (Ex, Ey), H = Hz."""
    ECMP = (0, 1)
    HCMP = (2,)


class ModelTM(_Model2D):
    """2D TM model (tm.jl:4-13): E = Ez, H = (Hx, Hy)."""
    ECMP = (2,)
    HCMP = (0, 1)


# setters (model.jl:86-89) on the 2D model
def set_ωpml(mdl, ωpml):
    M.set_ωpml(mdl.m3, ωpml)


def set_Npml(mdl, Npml):
    M.set_Npml(mdl.m3, ((Npml[0][0], Npml[0][1], 0), (Npml[1][0], Npml[1][1], 0)))


def set_kbloch(mdl, kbloch):
    M.set_kbloch(mdl.m3, (kbloch[0], kbloch[1], 0.0))


def set_boundft(mdl, boundft):
    M.set_boundft(mdl.m3, (boundft[0], boundft[1], M.EE))


set_omegapml = set_ωpml


def add_obj(mdl, matname, *shapes, ε=1.0, μ=1.0):
    """add_obj! with 2D shapes (extruded along z); ε, μ are scalars or 3x3 tensors."""
    M.add_obj(mdl.m3, matname, *[_extrude(s) for s in shapes], ε=ε, μ=μ)


class PointSrc2:
    """2D point source at c = (x, y) with polarisation p over the 2D model's
    current components (2 for the two-component field, 1 for the other)."""

    def __init__(self, c, p, Idr=1.0):
        self.c = tuple(float(v) for v in c)
        self.p = np.atleast_1d(np.asarray(p, dtype=np.float64))
        self.Idr = complex(Idr)


def _add_src(mdl, ft, src):
    cmps = mdl.ECMP if ft == M.EE else mdl.HCMP
    if len(src.p) != len(cmps):
        raise MfdArgumentError(L.MFD_ERR_ARG, f"polarisation needs {len(cmps)} components")
    p3 = np.zeros(3)
    for k, c in enumerate(cmps):
        p3[c] = src.p[k]
    M._add_src(mdl.m3, ft, M.PointSrc((src.c[0], src.c[1], 0.5), p3, src.Idr))


def add_srce(mdl, src):
    _add_src(mdl, M.EE, src)


def add_srcm(mdl, src):
    _add_src(mdl, M.HH, src)


def calc_matparams(mdl):
    M.calc_matparams(mdl.m3)


def create_srcs(mdl):
    je, jm = M.create_srcs(mdl.m3)
    return je[mdl.esel()], jm[mdl.hsel()]


def _restrict(A3, sel):
    A = A3.tocsr()[sel][:, sel].tocsc()
    A.eliminate_zeros()   # the z differences of the extrusion vanish identically
    A.sort_indices()
    return A


def create_A(ft, ω, mdl):
    """create_A (model.jl:225-246) of the 2D model, assembled on the device."""
    if ft != M.EE:
        raise NotImplementedError("2D models: ft = EE only")
    return _restrict(M.create_A(ft, ω, mdl.m3), mdl.esel())


def create_b(ft, ω, mdl, jₑ, jₘ):
    """create_b (model.jl:251-274) of the 2D model: b = −Cₘ μ⁻¹ jₘ − iω jₑ."""
    if ft != M.EE:
        raise NotImplementedError("2D models: ft = EE only")
    je3 = np.zeros(3 * mdl.grid.M, dtype=np.complex128)
    jm3 = np.zeros_like(je3)
    je3[mdl.esel()] = jₑ
    jm3[mdl.hsel()] = jₘ
    return M.create_b(ft, ω, mdl.m3, je3, jm3)[mdl.esel()]


def create_linsys(ft, ω, mdl):
    je, jm = create_srcs(mdl)
    return create_A(ft, ω, mdl), create_b(ft, ω, mdl, je, jm)


class _WindowOperator:
    """Device operator and solver of a 2D or 1D model: the single-cell 3D
    operator restricted to the model's E components (mfd_op_set_components),
    so the vectors hold K values per cell, not 3.  Model vectors are
    component-first (c + K·p); the device vectors are component-last
    (p + M·k), and the two orders are exchanged on the host.  Where the
    window is not built (μ ≠ 1, boundft = HH) the full single-cell operator
    serves, with the model's components embedded in three per cell."""

    def __init__(self, mdl, ω, m3, ecmp, esel, hsel):
        self.mdl, self.ω, self._m3 = mdl, ω, m3
        self.K = len(ecmp)
        self.Mc = m3.grid.M
        self.n = self.K * self.Mc
        self._esel, self._hsel = esel, hsel
        self._full = None   # the 3D operator, for the field conversions
        self.op = M.MaxwellOperator(m3, ω, order_cmpfirst=False)
        try:
            self.op.set_components(ecmp[0], self.K)
            self.windowed = True
        except MfdUnsupportedError:
            self.op.close()
            self.op = self._full = M.MaxwellOperator(m3, ω)
            self.windowed = False

    def close(self):
        self.op.close()
        if self._full is not None and self._full is not self.op:
            self._full.close()

    def _to_dev(self, x):
        x = np.atleast_2d(np.asarray(x, dtype=np.complex128))
        if not self.windowed:
            X = np.zeros((x.shape[0], 3 * self.Mc), dtype=np.complex128)
            X[:, self._esel] = x
            return X
        return np.ascontiguousarray(x.reshape(-1, self.Mc, self.K).transpose(0, 2, 1)).reshape(-1, self.n)

    def _from_dev(self, y, shape):
        y = np.atleast_2d(y)
        if not self.windowed:
            return y[:, self._esel].reshape(shape)
        return np.ascontiguousarray(y.reshape(-1, self.K, self.Mc).transpose(0, 2, 1)).reshape(shape)

    def apply(self, x):
        return self._from_dev(self.op.apply(self._to_dev(x)), np.shape(x))

    def solve(self, b, **kw):
        if kw.get("x0") is not None:
            kw["x0"] = self._to_dev(kw["x0"])
        e, st = self.op.solve(self._to_dev(b), **kw)
        return self._from_dev(e, np.shape(b)), st

    def h_from_e(self, e, jm=None):
        if self._full is None:
            self._full = M.MaxwellOperator(self._m3, self.ω)
        e3 = np.zeros(3 * self.Mc, dtype=np.complex128)
        e3[self._esel] = np.ravel(e)
        jm3 = None
        if jm is not None:
            jm3 = np.zeros(3 * self.Mc, dtype=np.complex128)
            jm3[self._hsel] = jm
        return self._full.h_from_e(e3, jm3)[self._hsel]


class MaxwellOperator2D(_WindowOperator):
    """Device operator and solver of a 2D model: (Ex, Ey) unknowns for
    ModelTE, Ez for ModelTM."""

    def __init__(self, mdl, ω):
        super().__init__(mdl, ω, mdl.m3, mdl.ECMP, mdl.esel(), mdl.hsel())
