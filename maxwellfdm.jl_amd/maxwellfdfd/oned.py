"""1D model: ModelTEM (E = Ex, H = Hy, variation along z) of MaxwellFDFD
(src/model/tem.jl:4-13).

A 1D model is the 3D model of a single x-y cell with Bloch (k = 0) x and y
boundaries and shapes extruded across x and y: every x and y difference then
vanishes, the 3D curl-curl operator splits into independent Ex, Ey and Ez
blocks, and the Ex block with the Hy field is the TEM system.  The 3D device
path (assembly, subpixel smoothing, matrix-free solve) serves it unchanged.
1D DOF vectors hold one component per cell (index k along z); 1D parameter
arrays are (Nz, 1, 1) for that component (tem.jl:4-13: ε rank 0 on Ex, μ
rank 0 on Hy).
"""
from dataclasses import dataclass

import numpy as np

from . import core as M
from ._lib import MfdArgumentError
from . import _lib as L
from .twod import _WindowOperator

ECMP = (0,)   # Ex
HCMP = (1,)   # Hy


class Grid1:
    """1D staggered grid from the primal z plane locations."""

    def __init__(self, lprim, isbloch=False):
        lprim = np.asarray(lprim, dtype=np.float64)
        self.grid3 = M.Grid([np.array([0.0, 1.0]), np.array([0.0, 1.0]), lprim], (True, True, bool(isbloch)))
        self.N = (self.grid3.N[2],)
        self.lprim = (self.grid3.lprim[2],)
        self.isbloch = (bool(isbloch),)

    @property
    def M(self):
        return self.N[0]


@dataclass
class Segment:
    """1D interval (GeometryPrimitives Cuboid in 1D): centre c, length s."""
    c: float
    s: float


def _extrude(shape):
    if isinstance(shape, Segment):
        return M.Cuboid((0.5, 0.5, float(shape.c)), (1e30, 1e30, float(shape.s)))
    raise TypeError("1D shapes must be Segment")


class ModelTEM:
    """1D TEM model (tem.jl:4-13): E = Ex, H = Hy, shapes along z."""

    def __init__(self, grid: Grid1):
        self.grid = grid
        self.m3 = M.ModelFull(grid.grid3)

    @property
    def εarr(self):
        return self.m3.εarr[0, 0, :, 0, 0].reshape(self.grid.N[0], 1, 1)

    @property
    def μarr(self):
        return self.m3.μarr[0, 0, :, 1, 1].reshape(self.grid.N[0], 1, 1)

    @staticmethod
    def _sel(c, n):
        return c + 3 * np.arange(n, dtype=np.int64)

    def esel(self):
        return self._sel(ECMP[0], self.grid.M)

    def hsel(self):
        return self._sel(HCMP[0], self.grid.M)


# setters (model.jl:86-89) on the 1D model
def set_ωpml(mdl, ωpml):
    M.set_ωpml(mdl.m3, ωpml)


def set_Npml(mdl, Npml):
    """Npml = (n_neg, n_pos) along z."""
    M.set_Npml(mdl.m3, ((0, 0, Npml[0]), (0, 0, Npml[1])))


def set_kbloch(mdl, kbloch):
    M.set_kbloch(mdl.m3, (0.0, 0.0, float(np.atleast_1d(kbloch)[0])))


def set_boundft(mdl, boundft):
    M.set_boundft(mdl.m3, (M.EE, M.EE, np.atleast_1d(boundft)[0]))


set_omegapml = set_ωpml


def add_obj(mdl, matname, *shapes, ε=1.0, μ=1.0):
    """add_obj! with 1D shapes (extruded across x and y); ε, μ are scalars."""
    M.add_obj(mdl.m3, matname, *[_extrude(s) for s in shapes], ε=ε, μ=μ)


class PointSrc1:
    """1D point source at z = c carrying the current I∆r (Jx for add_srce,
    My for add_srcm)."""

    def __init__(self, c, Idr=1.0):
        self.c = float(c)
        self.Idr = complex(Idr)


def _add_src(mdl, ft, src):
    if not isinstance(src, PointSrc1):
        raise MfdArgumentError(L.MFD_ERR_ARG, "1D models take PointSrc1 sources")
    p3 = np.zeros(3)
    p3[ECMP[0] if ft == M.EE else HCMP[0]] = 1.0
    M._add_src(mdl.m3, ft, M.PointSrc((0.5, 0.5, src.c), p3, src.Idr))


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
    A.eliminate_zeros()   # the x and y differences of the extrusion vanish identically
    A.sort_indices()
    return A


def create_A(ft, ω, mdl):
    """create_A (model.jl:225-246) of the 1D model, assembled on the device."""
    if ft != M.EE:
        raise NotImplementedError("1D model: ft = EE only")
    return _restrict(M.create_A(ft, ω, mdl.m3), mdl.esel())


def create_b(ft, ω, mdl, jₑ, jₘ):
    """create_b (model.jl:251-274) of the 1D model: b = −Cₘ μ⁻¹ jₘ − iω jₑ."""
    if ft != M.EE:
        raise NotImplementedError("1D model: ft = EE only")
    je3 = np.zeros(3 * mdl.grid.M, dtype=np.complex128)
    jm3 = np.zeros_like(je3)
    je3[mdl.esel()] = jₑ
    jm3[mdl.hsel()] = jₘ
    return M.create_b(ft, ω, mdl.m3, je3, jm3)[mdl.esel()]


def create_linsys(ft, ω, mdl):
    je, jm = create_srcs(mdl)
    return create_A(ft, ω, mdl), create_b(ft, ω, mdl, je, jm)


class MaxwellOperator1D(_WindowOperator):
    """Device operator and solver of the 1D model: the Ex unknowns only."""

    def __init__(self, mdl, ω):
        super().__init__(mdl, ω, mdl.m3, ECMP, mdl.esel(), mdl.hsel())
