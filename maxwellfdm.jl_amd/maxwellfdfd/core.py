"""Host-side mirror of the MaxwellFDFD assembly and solve API.

Names and argument meaning follow the reference package's exports
(src/model/model.jl:23-28) with Julia's `!` dropped and Unicode superscripts
spelled out (create_e⁻ⁱᵏᴸ -> create_eikL).  Every numerical entry point runs
on the GPU through libmfdfd's C ABI; the Python layer only builds argument
structs and wraps outputs (CSC results become scipy.sparse.csc_matrix with
0-based indices; the raw 1-based Julia CSC is available from the *_csc
functions).
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np

from . import _lib as L
from ._lib import MfdArgumentError, check, context, lib, ptr

EE, HH = 0, 1            # FieldType
PRIM, DUAL = 0, 1        # GridType
nX, nY, nZ = 0, 1, 2


# --------------------------------------------------------------------- grid

class Grid:
    """Staggered 3D grid from primal plane locations (MaxwellBase Grid(lprim, isbloch)).

    lprim[w] holds N[w]+1 primal planes including both bounds
    (test/source.jl:141-142: -10:10 gives 20 cells).  Conventions of the
    reference's use of Grid (test/source.jl:6-27, pointsrc.jl:66-68):
    l[PRIM][w] = primal vertices without the ghost at the positive bound;
    l[DUAL][w] = midpoints; ∆l[PRIM][w] = diff(dual incl. the negative ghost);
    ∆l[DUAL][w] = diff(primal incl. the positive ghost).
    """

    def __init__(self, lprim, isbloch=(False, False, False)):
        self.lprim = tuple(np.ascontiguousarray(np.asarray(a, dtype=np.float64)) for a in lprim)
        if len(self.lprim) != 3:
            raise ValueError("Grid needs 3 axes")
        self.isbloch = tuple(bool(b) for b in isbloch)
        self.N = tuple(len(a) - 1 for a in self.lprim)
        if min(self.N) < 1:
            raise ValueError("each axis needs at least one cell")
        for a in self.lprim:
            if np.any(np.diff(a) <= 0):
                raise ValueError("lprim must be strictly increasing")
        self.bounds = (np.array([a[0] for a in self.lprim]), np.array([a[-1] for a in self.lprim]))
        self.L = self.bounds[1] - self.bounds[0]
        lp = [a[:-1].copy() for a in self.lprim]
        ld = [0.5 * (a[:-1] + a[1:]) for a in self.lprim]
        ghost = [ld[w][-1] - self.L[w] if self.isbloch[w] else 2.0 * self.lprim[w][0] - ld[w][0] for w in range(3)]
        ldg = [np.concatenate(([ghost[w]], ld[w])) for w in range(3)]
        self.l = (tuple(lp), tuple(ld))
        self.dl = (tuple(np.diff(ldg[w]) for w in range(3)), tuple(np.diff(self.lprim[w]) for w in range(3)))
        self.ghosted_l = (tuple(self.lprim), tuple(ldg))

    @property
    def M(self):
        return int(np.prod(self.N))

    def __repr__(self):
        return f"Grid(N={self.N}, isbloch={self.isbloch})"


# ------------------------------------------------------------------- shapes

@dataclass
class Ball:
    """GeometryPrimitives Ball(c, r)."""
    c: Tuple[float, float, float]
    r: float


@dataclass
class Cuboid:
    """Axis-aligned GeometryPrimitives Cuboid(c, s) (centre, side lengths)."""
    c: Tuple[float, float, float]
    s: Tuple[float, float, float]


@dataclass
class Cylinder:
    """Cylinder along z: centre c, radius r, height h (a 2D disk extruded along z)."""
    c: Tuple[float, float, float]
    r: float
    h: float = 1e30


Sphere, Box = Ball, Cuboid


def _tensor(v):
    a = np.asarray(v, dtype=np.complex128)
    if a.ndim == 0:
        return np.eye(3, dtype=np.complex128) * a
    if a.shape == (3,):
        return np.diag(a)
    if a.shape == (3, 3):
        return a.copy()
    raise ValueError("material parameter must be a scalar, a 3-vector (diagonal) or a 3x3 tensor")


@dataclass
class Material:
    name: str
    eps: np.ndarray
    mu: np.ndarray


@dataclass
class Object:
    shape: object
    mat: Material


# ------------------------------------------------------------------ sources

class PointSrc:
    """Point dipole PointSrc(c, p, I∆r) (pointsrc.jl:51-62); p is normalised."""

    def __init__(self, c, p, Idr=1.0):
        self.c = np.asarray(c, dtype=np.float64)
        p = np.asarray(p, dtype=np.float64)
        self.p = p / np.sqrt(np.sum(p * p))
        self.Idr = complex(Idr)


class PlaneSrc:
    """Plane source PlaneSrc(n, c, p, J∆n) (planesrc.jl:14-33): n is a Cartesian
    normal, c the intercept along it, p the polarisation (normalised), J∆n the
    complex sheet current."""

    def __init__(self, n, c, p, Jdn=1.0):
        n = np.asarray(n, dtype=np.float64)
        if np.count_nonzero(n) != 1:
            raise MfdArgumentError(L.MFD_ERR_ARG, f"n = {list(n)} must be along Cartesian direction.")
        self.n = n / np.sqrt(np.sum(n * n))
        self.c = float(c)
        p = np.asarray(p, dtype=np.float64)
        self.p = p / np.sqrt(np.sum(p * p))
        self.Jdn = complex(Jdn)


def distweights(c, gt, bounds, l, dl, isbloch):
    """source.jl:153-154/208-306 on the C ABI: returns ((ind1, ind2), (wt1, wt2)),
    1-based indices; raises MfdArgumentError where the reference throws."""
    l = np.ascontiguousarray(l, dtype=np.float64)
    dl = np.ascontiguousarray(dl, dtype=np.float64)
    b = (C.c_double * 2)(float(bounds[0]), float(bounds[1]))
    ind = (C.c_int64 * 2)()
    wt = (C.c_double * 2)()
    check(lib().mfd_distweights(float(c), int(gt), b, ptr(l), ptr(dl), len(l), int(bool(isbloch)), ind, wt))
    return (ind[0], ind[1]), (wt[0], wt[1])


def create_field_array(N):
    return np.zeros(tuple(N) + (3,), dtype=np.complex128, order="F")


def create_param_array(N):
    return np.zeros(tuple(N) + (3, 3), dtype=np.complex128, order="F")


# -------------------------------------------------------------------- model

class Model:
    """MaxwellFDFD.Model for the full 3D equations (model.jl:32-73, full.jl:4-13)."""

    def __init__(self, grid: Grid):
        self.grid = grid
        self.ωpml = 0.0
        self.boundft = [EE, EE, EE]
        self.Npml = ((0, 0, 0), (0, 0, 0))
        self.kbloch = np.zeros(3)
        self.εarr = create_param_array(grid.N)
        self.μarr = create_param_array(grid.N)
        self.jₑarr = create_field_array(grid.N)
        self.jₘarr = create_field_array(grid.N)
        self.objs: List[Object] = []
        self.order_cmpfirst = True
        self.pml_m = 4.0
        self.pml_lnR = -16.0

    # C-ABI grid description (holds references to the coordinate arrays)
    def _desc(self):
        g = L.mfd_grid()
        for w in range(3):
            g.N[w] = self.grid.N[w]
            g.lprim[w] = self.grid.lprim[w].ctypes.data
            g.isbloch[w] = int(self.grid.isbloch[w])
            g.boundft[w] = int(self.boundft[w])
            g.npml[0][w] = int(self.Npml[0][w])
            g.npml[1][w] = int(self.Npml[1][w])
            g.kbloch[w] = float(self.kbloch[w])
        g.omega_pml = L.cplx(self.ωpml)
        g.pml_m = self.pml_m
        g.pml_lnR = self.pml_lnR
        g.order_cmpfirst = int(bool(self.order_cmpfirst))
        g._keep = self.grid.lprim
        return g

    def size(self, ft=EE):
        return (3,) + tuple(self.grid.N) if self.order_cmpfirst else tuple(self.grid.N) + (3,)

    def length(self, ft=EE):
        return 3 * self.grid.M


ModelFull = Model


# setters (model.jl:86-89)
def set_ωpml(mdl, ωpml):
    mdl.ωpml = ωpml


def set_boundft(mdl, boundft):
    mdl.boundft = [int(b) for b in boundft]


def set_Npml(mdl, Npml):
    mdl.Npml = (tuple(int(v) for v in Npml[0]), tuple(int(v) for v in Npml[1]))


def set_kbloch(mdl, kbloch):
    mdl.kbloch = np.asarray(kbloch, dtype=np.float64)


set_omegapml = set_ωpml


def create_eikL(mdl):
    """create_e⁻ⁱᵏᴸ (model.jl:91)."""
    return np.exp(-1j * mdl.kbloch * mdl.grid.L)


def clear_objs(mdl):
    mdl.εarr[...] = 0
    mdl.μarr[...] = 0
    mdl.objs.clear()


def clear_srcs(mdl):
    mdl.jₑarr[...] = 0
    mdl.jₘarr[...] = 0


def add_obj(mdl, matname, *shapes, ε=1.0, μ=1.0):
    """add_obj!(mdl, name, shapes...; ε, μ) (model.jl:107-120); later objects win."""
    if len(shapes) == 1 and isinstance(shapes[0], (list, tuple)):
        shapes = tuple(shapes[0])
    mat = Material(matname, _tensor(ε), _tensor(μ))
    for s in shapes:
        if not isinstance(s, (Ball, Cuboid, Cylinder)):
            raise TypeError("shapes must be Ball, Cuboid or Cylinder")
        mdl.objs.append(Object(s, mat))


def _add_src(mdl, ft, src):
    jarr = mdl.jₑarr if ft == EE else mdl.jₘarr
    p = (C.c_double * 3)(*src.p)
    g = mdl._desc()
    if isinstance(src, PlaneSrc):
        # the reference locates the normal with findfirst(n .== 1) (planesrc.jl:42),
        # so only a positive Cartesian normal names an axis
        if not np.any(src.n == 1):
            raise MfdArgumentError(L.MFD_ERR_ARG, f"plane normal {list(src.n)} must be a positive Cartesian direction")
        nn = int(np.flatnonzero(src.n == 1)[0])
        check(lib().mfd_add_planesrc(C.byref(g), ft, nn, src.c, p, L.cplx(src.Jdn), ptr(jarr)))
        return
    c = (C.c_double * 3)(*src.c)
    check(lib().mfd_add_pointsrc(C.byref(g), ft, c, p, L.cplx(src.Idr), ptr(jarr)))


def add_srce(mdl, src):
    """add_srce! (model.jl:184-191)."""
    _add_src(mdl, EE, src)


def add_srcm(mdl, src):
    """add_srcm! (model.jl:193-200)."""
    _add_src(mdl, HH, src)


def field_arr2vec(arr, order_cmpfirst=True):
    if order_cmpfirst:
        return np.ascontiguousarray(np.moveaxis(arr, 3, 0)).reshape(-1, order="F").copy()
    return np.asarray(arr).reshape(-1, order="F").copy()


def create_srcs(mdl):
    """create_srcs (model.jl:202-207): fresh DOF vectors (jₑ, jₘ)."""
    return field_arr2vec(mdl.jₑarr, mdl.order_cmpfirst), field_arr2vec(mdl.jₘarr, mdl.order_cmpfirst)


def create_stretched_dls(mdl):
    """Stretched ∆l tables ((PRIM x,y,z), (DUAL x,y,z)), complex (model.jl:126)."""
    g = mdl._desc()
    N = mdl.grid.N
    out = np.zeros(2 * sum(N), dtype=np.complex128)
    check(lib().mfd_stretched_dl(context(), C.byref(g), ptr(out)))
    res, off = [], 0
    for gt in range(2):
        row = []
        for w in range(3):
            row.append(out[off:off + N[w]].copy())
            off += N[w]
        res.append(tuple(row))
    return tuple(res)


def _csc_call(nnz_fn, csc_fn, args, n):
    nnz = C.c_int64()
    check(nnz_fn(*args, C.byref(nnz)))
    colptr = np.zeros(n + 1, dtype=np.int64)
    rowval = np.zeros(max(nnz.value, 1), dtype=np.int64)
    nzval = np.zeros(max(nnz.value, 1), dtype=np.complex128)
    check(csc_fn(*args, ptr(colptr), ptr(rowval), ptr(nzval)))
    return colptr, rowval[:nnz.value], nzval[:nnz.value]


def to_scipy(csc):
    import scipy.sparse as sp
    colptr, rowval, nzval = csc
    n = len(colptr) - 1
    return sp.csc_matrix((nzval, rowval - 1, colptr - 1), shape=(n, n))


def curl_csc(mdl, which):
    """Raw Julia-form CSC (1-based colptr, rowval, nzval) of Cₑ (0) or Cₘ (1)."""
    g = mdl._desc()
    return _csc_call(lib().mfd_curl_nnz, lib().mfd_curl_csc, (context(), C.byref(g), which), 3 * mdl.grid.M)


def Mcs_csc(mdl, which):
    """Raw Julia-form CSC of the corner-mean operator Mcₑ (0) or Mcₘ (1)."""
    g = mdl._desc()
    return _csc_call(lib().mfd_Mcs_nnz, lib().mfd_Mcs_csc, (context(), C.byref(g), which), 3 * mdl.grid.M)


def paramop_csc(mdl, which, parr):
    g = mdl._desc()
    parr = L.carr(parr, "F")
    return _csc_call(lib().mfd_paramop_nnz, lib().mfd_paramop_csc, (context(), C.byref(g), which, ptr(parr)),
                     3 * mdl.grid.M)


def A_csc(mdl, ω, εarr, μarr, ft=EE):
    g = mdl._desc()
    e = L.carr(εarr, "F")
    u = L.carr(μarr, "F")
    return _csc_call(lib().mfd_A_nnz, lib().mfd_A_csc, (context(), C.byref(g), ft, L.cplx(ω), ptr(e), ptr(u)),
                     3 * mdl.grid.M)


def _shapes_tables(mdl):
    n = len(mdl.objs)
    arr = (L.mfd_shape * max(n, 1))()
    et = np.zeros((max(n, 1), 9), dtype=np.complex128)
    mt = np.zeros((max(n, 1), 9), dtype=np.complex128)
    for i, o in enumerate(mdl.objs):
        s = o.shape
        if isinstance(s, Ball):
            arr[i].kind = 0
            arr[i].r = float(s.r)
        elif isinstance(s, Cylinder):
            arr[i].kind = 2
            arr[i].r = float(s.r)
            arr[i].s[2] = float(s.h)
        else:
            arr[i].kind = 1
            for a in range(3):
                arr[i].s[a] = float(s.s[a])
        for a in range(3):
            arr[i].c[a] = float(s.c[a])
        et[i] = o.mat.eps.reshape(-1, order="F")
        mt[i] = o.mat.mu.reshape(-1, order="F")
    return n, arr, et, mt


def calc_matparams(mdl):
    """calc_matparams! (full.jl:16-70): assignment + subpixel smoothing into εarr, μarr."""
    g = mdl._desc()
    n, arr, et, mt = _shapes_tables(mdl)
    check(lib().mfd_matparams(context(), C.byref(g), n, C.cast(arr, C.c_void_p), ptr(et), ptr(mt),
                              ptr(mdl.εarr), ptr(mdl.μarr)))


def create_paramops(mdl):
    """create_paramops (model.jl:141-158): updates εarr/μarr and returns (Pε, Pμ)."""
    calc_matparams(mdl)
    return to_scipy(paramop_csc(mdl, 0, mdl.εarr)), to_scipy(paramop_csc(mdl, 1, mdl.μarr))


def create_curls(mdl):
    """create_curls (model.jl:160-175): (Cₑ, Cₘ)."""
    return to_scipy(curl_csc(mdl, 0)), to_scipy(curl_csc(mdl, 1))


def create_Mcs(mdl):
    """create_Mcs (model.jl:287-306): (Mcₑ, Mcₘ), the operators that interpolate
    the solution fields E and H to the voxel corners."""
    return to_scipy(Mcs_csc(mdl, 0)), to_scipy(Mcs_csc(mdl, 1))


class _CscArg:
    """A sparse matrix as the Julia SparseMatrixCSC{ComplexF64,Int64} layout
    (1-based colptr/rowval, rows ascending, explicit zeros kept) handed to the
    C ABI; holds the arrays while the call runs."""

    def __init__(self, A):
        import scipy.sparse as sp
        A = sp.csc_matrix(A, dtype=np.complex128, copy=True)
        A.sum_duplicates()
        A.sort_indices()
        self.colptr = (A.indptr.astype(np.int64) + 1)
        self.rowval = (A.indices.astype(np.int64) + 1)
        self.nzval = np.ascontiguousarray(A.data, dtype=np.complex128)
        self.s = L.mfd_csc(A.shape[0], A.shape[1], A.nnz, self.colptr.ctypes.data,
                           self.rowval.ctypes.data if A.nnz else None, self.nzval.ctypes.data if A.nnz else None)

    def ref(self):
        return C.byref(self.s)


def _is_model(x):
    return isinstance(x, Model)


def _unpack(args, k):
    """The reference's tuple forms: create_A(ft, ω, (Pε, Pμ), (Cₑ, Cₘ)) etc."""
    out = []
    for a in args:
        out.extend(a if isinstance(a, tuple) else (a,))
    if len(out) != k:
        raise MfdArgumentError(L.MFD_ERR_ARG, f"expected {k} matrices/vectors, got {len(out)}")
    return out


def create_A(ft, ω, *args, εarr=None, μarr=None):
    """create_A (model.jl:222-246).

    create_A(ft, ω, Pε, Pμ, Cₑ, Cₘ) or create_A(ft, ω, (Pε, Pμ), (Cₑ, Cₘ)):
    the reference's signature; A = Cₘ (Pμ \\ Cₑ) − ω² Pε (EE) or
    Cₑ (Pε \\ Cₘ) − ω² Pμ (HH) formed on the device from the given matrices.
    create_A(ft, ω, mdl): the same A assembled on the device in closed form
    from the model's grid and parameter arrays (no matrices needed).
    """
    if len(args) == 1 and _is_model(args[0]):
        mdl = args[0]
        εarr = mdl.εarr if εarr is None else εarr
        μarr = mdl.μarr if μarr is None else μarr
        return to_scipy(A_csc(mdl, ω, εarr, μarr, ft))
    Pe, Pm, Ce, Cm = (_CscArg(a) for a in _unpack(args, 4))
    n = Pe.s.n
    return to_scipy(_csc_call(lib().mfd_A_from_csc_nnz, lib().mfd_A_from_csc,
                              (context(), ft, L.cplx(ω), Pe.ref(), Pm.ref(), Ce.ref(), Cm.ref()), n))


def create_b(ft, ω, *args, μarr=None, εarr=None):
    """create_b (model.jl:248-274): b = −Cₘ (Pμ \\ jₘ) − iω jₑ (ft = EE) or
    b = Cₑ (Pε \\ jₑ) − iω jₘ (ft = HH), on the device.

    create_b(ft, ω, Pε, Pμ, Cₑ, Cₘ, jₑ, jₘ) or create_b(ft, ω, (Pε, Pμ),
    (Cₑ, Cₘ), (jₑ, jₘ)): from the given matrices, as the reference.
    create_b(ft, ω, mdl, jₑ, jₘ): from the model's grid and parameter arrays.
    """
    if _is_model(args[0]):
        return _create_b_model(ft, ω, *args, μarr=μarr, εarr=εarr)
    Pe, Pm, Ce, Cm, je, jm = _unpack(args, 6)
    je = L.carr(je)
    jm = L.carr(jm)
    mats = [_CscArg(a) for a in (Pe, Pm, Ce, Cm)]
    if je.size != mats[0].s.n or jm.size != mats[0].s.n:
        raise MfdArgumentError(L.MFD_ERR_ARG, "create_b: jₑ and jₘ must have n entries")
    b = np.zeros_like(je)
    check(lib().mfd_b_from_csc(context(), ft, L.cplx(ω), *(m.ref() for m in mats), ptr(je), ptr(jm), ptr(b)))
    return b


def _field_csc(fn, v, ω, P, Cx, j):
    v = L.carr(v)
    Pa, Ca = _CscArg(P), _CscArg(Cx)
    if v.size != Ca.s.n:
        raise MfdArgumentError(L.MFD_ERR_ARG, "field vector must have n entries")
    j = None if j is None else L.carr(j)
    if j is not None and j.size != v.size:
        raise MfdArgumentError(L.MFD_ERR_ARG, "source vector must have n entries")
    out = np.zeros_like(v)
    check(fn(context(), ptr(v), L.cplx(ω), Pa.ref(), Ca.ref(), None if j is None else ptr(j), ptr(out)))
    return out


def h_from_e(e, ω, *args):
    """h_from_e (model.jl:276-279): (i/ω) (Pμ \\ (Cₑ e + jₘ)) on the device.
    h_from_e(e, ω, Pμ, Cₑ, jₘ) or h_from_e(e, ω, (Pε, Pμ), (Cₑ, Cₘ), (jₑ, jₘ))."""
    if len(args) == 3 and all(isinstance(a, tuple) for a in args):
        M, Cs, js = args
        return _field_csc(lib().mfd_h_from_e_csc, e, ω, M[1], Cs[0], js[1])
    Pm, Ce, jm = args
    return _field_csc(lib().mfd_h_from_e_csc, e, ω, Pm, Ce, jm)


def e_from_h(h, ω, *args):
    """e_from_h (model.jl:281-284): (−i/ω) (Pε \\ (Cₘ h − jₑ)) on the device.
    e_from_h(h, ω, Pε, Cₘ, jₑ) or e_from_h(h, ω, (Pε, Pμ), (Cₑ, Cₘ), (jₑ, jₘ))."""
    if len(args) == 3 and all(isinstance(a, tuple) for a in args):
        M, Cs, js = args
        return _field_csc(lib().mfd_e_from_h_csc, h, ω, M[0], Cs[1], js[0])
    Pe, Cm, je = args
    return _field_csc(lib().mfd_e_from_h_csc, h, ω, Pe, Cm, je)


def _create_b_model(ft, ω, mdl, jₑ, jₘ, μarr=None, εarr=None):
    μarr = mdl.μarr if μarr is None else μarr
    εarr = mdl.εarr if εarr is None else εarr
    g = mdl._desc()
    je = L.carr(jₑ)
    jm = L.carr(jₘ)
    b = np.zeros_like(je)
    u = L.carr(μarr, "F")
    e = L.carr(εarr, "F")
    check(lib().mfd_create_b(context(), C.byref(g), ft, L.cplx(ω), ptr(e), ptr(u), ptr(je), ptr(jm), ptr(b)))
    return b


def create_linsys(ft, ω, *args):
    """create_linsys (model.jl:209-220): (A, b).

    create_linsys(ft, ω, Pε, Pμ, Cₑ, Cₘ, jₑ, jₘ) or with the tuples (Ps, Cs,
    js), as the reference; create_linsys(ft, ω, mdl[, jₑ, jₘ]) from a model.
    """
    if _is_model(args[0]):
        mdl = args[0]
        jₑ, jₘ = args[1:3] if len(args) >= 3 else create_srcs(mdl)
        return create_A(ft, ω, mdl), create_b(ft, ω, mdl, jₑ, jₘ)
    Pe, Pm, Ce, Cm, je, jm = _unpack(args, 6)
    return create_A(ft, ω, Pe, Pm, Ce, Cm), create_b(ft, ω, Pe, Pm, Ce, Cm, je, jm)


# ------------------------------------------------------- matrix-free solve

def _solve_batch(ops, bs, es, dev, o, raise_on_noconv=True):
    count = len(ops)
    if not (count == len(bs) == len(es)):
        raise MfdArgumentError(L.MFD_ERR_ARG, "solve_batch: one b and one e per operator")
    hs = (C.c_void_p * count)(*[op._h for op in ops])
    if dev:
        bp = (C.c_void_p * count)(*bs)
        ep = (C.c_void_p * count)(*es)
    else:
        bp = (C.c_void_p * count)(*[ptr(b) for b in bs])
        ep = (C.c_void_p * count)(*[ptr(e) for e in es])
    st = (L.mfd_solve_stats * count)()
    fn = lib().mfd_solve_batch_dev if dev else lib().mfd_solve_batch
    rc = fn(count, hs, bp, ep, C.byref(o), st)
    if rc != L.MFD_OK and (raise_on_noconv or rc != L.MFD_ERR_NOCONV):
        check(rc)
    return list(st)


def solve_batch(ops, bs, tol=1e-8, precision="fp64", maxit=None, inner_tol=None, check_every=None,
                use_graph=True, raise_on_noconv=True):
    """Solve independent systems ops[p] e_p = bs[p] (each its own grid, ε, ω)
    in one lockstep device loop; returns ([e_p], [stats_p])."""
    bs = [L.carr(b) for b in bs]
    for op, b in zip(ops, bs):
        if b.size != op.n:
            raise MfdArgumentError(L.MFD_ERR_ARG, "solve_batch: b has the wrong length for its operator")
    es = [np.zeros_like(b) for b in bs]
    o = MaxwellOperator._opts(tol, precision, maxit, inner_tol, check_every, use_graph)
    st = _solve_batch(list(ops), bs, es, False, o, raise_on_noconv)
    return es, st


def solve_linsys(A, b, tol=1e-8, precision="fp64", **kw):
    """The user's `A \\ b` on the device: A a MaxwellOperator (matrix-free) or
    the sparse A of create_A / create_linsys (uploaded as a sparse operator).
    Returns (e, stats)."""
    if isinstance(A, MaxwellOperator):
        return A.solve(b, tol=tol, precision=precision, **kw)
    op = MaxwellOperator.from_csc(A)
    try:
        return op.solve(b, tol=tol, precision=precision, **kw)
    finally:
        op.close()


class MaxwellOperator:
    """Device-resident A of create_A (model.jl:225-246) with its solver:
    ft = EE, A = Cₘ μ⁻¹ Cₑ − ω² Pε; ft = HH, A = Cₑ ε⁻¹ Cₘ − ω² Pμ (built as
    the identical EE operator of the model with every boundft flipped and ε,
    μ swapped).  Any boundft.

    solve() is the user's `A \\ b` of the reference workflow (iterative:
    Jacobi-scaled symmetric QMR; precision "fp64" or "mixed" = FP32 inner
    iterations with FP64 residual and update).
    """

    def __init__(self, mdl, ω, ft=EE, εarr=None, μarr=None, order_cmpfirst=None):
        self.mdl = mdl
        self.ω = complex(ω)
        self.n = 3 * mdl.grid.M
        e = L.carr(mdl.εarr if εarr is None else εarr, "F")
        u = L.carr(mdl.μarr if μarr is None else μarr, "F")
        g = mdl._desc()
        if order_cmpfirst is not None:   # DOF order of this operator's vectors, else the model's
            g.order_cmpfirst = int(bool(order_cmpfirst))
        h = C.c_void_p()
        check(lib().mfd_op_create(context(), C.byref(g), ft, L.cplx(ω), ptr(e), ptr(u), C.byref(h)))
        self._h = h

    @classmethod
    def from_csc(cls, A):
        """Operator from an assembled sparse A (mfd_op_create_csc): applied as a
        sparse product; SQMR when A equals its transpose, BiCGSTAB otherwise."""
        a = _CscArg(A)
        self = object.__new__(cls)
        self.mdl = None
        self.ω = None
        self.n = a.s.n
        h = C.c_void_p()
        check(lib().mfd_op_create_csc(context(), a.ref(), C.byref(h)))
        self._h = h
        return self

    def _nrhs(self, v, what):
        """Number of stacked DOF vectors in v; raises unless v holds a positive whole number of them."""
        if v.size == 0 or v.size % self.n:
            raise MfdArgumentError(L.MFD_ERR_ARG, f"{what} must hold k·{self.n} values (k ≥ 1), got {v.size}")
        return v.size // self.n

    def _one(self, v, what):
        if v.size != self.n:
            raise MfdArgumentError(L.MFD_ERR_ARG, f"{what} must hold {self.n} values, got {v.size}")

    def close(self):
        if getattr(self, "_h", None):
            lib().mfd_op_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def apply(self, x):
        x = L.carr(x)
        nrhs = self._nrhs(x, "x")
        y = np.zeros_like(x)
        check(lib().mfd_op_apply(self._h, ptr(x), ptr(y), nrhs))
        return y

    def apply_dev(self, x_ptr, y_ptr, nrhs=1, precision="fp64"):
        check(lib().mfd_op_apply_dev(self._h, C.c_void_p(x_ptr), C.c_void_p(y_ptr), nrhs,
                                     L.MFD_FP64 if precision == "fp64" else L.MFD_MIXED))

    def set_components(self, c0, nc):
        """Restrict the unknowns to the E components c0 .. c0+nc−1 (the 2D and 1D
        models; see mfd_op_set_components): vectors of nc·M values, component-last."""
        check(lib().mfd_op_set_components(self._h, int(c0), int(nc)))
        self.n = int(nc) * self.mdl.grid.M

    def set_graddiv(self, enable=True):
        """Grad-div augmented inner operator of the solver (see mfd_op_set_graddiv)."""
        check(lib().mfd_op_set_graddiv(self._h, int(bool(enable))))

    def set_tile(self, precision, ty, kz, xcd_order=1):
        """Stencil tile shape: TY rows per block (1, 2, 4, 8; 0 = direct kernel), KZ planes per
        block, and tile order (bit 0: a contiguous run of tiles per XCD, else dispatch order;
        bit 1: 32-wide tiles of 2·TY rows, for diagonal ε)."""
        check(lib().mfd_op_set_tile(self._h, L.MFD_FP64 if precision == "fp64" else L.MFD_MIXED, ty, kz,
                                    xcd_order))

    def autotune(self, nrhs=1, iters=10):
        """Pick the stencil tile shape by timing; returns ((TY, KZ, order) fp64,
        (TY, KZ, order) fp32, (ms64, ms32))."""
        t64 = (C.c_int * 3)()
        t32 = (C.c_int * 3)()
        ms = (C.c_double * 2)()
        check(lib().mfd_op_autotune(self._h, nrhs, iters, t64, t32, ms))
        return tuple(t64), tuple(t32), tuple(ms)

    def bench_apply(self, nrhs=1, precision="fp64", iters=20):
        ms = C.c_double()
        check(lib().mfd_op_bench_apply(self._h, nrhs, L.MFD_FP64 if precision == "fp64" else L.MFD_MIXED, iters,
                                       C.byref(ms)))
        return ms.value

    def bench_iteration(self, nrhs=1, precision="fp64", iters=20):
        """Average ms of the three kernels of one solver iteration:
        (stencil with qᵀPq, r update, d/z/q update).  Overwrites the solver workspace."""
        ms = (C.c_double * 3)()
        check(lib().mfd_op_bench_iteration(self._h, nrhs, L.MFD_FP64 if precision == "fp64" else L.MFD_MIXED,
                                           iters, ms))
        return tuple(ms)

    @staticmethod
    def _opts(tol, precision, maxit, inner_tol, check_every, use_graph, method=None):
        o = L.mfd_solve_opts()
        lib().mfd_solve_opts_default(C.byref(o))
        if method is not None:
            o.method = {"sqmr": L.MFD_SQMR, "cocg": L.MFD_COCG, "bicgstab": L.MFD_BICGSTAB}[method]
        o.tol = tol
        o.precision = L.MFD_MIXED if precision == "mixed" else L.MFD_FP64
        if maxit is not None:
            o.maxit = maxit
        if inner_tol is not None:
            o.inner_tol = inner_tol
        if check_every is not None:
            o.check_every = check_every
        o.use_graph = int(bool(use_graph))
        return o

    def solve(self, b, tol=1e-8, precision="fp64", x0=None, maxit=None, inner_tol=None,
              check_every=None, use_graph=True, raise_on_noconv=True, method=None):
        b = L.carr(b)
        nrhs = self._nrhs(b, "b")
        e = np.zeros_like(b) if x0 is None else L.carr(x0).copy()
        if e.size != b.size:
            raise MfdArgumentError(L.MFD_ERR_ARG, f"x0 must hold as many values as b ({b.size}), got {e.size}")
        o = self._opts(tol, precision, maxit, inner_tol, check_every, use_graph, method)
        st = L.mfd_solve_stats()
        rc = lib().mfd_solve(self._h, ptr(b), ptr(e), nrhs, C.byref(o), C.byref(st))
        if rc != L.MFD_OK and (raise_on_noconv or rc != L.MFD_ERR_NOCONV):
            check(rc)
        return e, st

    def solve_dev(self, b_ptr, e_ptr, nrhs=1, tol=1e-8, precision="fp64", maxit=None, inner_tol=None,
                  check_every=None, use_graph=True, method=None, raise_on_noconv=True):
        o = self._opts(tol, precision, maxit, inner_tol, check_every, use_graph, method)
        st = L.mfd_solve_stats()
        rc = lib().mfd_solve_dev(self._h, C.c_void_p(b_ptr), C.c_void_p(e_ptr), nrhs, C.byref(o), C.byref(st))
        if rc != L.MFD_OK and (raise_on_noconv or rc != L.MFD_ERR_NOCONV):
            check(rc)
        return st

    def solve_dev_noalloc(self, b_ptr, e_ptr, nrhs=1, tol=1e-8, precision="fp64", maxit=None, inner_tol=None,
                          check_every=None, method=None, raise_on_noconv=True):
        """solve_dev without any device allocation: the workspace lives in the
        library's fixed per-GPU scratch array (see mfd_solve_dev_noalloc)."""
        o = self._opts(tol, precision, maxit, inner_tol, check_every, False, method)
        st = L.mfd_solve_stats()
        rc = lib().mfd_solve_dev_noalloc(self._h, C.c_void_p(b_ptr), C.c_void_p(e_ptr), nrhs, C.byref(o),
                                         C.byref(st))
        if rc != L.MFD_OK and (raise_on_noconv or rc != L.MFD_ERR_NOCONV):
            check(rc)
        return st

    def solve_batch_dev(self, others, b_ptrs, e_ptrs, tol=1e-8, precision="fp64", maxit=None, inner_tol=None,
                        check_every=None, use_graph=True):
        """Device-pointer form of `solve_batch` with this operator first."""
        return _solve_batch([self] + list(others), b_ptrs, e_ptrs, True,
                            self._opts(tol, precision, maxit, inner_tol, check_every, use_graph))

    def solve_adjoint(self, g, tol=1e-8, precision="fp64", maxit=None, method=None):
        g = L.carr(g)
        self._one(g, "g")
        lam = np.zeros_like(g)
        o = self._opts(tol, precision, maxit, None, None, True, method)
        st = L.mfd_solve_stats()
        check(lib().mfd_solve_adjoint(self._h, ptr(g), ptr(lam), C.byref(o), C.byref(st)))
        return lam, st

    def h_from_e(self, e, jm=None):
        """h_from_e (model.jl:276-279): h = (i/ω) Pμ⁻¹ (Cₑ e + jₘ), on the device."""
        e = L.carr(e)
        self._one(e, "e")
        h = np.zeros_like(e)
        jm = None if jm is None else L.carr(jm)
        if jm is not None:
            self._one(jm, "jm")
        check(lib().mfd_h_from_e(self._h, ptr(e), None if jm is None else ptr(jm), ptr(h)))
        return h

    def e_from_h(self, h, je=None):
        """e_from_h (model.jl:281-284): e = (−i/ω) Pε⁻¹ (Cₘ h − jₑ), on the device (diagonal ε)."""
        h = L.carr(h)
        self._one(h, "h")
        e = np.zeros_like(h)
        je = None if je is None else L.carr(je)
        if je is not None:
            self._one(je, "je")
        check(lib().mfd_e_from_h(self._h, ptr(h), None if je is None else ptr(je), ptr(e)))
        return e

    def solve_with_pullback(self, b, tol=1e-10, precision="fp64"):
        """The rrule of ε ↦ e = A(ε)⁻¹ b: returns e and pullback(g), which for
        g = (∂f/∂e)ᵀ of a real objective f solves Aᵀλ = g and returns ∂f/∂ε per
        E DOF (2 Re(ω² λ ∘ e))."""
        e, _ = self.solve(b, tol=tol, precision=precision)

        def pullback(g, tol_adjoint=tol):
            lam, _ = self.solve_adjoint(g, tol=tol_adjoint)
            return self.adjoint_grad(e, lam)

        return e, pullback

    def adjoint_grad(self, e, lam):
        """∂f/∂ε per E DOF: 2 Re(ω² λ ∘ e) for Aᵀλ = (∂f/∂e)ᵀ."""
        e = L.carr(e)
        lam = L.carr(lam)
        self._one(e, "e")
        self._one(lam, "lam")
        grad = np.zeros(e.size, dtype=np.float64)
        check(lib().mfd_adjoint_grad(self._h, ptr(e), ptr(lam), ptr(grad)))
        return grad
