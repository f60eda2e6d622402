"""ctypes view of the C restatement in oracle.c.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
cpu_baseline leg of bench.py, and only as the checker.  The product package
never imports this module.

All functions take a plain grid spec (any object with the attributes
``lprim`` (3 arrays of N+1 primal plane locations), ``isbloch``, ``boundft``,
``npml`` ((neg,pos) x 3), ``omega_pml`` (complex), ``kbloch``, ``pml_m``,
``pml_lnR`` and ``order_cmpfirst``) and return numpy arrays.
"""
import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        _LIB = C.CDLL(path)
        _LIB.og_volfrac.restype = C.c_double
    return _LIB


class Desc(C.Structure):
    _fields_ = [("N", C.c_int64 * 3), ("lprim", C.c_void_p * 3), ("isbloch", C.c_int32 * 3),
                ("boundft", C.c_int32 * 3), ("npml", (C.c_int64 * 3) * 2),
                ("omega_pml", C.c_double * 2), ("kbloch", C.c_double * 3),
                ("pml_m", C.c_double), ("pml_lnR", C.c_double), ("order_cmpfirst", C.c_int32)]


class Shape(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("c", C.c_double * 3),
                ("r", C.c_double), ("s", C.c_double * 3)]


class CSC(C.Structure):
    _fields_ = [("nnz", C.c_int64), ("colptr", C.POINTER(C.c_int64)),
                ("rowval", C.POINTER(C.c_int64)), ("nzval", C.c_void_p)]


def _desc(spec):
    lp = [np.ascontiguousarray(np.asarray(a, dtype=np.float64)) for a in spec.lprim]
    d = Desc()
    for w in range(3):
        d.N[w] = len(lp[w]) - 1
        d.lprim[w] = lp[w].ctypes.data
        d.isbloch[w] = int(bool(spec.isbloch[w]))
        d.boundft[w] = int(spec.boundft[w])
        d.npml[0][w] = int(spec.npml[0][w])
        d.npml[1][w] = int(spec.npml[1][w])
        d.kbloch[w] = float(spec.kbloch[w])
    om = complex(spec.omega_pml)
    d.omega_pml[0], d.omega_pml[1] = om.real, om.imag
    d.pml_m = float(spec.pml_m)
    d.pml_lnR = float(spec.pml_lnR)
    d.order_cmpfirst = int(bool(spec.order_cmpfirst))
    d._keep = lp
    return d


def _N(spec):
    return tuple(len(a) - 1 for a in spec.lprim)


def _ptr(a):
    return C.c_void_p(a.ctypes.data)


def _check(rc, what):
    if rc != 0:
        raise ValueError(f"oracle {what} failed with code {rc}")


def _take_csc(m, n):
    nnz = m.nnz
    colptr = np.ctypeslib.as_array(m.colptr, shape=(n + 1,)).copy()
    rowval = np.ctypeslib.as_array(m.rowval, shape=(max(nnz, 1),))[:nnz].copy()
    nz = np.ctypeslib.as_array(C.cast(m.nzval, C.POINTER(C.c_double)), shape=(2 * max(nnz, 1),))
    nzval = nz[: 2 * nnz].copy().view(np.complex128)
    lib().og_csc_free(C.byref(m))
    return colptr, rowval, nzval


def grid_tables(spec):
    """l, ∆l and stretched ∆l tables: dict keyed (gt, axis)."""
    d = _desc(spec)
    N = _N(spec)
    tot = 2 * sum(N)
    l = np.zeros(tot)
    dl = np.zeros(tot)
    sdl = np.zeros(tot, dtype=np.complex128)
    _check(lib().og_grid_tables(C.byref(d), _ptr(l), _ptr(dl), _ptr(sdl)), "grid_tables")
    out = {}
    off = 0
    for gt in range(2):
        for w in range(3):
            out[("l", gt, w)] = l[off:off + N[w]]
            out[("dl", gt, w)] = dl[off:off + N[w]]
            out[("sdl", gt, w)] = sdl[off:off + N[w]]
            off += N[w]
    return out


def distweights(c, gt, bounds, l, dl, isbloch):
    """source.jl:208-306; returns (1-based indices, weights) or raises ValueError."""
    l = np.ascontiguousarray(l, dtype=np.float64)
    dl = np.ascontiguousarray(dl, dtype=np.float64)
    b = (C.c_double * 2)(float(bounds[0]), float(bounds[1]))
    ind = (C.c_int64 * 2)()
    wt = (C.c_double * 2)()
    rc = lib().og_distweights(C.c_double(float(c)), C.c_int(int(gt)), b, _ptr(l), _ptr(dl),
                              C.c_int64(len(l)), C.c_int(int(bool(isbloch))), ind, wt)
    if rc != 0:
        raise ValueError(f"distweights rejected c={c} (code {rc})")
    return (ind[0] + 1, ind[1] + 1), (wt[0], wt[1])


def add_pointsrc(spec, ft, c, p, Idr, jarr):
    """pointsrc.jl:64-103 into jarr (Nx,Ny,Nz,3) Fortran-ordered complex."""
    assert jarr.flags.f_contiguous and jarr.dtype == np.complex128
    p = np.asarray(p, dtype=np.float64)
    p = p / np.sqrt(np.sum(p * p))
    d = _desc(spec)
    cc = (C.c_double * 3)(*[float(v) for v in c])
    pp = (C.c_double * 3)(*[float(v) for v in p])
    Idr = complex(Idr)
    rc = lib().og_add_pointsrc(C.byref(d), C.c_int(int(ft)), cc, pp, C.c_double(Idr.real),
                               C.c_double(Idr.imag), _ptr(jarr))
    if rc != 0:
        raise ValueError(f"point source rejected (code {rc})")


def add_planesrc(spec, ft, n, c, p, Jdn, jarr):
    """planesrc.jl:41-79 into jarr (Nx,Ny,Nz,3); n is a Cartesian unit normal."""
    assert jarr.flags.f_contiguous and jarr.dtype == np.complex128
    n = np.asarray(n, dtype=np.float64)
    n = n / np.sqrt(np.sum(n * n))
    nn = int(np.flatnonzero(n == 1)[0])
    p = np.asarray(p, dtype=np.float64)
    p = p / np.sqrt(np.sum(p * p))
    pp = (C.c_double * 3)(*[float(v) for v in p])
    Jdn = complex(Jdn)
    rc = lib().og_add_planesrc(C.byref(_desc(spec)), C.c_int(int(ft)), C.c_int(nn), C.c_double(float(c)), pp,
                               C.c_double(Jdn.real), C.c_double(Jdn.imag), _ptr(jarr))
    if rc != 0:
        raise ValueError(f"plane source rejected (code {rc})")


def field_arr2vec(spec, jarr):
    d = _desc(spec)
    v = np.zeros(jarr.size, dtype=np.complex128)
    _check(lib().og_field_arr2vec(C.byref(d), _ptr(np.asfortranarray(jarr)), _ptr(v)), "arr2vec")
    return v


def curl(spec, which):
    d = _desc(spec)
    n = 3 * int(np.prod(_N(spec)))
    m = CSC()
    _check(lib().og_curl(C.byref(d), C.c_int(which), C.byref(m)), "curl")
    return _take_csc(m, n)


def Mcs(spec, which):
    """Corner-mean operator Mcₑ (which 0) or Mcₘ (1) of create_Mcs (model.jl:287-306)."""
    d = _desc(spec)
    n = 3 * int(np.prod(_N(spec)))
    m = CSC()
    _check(lib().og_Mcs(C.byref(d), C.c_int(which), C.byref(m)), "Mcs")
    return _take_csc(m, n)


def paramop(spec, which, parr):
    d = _desc(spec)
    parr = np.asfortranarray(parr, dtype=np.complex128)
    n = 3 * int(np.prod(_N(spec)))
    m = CSC()
    _check(lib().og_paramop(C.byref(d), C.c_int(which), _ptr(parr), C.byref(m)), "paramop")
    return _take_csc(m, n)


def create_A(spec, omega, eps_arr, mu_arr, ft=0):
    """model.jl:225-246; ft = 0 (EE) or 1 (HH)."""
    d = _desc(spec)
    e = np.asfortranarray(eps_arr, dtype=np.complex128)
    u = np.asfortranarray(mu_arr, dtype=np.complex128)
    n = 3 * int(np.prod(_N(spec)))
    m = CSC()
    om = complex(omega)
    f = lib().og_create_A_hh if ft else lib().og_create_A
    _check(f(C.byref(d), C.c_double(om.real), C.c_double(om.imag), _ptr(e), _ptr(u), C.byref(m)), "create_A")
    return _take_csc(m, n)


def create_b(spec, omega, mu_arr, je, jm, ft=0, eps_arr=None):
    """model.jl:251-274; EE uses mu_arr, HH uses eps_arr."""
    d = _desc(spec)
    p = np.asfortranarray(eps_arr if ft else mu_arr, dtype=np.complex128)
    je = np.ascontiguousarray(je, dtype=np.complex128)
    jm = np.ascontiguousarray(jm, dtype=np.complex128)
    b = np.zeros_like(je)
    om = complex(omega)
    f = lib().og_create_b_hh if ft else lib().og_create_b
    _check(f(C.byref(d), C.c_double(om.real), C.c_double(om.imag), _ptr(p), _ptr(je), _ptr(jm), _ptr(b)),
           "create_b")
    return b


def matparams(spec, shapes, eps_t, mu_t):
    """shapes: list of ('ball', c, r) / ('cuboid', c, s) / ('cylinder', c, r, h);
    eps_t/mu_t: (nobj,3,3)."""
    d = _desc(spec)
    N = _N(spec)
    arr = (Shape * max(len(shapes), 1))()
    for i, s in enumerate(shapes):
        if s[0] == "ball":
            arr[i].kind = 0
            arr[i].r = float(s[2])
        elif s[0] == "cylinder":
            arr[i].kind = 2
            arr[i].r = float(s[2])
            arr[i].s[2] = float(s[3])
        else:
            arr[i].kind = 1
            for a in range(3):
                arr[i].s[a] = float(s[2][a])
        for a in range(3):
            arr[i].c[a] = float(s[1][a])
    et = np.asfortranarray(np.asarray(eps_t, dtype=np.complex128).transpose(1, 2, 0))
    mt = np.asfortranarray(np.asarray(mu_t, dtype=np.complex128).transpose(1, 2, 0))
    ea = np.zeros(N + (3, 3), dtype=np.complex128, order="F")
    ma = np.zeros(N + (3, 3), dtype=np.complex128, order="F")
    _check(lib().og_matparams(C.byref(d), C.c_int64(len(shapes)), arr, _ptr(et), _ptr(mt),
                              _ptr(ea), _ptr(ma)), "matparams")
    return ea, ma


def volfrac(vmin, vmax, n, r0):
    f = lambda v: (C.c_double * 3)(*[float(x) for x in v])
    return lib().og_volfrac(f(vmin), f(vmax), f(n), f(r0))


def kottke(efg, ebg, r, n):
    a = np.asfortranarray(efg, dtype=np.complex128)
    b = np.asfortranarray(ebg, dtype=np.complex128)
    out = np.zeros((3, 3), dtype=np.complex128, order="F")
    nn = (C.c_double * 3)(*[float(x) for x in n])
    lib().og_kottke(_ptr(a), _ptr(b), C.c_double(float(r)), nn, _ptr(out))
    return out


def solve_sqmr(spec, csc, b, tol=1e-8, maxit=100000, nthreads=0):
    """CPU FP64 SQMR; returns (x, iterations, relative residual)."""
    d = _desc(spec)
    colptr, rowval, nzval = csc
    n = len(colptr) - 1
    colptr = np.ascontiguousarray(colptr, dtype=np.int64)
    rowval = np.ascontiguousarray(rowval, dtype=np.int64)
    nzval = np.ascontiguousarray(nzval, dtype=np.complex128)
    m = CSC()
    m.nnz = len(rowval)
    m.colptr = colptr.ctypes.data_as(C.POINTER(C.c_int64))
    m.rowval = rowval.ctypes.data_as(C.POINTER(C.c_int64))
    m.nzval = nzval.ctypes.data
    b = np.ascontiguousarray(b, dtype=np.complex128)
    x = np.zeros(n, dtype=np.complex128)
    it = C.c_int64()
    rr = C.c_double()
    lib().og_solve_sqmr(C.byref(d), C.byref(m), _ptr(b), _ptr(x), C.c_double(tol), C.c_int64(maxit),
                        C.c_int(nthreads), C.byref(it), C.byref(rr))
    return x, it.value, rr.value


def to_scipy(csc):
    import scipy.sparse as sp
    colptr, rowval, nzval = csc
    n = len(colptr) - 1
    return sp.csc_matrix((nzval, rowval - 1, colptr - 1), shape=(n, n))
