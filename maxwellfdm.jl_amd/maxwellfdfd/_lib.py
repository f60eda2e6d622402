"""ctypes binding of libmfdfd (include/mfdfd.h).

The library is required: importing the bindings succeeds without it, but
every call that needs it raises RuntimeError naming the missing file.  There
is no CPU fallback for any device entry point.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "lib", "libmfdfd.so"))

MFD_OK, MFD_ERR_ARG, MFD_ERR_UNSUPPORTED, MFD_ERR_HIP, MFD_ERR_NOCONV, MFD_ERR_ALLOC = range(6)
MFD_FP64, MFD_MIXED = 0, 1
MFD_SQMR, MFD_COCG, MFD_BICGSTAB = 0, 1, 2


class c64(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class mfd_grid(C.Structure):
    _fields_ = [("N", C.c_int64 * 3), ("lprim", C.c_void_p * 3), ("isbloch", C.c_int32 * 3),
                ("boundft", C.c_int32 * 3), ("npml", (C.c_int64 * 3) * 2), ("omega_pml", c64),
                ("kbloch", C.c_double * 3), ("pml_m", C.c_double), ("pml_lnR", C.c_double),
                ("order_cmpfirst", C.c_int32)]


class mfd_csc(C.Structure):
    """A Julia SparseMatrixCSC{ComplexF64,Int64} as it lies in memory (1-based)."""
    _fields_ = [("m", C.c_int64), ("n", C.c_int64), ("nnz", C.c_int64), ("colptr", C.c_void_p),
                ("rowval", C.c_void_p), ("nzval", C.c_void_p)]


class mfd_shape(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("c", C.c_double * 3), ("r", C.c_double),
                ("s", C.c_double * 3)]


class mfd_solve_opts(C.Structure):
    _fields_ = [("tol", C.c_double), ("maxit", C.c_int64), ("precision", C.c_int32), ("method", C.c_int32),
                ("inner_tol", C.c_double), ("check_every", C.c_int32), ("use_graph", C.c_int32)]


class mfd_solve_stats(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("outer_cycles", C.c_int64), ("relres", C.c_double),
                ("seconds", C.c_double), ("converged", C.c_int32), ("iterations_fp32", C.c_int64)]


class MfdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class MfdArgumentError(MfdError, ValueError):
    """Raised where the reference throws ArgumentError."""


class MfdUnsupportedError(MfdError, NotImplementedError):
    pass


class MfdNoConvergence(MfdError):
    pass


_lib = None

# every exported symbol with its (restype, argtypes)
_P = C.c_void_p
_I64P = C.POINTER(C.c_int64)
SIGNATURES = {
    "mfd_version": (C.c_char_p, []),
    "mfd_last_error": (C.c_char_p, []),
    "mfd_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "mfd_ctx_destroy": (C.c_int, [_P]),
    "mfd_stretched_dl": (C.c_int, [_P, C.POINTER(mfd_grid), _P]),
    "mfd_distweights": (C.c_int, [C.c_double, C.c_int, C.POINTER(C.c_double), _P, _P, C.c_int64, C.c_int,
                                  _I64P, C.POINTER(C.c_double)]),
    "mfd_add_pointsrc": (C.c_int, [C.POINTER(mfd_grid), C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                   c64, _P]),
    "mfd_add_planesrc": (C.c_int, [C.POINTER(mfd_grid), C.c_int, C.c_int, C.c_double, C.POINTER(C.c_double),
                                   c64, _P]),
    "mfd_curl_nnz": (C.c_int, [_P, C.POINTER(mfd_grid), C.c_int, _I64P]),
    "mfd_curl_csc": (C.c_int, [_P, C.POINTER(mfd_grid), C.c_int, _P, _P, _P]),
    "mfd_Mcs_nnz": (C.c_int, [_P, C.POINTER(mfd_grid), C.c_int, _I64P]),
    "mfd_Mcs_csc": (C.c_int, [_P, C.POINTER(mfd_grid), C.c_int, _P, _P, _P]),
    "mfd_matparams": (C.c_int, [_P, C.POINTER(mfd_grid), C.c_int64, _P, _P, _P, _P, _P]),
    "mfd_paramop_nnz": (C.c_int, [_P, C.POINTER(mfd_grid), C.c_int, _P, _I64P]),
    "mfd_paramop_csc": (C.c_int, [_P, C.POINTER(mfd_grid), C.c_int, _P, _P, _P, _P]),
    "mfd_A_nnz": (C.c_int, [_P, C.POINTER(mfd_grid), C.c_int, c64, _P, _P, _I64P]),
    "mfd_A_csc": (C.c_int, [_P, C.POINTER(mfd_grid), C.c_int, c64, _P, _P, _P, _P, _P]),
    "mfd_create_b": (C.c_int, [_P, C.POINTER(mfd_grid), C.c_int, c64, _P, _P, _P, _P, _P]),
    "mfd_op_create": (C.c_int, [_P, C.POINTER(mfd_grid), C.c_int, c64, _P, _P, C.POINTER(_P)]),
    "mfd_op_create_csc": (C.c_int, [_P, C.POINTER(mfd_csc), C.POINTER(_P)]),
    "mfd_A_from_csc_nnz": (C.c_int, [_P, C.c_int, c64] + [C.POINTER(mfd_csc)] * 4 + [_I64P]),
    "mfd_A_from_csc": (C.c_int, [_P, C.c_int, c64] + [C.POINTER(mfd_csc)] * 4 + [_P, _P, _P]),
    "mfd_b_from_csc": (C.c_int, [_P, C.c_int, c64] + [C.POINTER(mfd_csc)] * 4 + [_P, _P, _P]),
    "mfd_h_from_e_csc": (C.c_int, [_P, _P, c64, C.POINTER(mfd_csc), C.POINTER(mfd_csc), _P, _P]),
    "mfd_e_from_h_csc": (C.c_int, [_P, _P, c64, C.POINTER(mfd_csc), C.POINTER(mfd_csc), _P, _P]),
    "mfd_op_destroy": (C.c_int, [_P]),
    "mfd_op_apply": (C.c_int, [_P, _P, _P, C.c_int64]),
    "mfd_op_apply_dev": (C.c_int, [_P, _P, _P, C.c_int64, C.c_int]),
    "mfd_op_set_graddiv": (C.c_int, [_P, C.c_int]),
    "mfd_op_set_components": (C.c_int, [_P, C.c_int, C.c_int]),
    "mfd_op_set_tile": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int]),
    "mfd_op_autotune": (C.c_int, [_P, C.c_int64, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_double)]),
    "mfd_op_bench_apply": (C.c_int, [_P, C.c_int64, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "mfd_op_bench_iteration": (C.c_int, [_P, C.c_int64, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "mfd_solve_opts_default": (None, [C.POINTER(mfd_solve_opts)]),
    "mfd_solve": (C.c_int, [_P, _P, _P, C.c_int64, C.POINTER(mfd_solve_opts), C.POINTER(mfd_solve_stats)]),
    "mfd_solve_dev": (C.c_int, [_P, _P, _P, C.c_int64, C.POINTER(mfd_solve_opts), C.POINTER(mfd_solve_stats)]),
    "mfd_solve_dev_noalloc": (C.c_int, [_P, _P, _P, C.c_int64, C.POINTER(mfd_solve_opts),
                                        C.POINTER(mfd_solve_stats)]),
    "mfd_solve_batch": (C.c_int, [C.c_int, _P, _P, _P, C.POINTER(mfd_solve_opts), _P]),
    "mfd_solve_batch_dev": (C.c_int, [C.c_int, _P, _P, _P, C.POINTER(mfd_solve_opts), _P]),
    "mfd_h_from_e": (C.c_int, [_P, _P, _P, _P]),
    "mfd_e_from_h": (C.c_int, [_P, _P, _P, _P]),
    "mfd_solve_adjoint": (C.c_int, [_P, _P, _P, C.POINTER(mfd_solve_opts), C.POINTER(mfd_solve_stats)]),
    "mfd_adjoint_grad": (C.c_int, [_P, _P, _P, _P]),
}


def lib():
    """Load libmfdfd.so (raises RuntimeError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libmfdfd.so not found at {LIB_PATH}; run __graft_entry__.build() "
                               "(make -C maxwellfdm.jl_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc == MFD_OK:
        return
    msg = lib().mfd_last_error().decode("utf-8", "replace")
    if rc == MFD_ERR_ARG:
        raise MfdArgumentError(rc, msg)
    if rc == MFD_ERR_UNSUPPORTED:
        raise MfdUnsupportedError(rc, msg)
    if rc == MFD_ERR_NOCONV:
        raise MfdNoConvergence(rc, msg)
    raise MfdError(rc, msg)


_ctx = None


def context():
    """The process-wide context on device 0 (one HIP stream)."""
    global _ctx
    if _ctx is None:
        h = C.c_void_p()
        check(lib().mfd_ctx_create(0, C.byref(h)))
        _ctx = h
    return _ctx


def ptr(a):
    return C.c_void_p(a.ctypes.data)


def cplx(z):
    z = complex(z)
    return c64(z.real, z.imag)


def carr(a, order="C"):
    """complex128 array in the requested memory order (no copy when it already is)."""
    return np.require(a, dtype=np.complex128, requirements=[order])
