import os
import sys
import types

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "maxwellfdm.jl_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


def spec_of(mdl):
    """Oracle grid spec of a maxwellfdfd Model."""
    g = mdl.grid
    return types.SimpleNamespace(lprim=g.lprim, isbloch=g.isbloch, boundft=tuple(mdl.boundft), npml=mdl.Npml,
                                 omega_pml=mdl.ωpml, kbloch=tuple(mdl.kbloch), pml_m=mdl.pml_m,
                                 pml_lnR=mdl.pml_lnR, order_cmpfirst=mdl.order_cmpfirst)


def shapes_of(mdl):
    out, et, mt = [], [], []
    for o in mdl.objs:
        s = o.shape
        if hasattr(s, "h"):
            out.append(("cylinder", tuple(s.c), s.r, s.h))
        elif hasattr(s, "r"):
            out.append(("ball", tuple(s.c), s.r))
        else:
            out.append(("cuboid", tuple(s.c), tuple(s.s)))
        et.append(o.mat.eps)
        mt.append(o.mat.mu)
    return out, np.array(et), np.array(mt)


@pytest.fixture(scope="session")
def pkg():
    import maxwellfdfd
    return maxwellfdfd


@pytest.fixture(scope="session")
def O():
    import oracle
    return oracle


class DeviceBuffers:
    """Device memory for the tests that call the device-pointer entry points,
    through the HIP runtime the library itself loaded (a second HIP runtime
    in the same process, such as torch's, does not see the GPU once the
    library's has started)."""

    def __init__(self, libpath):
        import ctypes as C
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64" in line:
                    path = line.split()[-1]
                    break
        if path is None:
            raise RuntimeError("the HIP runtime is not loaded")
        self.C = C
        self.rt = C.CDLL(path)
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.live = []

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with HIP error {rc}")

    def put(self, a):
        """A device copy of the array a; returns its address."""
        a = np.ascontiguousarray(a)
        p = self.C.c_void_p()
        self._ok(self.rt.hipMalloc(self.C.byref(p), max(a.nbytes, 1)), "hipMalloc")
        self.live.append(p.value)
        self._ok(self.rt.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p.value

    def get(self, ptr, like):
        out = np.empty_like(np.ascontiguousarray(like))
        self._ok(self.rt.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


@pytest.fixture()
def devmem(pkg):
    from maxwellfdfd._lib import context
    context()   # the library and its HIP runtime are loaded, device 0 current
    d = DeviceBuffers(pkg.LIB_PATH)
    yield d
    d.free()
