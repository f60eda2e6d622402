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

