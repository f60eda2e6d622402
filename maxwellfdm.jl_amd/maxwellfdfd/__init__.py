"""MI355X implementation of MaxwellFDFD's frequency-domain hot path.

Host-side mirror of the reference package's API (src/model/model.jl:23-28)
over libmfdfd's C ABI (include/mfdfd.h).  Import this package after putting
its parent directory (maxwellfdm.jl_amd/) on sys.path.
"""
from ._lib import (MfdArgumentError, MfdError, MfdNoConvergence, MfdUnsupportedError, LIB_PATH, SIGNATURES,
                   lib)
from .core import (DUAL, EE, HH, PRIM, Ball, Box, Cuboid, Grid, MaxwellOperator, Model, ModelFull, PlaneSrc, PointSrc,
                   Sphere, Cylinder, A_csc, add_obj, add_srce, add_srcm, calc_matparams, clear_objs, clear_srcs,
                   create_A, create_b, create_curls, create_Mcs, Mcs_csc, create_eikL, create_field_array, create_linsys,
                   create_param_array, create_paramops, create_srcs, create_stretched_dls, curl_csc,
                   distweights, field_arr2vec, nX, nY, nZ, paramop_csc, set_boundft, set_kbloch, set_Npml,
                   set_omegapml, set_ωpml, solve_batch, solve_linsys, to_scipy, h_from_e, e_from_h)

from . import twod
from . import oned

__all__ = [n for n in dir() if not n.startswith("_")]
