"""Host-side argument checks of the Python layer: they run before any call
into libmfdfd, so mismatched buffers never reach the C ABI (which copies
n·nrhs values in and out of the caller's arrays)."""
import numpy as np
import pytest


@pytest.fixture()
def op(pkg):
    # an operator object without a device handle: every check under test
    # raises before the library is called
    o = object.__new__(pkg.MaxwellOperator)
    o.n = 24
    o._h = None
    return o


def test_solve_rejects_bad_lengths(op, pkg):
    b = np.ones(48, complex)
    with pytest.raises(pkg.MfdArgumentError):
        op.solve(b, x0=np.zeros(24, complex))        # short x0: two RHS, one guess
    with pytest.raises(pkg.MfdArgumentError):
        op.solve(np.ones(30, complex))                # not a whole number of DOF vectors
    with pytest.raises(pkg.MfdArgumentError):
        op.solve(np.ones(0, complex))                 # empty


def test_apply_and_field_conversions_reject_bad_lengths(op, pkg):
    with pytest.raises(pkg.MfdArgumentError):
        op.apply(np.ones(25, complex))
    with pytest.raises(pkg.MfdArgumentError):
        op.h_from_e(np.ones(48, complex))
    with pytest.raises(pkg.MfdArgumentError):
        op.h_from_e(np.ones(24, complex), jm=np.ones(12, complex))
    with pytest.raises(pkg.MfdArgumentError):
        op.e_from_h(np.ones(23, complex))
    with pytest.raises(pkg.MfdArgumentError):
        op.e_from_h(np.ones(24, complex), je=np.ones(48, complex))


def test_adjoint_rejects_bad_lengths(op, pkg):
    with pytest.raises(pkg.MfdArgumentError):
        op.solve_adjoint(np.ones(48, complex))
    with pytest.raises(pkg.MfdArgumentError):
        op.adjoint_grad(np.ones(24, complex), np.ones(12, complex))
