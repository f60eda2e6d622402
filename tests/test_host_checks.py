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


@pytest.mark.parametrize("K", [1, 2])
@pytest.mark.parametrize("windowed", [True, False])
def test_window_order_exchange(K, windowed, pkg):
    """The 2D/1D operator's host-side order exchange: model vectors are
    component-first (c + K·p), the windowed device vectors component-last
    (p + M·k), and the fallback embeds them at 3p + c of the 3D vector; the
    exchange is a bijection that puts every value where the device expects
    it, for one and for several right-hand sides."""
    W = pkg.twod._WindowOperator
    Mc = 7
    ecmp = (0, 1)[:K] if K == 2 else (2,)
    w = object.__new__(W)
    w.K, w.Mc, w.n, w.windowed = K, Mc, K * Mc, windowed
    w._esel = np.concatenate([[c + 3 * p for c in ecmp] for p in range(Mc)]).astype(np.int64)
    rng = np.random.default_rng(5)
    x = rng.normal(size=(3, K * Mc)) + 1j * rng.normal(size=(3, K * Mc))
    d = w._to_dev(x)
    for r in range(3):
        for p in range(Mc):
            for k in range(K):
                if windowed:
                    assert d[r, p + Mc * k] == x[r, k + K * p]
                else:
                    assert d[r, 3 * p + ecmp[k]] == x[r, k + K * p]
    assert np.array_equal(w._from_dev(d, x.shape), x)
    assert np.array_equal(w._from_dev(w._to_dev(x[0]), x[0].shape), x[0])
