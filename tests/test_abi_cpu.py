"""C-ABI checks that need no GPU: the library loads, exports every entry point
include/dage_ops.h declares, and its host-only numerics (the t-quadrature)
equal the oracle's bit for bit."""
import ctypes

import numpy as np
import pytest

import dage_amd
import oracle as O
from dage_amd import _lib


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _lib.header_symbols()
    assert len(declared) >= 25
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    # and the Python binding covers the whole header
    assert sorted(_lib._SIGNATURES) == declared


def test_abi_version():
    assert _lib.load().dage_abi_version() == 1


@pytest.mark.parametrize("args", [(10, 8, 0.0, 2.0, 1000.0), (20, 20, 0.0, 2.0, 500.0),
                                  (25, 25, 0.0, 2.0, 500.0), (7, 3, 0.5, 3.0, 80.0)])
def test_gauss_quad_bitwise_equal_to_oracle(args):
    q = dage_amd.GaussQuad(nlin=args[0], nlog=args[1], tstart=args[2], tlin=args[3], tlog=args[4])
    tp, tw = O.gauss_quad(*args)
    assert np.array_equal(q.tp, tp) and np.array_equal(q.tw, tw)


def test_gauss_quad_rejects_bad_intervals():
    with pytest.raises(dage_amd.DageError) as e:
        dage_amd.GaussQuad(nlin=10, nlog=8, tlin=2.0, tlog=1.0)
    assert e.value.code == _lib.DAGE_ERR_ARG
    assert "tlin" in str(e.value)


def test_null_handles_are_rejected():
    lib = _lib.load()
    assert lib.dage_operator_apply(None, None, None, 0) == _lib.DAGE_ERR_ARG
    assert b"null" in lib.dage_last_error()
    assert lib.dage_ctx_synchronize(None) == _lib.DAGE_ERR_ARG
    # destroy of a null handle is a no-op
    assert lib.dage_operator_destroy(None) == 0
