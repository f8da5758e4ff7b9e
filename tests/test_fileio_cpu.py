"""The binary file formats (row f3) against files a Fortran compiler wrote.

tests/golden/fixture.fun and fixture.g3d were written by amdflang from
tests/golden/fun_fixture.F90 with unformatted stream I/O in the layout of
Function3D_dump / Bubbles_write / Grid3D_dump (src/bubbles/function3d.F90:
812-879, bubbles.F90:1251-1276, grid.F90:1128-1145).  The library's writer
must reproduce them byte for byte and its reader must return their values.
Host code only (no GPU).
"""
import ctypes
import os

import numpy as np

from dage_amd import _lib

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_dp = ctypes.POINTER(ctypes.c_double)


def fixture_values():
    i, j, k = np.meshgrid(np.arange(1, 4), np.arange(1, 5), np.arange(1, 6), indexing="ij")
    cube = np.asfortranarray(0.25 * i - 1.5 * j + 0.125 * k * k)
    charges = np.array([3.0, 1.0])
    centers = np.array([0.5, -0.25, 1.0, -1.0, 2.0, 0.0])
    ii, jj = np.meshgrid(np.arange(1, 14), np.arange(1, 5), indexing="ij")
    bfs = [np.asfortranarray(ib * 100.0 + jj * 10.0 + ii * 0.5) for ib in (1, 2)]
    return cube, charges, centers, bfs


def test_fun_writer_matches_fortran_bytes(tmp_path):
    lib = _lib.load()
    cube, charges, centers, bfs = fixture_values()
    path = str(tmp_path / "out.fun")
    shape = (ctypes.c_int * 3)(3, 4, 5)
    npts = (ctypes.c_int * 2)(13, 13)
    ptrs = (_dp * 2)(*[b.ctypes.data_as(_dp) for b in bfs])
    _lib.call("dage_fun_write", path.encode(), b"fixture", 1, shape, cube.ctypes.data_as(_dp), 2,
              1, -1, charges.ctypes.data_as(_dp), centers.ctypes.data_as(_dp), npts, ptrs)
    with open(path, "rb") as f, open(os.path.join(HERE, "fixture.fun"), "rb") as g:
        assert f.read() == g.read()
    del lib


def test_fun_reader_returns_fortran_values():
    path = os.path.join(HERE, "fixture.fun").encode()
    label = ctypes.create_string_buffer(21)
    typ, nb, lmax, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    shape = (ctypes.c_int * 3)()
    _lib.call("dage_fun_read_info", path, label, ctypes.byref(typ), shape, ctypes.byref(nb),
              ctypes.byref(lmax), ctypes.byref(k))
    assert label.value == b"fixture             " and typ.value == 1
    assert tuple(shape) == (3, 4, 5) and (nb.value, lmax.value, k.value) == (2, 1, -1)
    cube = np.zeros((3, 4, 5), order="F")
    charges, centers = np.zeros(2), np.zeros(6)
    bfs = [np.zeros((13, 4), order="F") for _ in range(2)]
    npts = (ctypes.c_int * 2)(13, 13)
    ptrs = (_dp * 2)(*[b.ctypes.data_as(_dp) for b in bfs])
    _lib.call("dage_fun_read", path, cube.ctypes.data_as(_dp), charges.ctypes.data_as(_dp),
              centers.ctypes.data_as(_dp), npts, ptrs)
    rc, rq, rcen, rbf = fixture_values()
    assert np.array_equal(cube, rc) and np.array_equal(charges, rq)
    assert np.array_equal(centers, rcen)
    assert all(np.array_equal(a, b) for a, b in zip(bfs, rbf))


def test_g3d_roundtrip_matches_fortran(tmp_path):
    steps = [np.array([0.1, 0.2]), np.array([0.3, 0.15]), np.array([0.05, 0.5])]
    qmin = np.array([-1.0, -2.0, -3.0])
    path = str(tmp_path / "g.g3d")
    ptrs = (_dp * 3)(*[s.ctypes.data_as(_dp) for s in steps])
    _lib.call("dage_g3d_write", path.encode(), 7, 1, (ctypes.c_int * 3)(2, 2, 2),
              qmin.ctypes.data_as(_dp), ptrs)
    with open(path, "rb") as f, open(os.path.join(HERE, "fixture.g3d"), "rb") as g:
        assert f.read() == g.read()
    nlip, gt = ctypes.c_int(), ctypes.c_int()
    ncell = (ctypes.c_int * 3)()
    got = [np.zeros(2) for _ in range(3)]
    q = np.zeros(3)
    gptr = (_dp * 3)(*[s.ctypes.data_as(_dp) for s in got])
    _lib.call("dage_g3d_read", os.path.join(HERE, "fixture.g3d").encode(), ctypes.byref(nlip),
              ctypes.byref(gt), ncell, q.ctypes.data_as(_dp), gptr)
    assert (nlip.value, gt.value, tuple(ncell)) == (7, 1, (2, 2, 2))
    assert np.array_equal(q, qmin) and all(np.array_equal(a, b) for a, b in zip(got, steps))
