"""Function3D mirror: a function = cube on a Grid3D + atom-centred bubbles.

Mirrors the reference's Function3D class (src/bubbles/function3d.F90:86-185)
over the C-ABI: the cube lives on the GPU (a DeviceCube), the bubbles in a
dage Bubbles handle.  Operations:

* operate_on(op, f)            Operator3D_operate_on_Function3D
                               (function3d.F90:3727-3792) / Helmholtz3D_operate
                               (helmholtz3d.F90:446-481): bubble transform +
                               cube transform in one C-ABI call
* f.inject_bubbles()           Function3D_inject_bubbles (:1760-1791)
* f.multiply(g)                Function3D_multiply_sub (:1097-1308), no
                               Taylor-series terms
* f.inject_extra_bubbles(l)    Function3D_inject_extra_bubbles (:2198-2222)
* f.dispose_extra_bubbles(l)   Function3D_dispose_extra_bubbles (:2225-2240)
* f.dump(folder, name)         Function3D_dump (:812-879), .fun format
* Function3D.load(...)         Function3D_load (:881-935)
* write_grid / read_grid       Grid3D_dump / Grid3D_init_read (grid.F90:1072-1145)
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._lib import call
from .operators import Bubbles, Context, DeviceCube, Grid3D

F3D_TYPE_NONE, F3D_TYPE_CUSP, F3D_TYPE_SPOT, F3D_TYPE_NUCP = 0, 1, 2, 3
_dp = ctypes.POINTER(ctypes.c_double)
STORE_ALL, STORE_ONLY_BUBBLES, STORE_ONLY_CUBES = 1, 2, 3


class Function3D:
    def __init__(self, grid: Grid3D, cube: DeviceCube | None = None,
                 bubbles: Bubbles | None = None, type: int = F3D_TYPE_NONE,
                 label: str = "Function"):
        self.grid, self.ctx = grid, grid.ctx
        self.cube = cube if cube is not None else DeviceCube(grid.ctx, grid.get_shape())
        self.bubbles = bubbles
        self.type, self.label = type, label
        self.taylor_order = 2  # Function3D%taylor_order default (function3d.F90:109)

    def close(self):
        for obj in (self.cube, self.bubbles):
            if obj is not None:
                obj.close()
        self.cube = self.bubbles = None

    # ------------------------------------------------------------- arithmetic
    def copy(self) -> "Function3D":
        """Function3D(self, copy_content=.true.)"""
        f = Function3D(self.grid, type=self.type, label=self.label)
        f.cube.copy_from(self.cube)
        if self.bubbles is not None:
            f.bubbles = Bubbles.like(self.bubbles, k=self.bubbles.k)
            f.bubbles.axpby(1.0, self.bubbles, 0.0)
        return f

    def axpby(self, a: float, x: "Function3D", b: float):
        """self = a x + b self, cube and bubbles (add_in_place,
        subtract_in_place, product_in_place_REAL64 of function3d.F90)."""
        n = int(np.prod(self.grid.get_shape()))
        call("dage_vec_axpby", self.ctx.handle, n, float(a), ctypes.c_void_p(x.cube.ptr),
             float(b), ctypes.c_void_p(self.cube.ptr))
        if x.bubbles is not None and self.bubbles is not None:
            self.bubbles.axpby(a, x.bubbles, b)
        elif self.bubbles is not None:
            self.bubbles.axpby(0.0, self.bubbles, b)
        elif x.bubbles is not None:
            self.bubbles = Bubbles.like(x.bubbles, k=x.bubbles.k)
            self.bubbles.axpby(a, x.bubbles, 0.0)

    def scale(self, a: float):
        self.axpby(0.0, self, a)

    def integrate(self) -> float:
        """Function3D_integrate (function3d.F90:1793-1846): the cube's LIP
        integral plus the bubbles' radial integrals."""
        val = self.grid.integrate(self.cube)
        if self.bubbles is not None:
            val += self.bubbles.integrate()
        return val

    def dot(self, f2: "Function3D", taylor: bool | None = None) -> float:
        """Function3D_dot_product (function3d.F90:1497-1530): integral of the
        product (bubbles multiplied centre by centre; `taylor` as in
        multiply)."""
        p = self.multiply(f2, taylor=taylor)
        val = p.integrate()
        p.close()
        return val

    # ---------------------------------------------------------------- bubbles
    def inject_bubbles(self) -> DeviceCube:
        """A new cube holding the bubbles evaluated on the grid (the cube part
        is not included, as in the reference)."""
        out = DeviceCube(self.ctx, self.grid.get_shape())
        out.set_to_zero()
        if self.bubbles is not None:
            self.bubbles.inject(self.grid, out)
        return out

    def multiply(self, f2: "Function3D", taylor: bool | None = None) -> "Function3D":
        """Function3D_multiply_sub (function3d.F90:1097-1308): bubbles
        multiplied centre by centre, cube = (cube2 + bubbles2)(cube1 +
        bubbles1) - product bubbles.  When both factors are cusp-type (CUSP or
        NUCP, as the reference decides in Function3D_multiply_bubbles,
        function3d.F90:1406-1470) or taylor=True, the product bubbles carry
        the Taylor-series terms of both factors (order `taylor_order`):
        (B1 + T1)(B2 + T2) - T1 T2, which keeps the cube free of the
        singular products near the nuclei."""
        out = Function3D(self.grid, label=self.label)
        if F3D_TYPE_CUSP in (self.type, f2.type):
            out.type = F3D_TYPE_CUSP
        both = self.bubbles is not None and f2.bubbles is not None
        cusp = (F3D_TYPE_CUSP, F3D_TYPE_NUCP)
        if taylor is None:
            taylor = self.type in cusp and f2.type in cusp
        if both and taylor:
            t = self.taylor_order
            bo = Bubbles.like(self.bubbles,
                              lmax=max(self.bubbles.lmax, t) + max(f2.bubbles.lmax, t),
                              k=self.bubbles.k + f2.bubbles.k)
            call("dage_function3d_multiply_taylor", self.grid.handle,
                 ctypes.c_void_p(self.cube.ptr), self.bubbles.handle,
                 ctypes.c_void_p(f2.cube.ptr), f2.bubbles.handle, int(t),
                 ctypes.c_void_p(out.cube.ptr), bo.handle)
            out.bubbles = bo
            return out
        bo = Bubbles.like(self.bubbles, lmax=self.bubbles.lmax + f2.bubbles.lmax,
                          k=self.bubbles.k + f2.bubbles.k) if both else None
        call("dage_function3d_multiply", self.grid.handle, ctypes.c_void_p(self.cube.ptr),
             self.bubbles.handle if self.bubbles is not None else None,
             ctypes.c_void_p(f2.cube.ptr), f2.bubbles.handle if f2.bubbles is not None else None,
             ctypes.c_void_p(out.cube.ptr), bo.handle if bo is not None else None)
        out.bubbles = bo
        return out

    def contaminants(self, order: int | None = None) -> np.ndarray:
        """Function3D_get_cube_contaminants + Bubbles_get_contaminants: the
        derivatives d^kappa (cube + other centres' bubbles) at every centre,
        |kappa| <= order, rows in CartIter order, one column per centre."""
        t = self.taylor_order if order is None else order
        nd = (t + 1) * (t + 2) * (t + 3) // 6
        out = np.zeros((nd, self.bubbles.nbub), order="F")
        call("dage_function3d_contaminants", self.grid.handle, ctypes.c_void_p(self.cube.ptr),
             self.bubbles.handle, int(t), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def taylor_series_bubbles(self, order: int | None = None) -> Bubbles:
        """Function3D_get_taylor_series_bubbles (function3d.F90:1338-1404)."""
        t = self.taylor_order if order is None else order
        h = ctypes.c_void_p()
        call("dage_function3d_taylor_bubbles", self.grid.handle, ctypes.c_void_p(self.cube.ptr),
             self.bubbles.handle, int(t), ctypes.byref(h))
        b = Bubbles.__new__(Bubbles)
        b.handle, b.ctx = h, self.ctx
        b.nbub, b.lmax, b.k, b.nlip = (self.bubbles.nbub, max(self.bubbles.lmax, t),
                                       self.bubbles.k, self.bubbles.nlip)
        return b

    def inject_extra_bubbles(self, lmax: int, injected: Bubbles | None = None):
        """Injects the l > lmax part of `injected` (default: own bubbles) into
        the cube and keeps the l <= lmax part as the function's bubbles."""
        src = injected if injected is not None else self.bubbles
        if src is None:
            return
        kept = ctypes.c_void_p()
        call("dage_function3d_inject_extra_bubbles", self.grid.handle,
             ctypes.c_void_p(self.cube.ptr), src.handle, int(lmax), ctypes.byref(kept))
        new = Bubbles.__new__(Bubbles)
        new.handle, new.ctx = kept, self.ctx
        new.nbub, new.lmax, new.k, new.nlip = src.nbub, min(lmax, src.lmax), src.k, src.nlip
        if self.bubbles is not None and self.bubbles is not injected:
            self.bubbles.close()
        self.bubbles = new

    def dispose_extra_bubbles(self, lmax: int):
        """Drops the bubbles' l > lmax part."""
        if self.bubbles is None or self.bubbles.lmax <= lmax:
            return
        kept = ctypes.c_void_p()
        call("dage_bubbles_copy_lmax", self.bubbles.handle, int(lmax), ctypes.byref(kept))
        new = Bubbles.__new__(Bubbles)
        new.handle, new.ctx = kept, self.ctx
        new.nbub, new.lmax, new.k, new.nlip = (self.bubbles.nbub, lmax, self.bubbles.k,
                                               self.bubbles.nlip)
        self.bubbles.close()
        self.bubbles = new

    # ------------------------------------------------------------------ files
    def dump(self, folder: str, funcname: str, mode: int = STORE_ALL) -> str:
        """Writes folder/funcname.fun (the reference's binary format)."""
        name = funcname if funcname.endswith(".fun") else funcname + ".fun"
        path = os.path.join(folder, name)
        shape = (ctypes.c_int * 3)(*self.grid.get_shape())
        cube = None
        if mode != STORE_ONLY_BUBBLES:
            cube = np.asfortranarray(self.cube.download())
        nbub, lmax, k = -1, 0, 0
        charges = centers = npts = ptrs = None
        keep = []
        if mode != STORE_ONLY_CUBES and self.bubbles is not None:
            b = self.bubbles
            nbub, lmax, k = b.nbub, b.lmax, b.k
            charges, centers = np.zeros(nbub), np.zeros(3 * nbub)
            npts = (ctypes.c_int * max(nbub, 1))()
            ptrs = (_dp * max(nbub, 1))()
            for i in range(nbub):
                c, q, _, bf = b.get(i)
                charges[i], centers[3 * i:3 * i + 3] = q, c
                npts[i] = bf.shape[0]
                keep.append(bf)
                ptrs[i] = bf.ctypes.data_as(_dp)
        call("dage_fun_write", path.encode(), self.label.encode()[:20], int(self.type),
             shape if cube is not None else None,
             cube.ctypes.data_as(_dp) if cube is not None else None, nbub, lmax, k,
             charges.ctypes.data_as(_dp) if charges is not None else None,
             centers.ctypes.data_as(_dp) if centers is not None else None, npts, ptrs)
        return path

    @classmethod
    def load(cls, grid: Grid3D, folder: str, filename: str,
             mould: Bubbles | None = None) -> "Function3D":
        """Reads a .fun file onto `grid`; the bubbles' radial grids come from
        `mould` (the file does not hold them)."""
        path = os.path.join(folder, filename)
        label = ctypes.create_string_buffer(21)
        typ, nbub, lmax, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        shape = (ctypes.c_int * 3)()
        call("dage_fun_read_info", path.encode(), label, ctypes.byref(typ), shape,
             ctypes.byref(nbub), ctypes.byref(lmax), ctypes.byref(k))
        f = cls(grid, type=typ.value, label=label.value.decode().rstrip())
        if shape[0] >= 0:
            if tuple(shape) != tuple(grid.get_shape()):
                raise ValueError(f"cube shape {tuple(shape)} != grid {grid.get_shape()}")
            cube = np.zeros(tuple(shape), order="F")
            call("dage_fun_read", path.encode(), cube.ctypes.data_as(_dp), None, None, None, None)
            f.cube.upload(cube)
        if nbub.value >= 0:
            if mould is None or mould.nbub != nbub.value:
                raise ValueError("a mould with the file's bubble count is needed")
            nb, L = nbub.value, lmax.value
            steps = [mould.get(i)[2] for i in range(nb)]
            npts = (ctypes.c_int * max(nb, 1))(*[len(s) * (mould.nlip - 1) + 1 for s in steps])
            bfs = [np.zeros((npts[i], (L + 1) ** 2), order="F") for i in range(nb)]
            ptrs = (_dp * max(nb, 1))(*[b.ctypes.data_as(_dp) for b in bfs])
            charges, centers = np.zeros(max(nb, 1)), np.zeros(3 * max(nb, 1))
            call("dage_fun_read", path.encode(), None, charges.ctypes.data_as(_dp),
                 centers.ctypes.data_as(_dp), npts, ptrs)
            bub = Bubbles(grid.ctx, nb, L, k.value, mould.nlip)
            for i in range(nb):
                bub.set_bubble(i, centers[3 * i:3 * i + 3], charges[i], steps[i], bfs[i])
            f.bubbles = bub
        return f


def operate_on(op, func: Function3D, only_cube: bool = False,
               out: Function3D | None = None) -> Function3D:
    """Operator3D%operate_on: the bubble transform of the operator (Coulomb:
    radial Poisson; Helmholtz after set_energy: Bessel convolution) and the
    cube transform, one C-ABI call.  The result's bubbles have k = 0."""
    grid_out = op.gridout
    if out is None:
        out = Function3D(grid_out, label=(func.label + ".copy")[:20])
    bub_out = None
    if func.bubbles is not None and not only_cube:
        bub_out = Bubbles.like(func.bubbles, k=0)
    call("dage_operator_operate_on", op.handle, ctypes.c_void_p(func.cube.ptr),
         func.bubbles.handle if (func.bubbles is not None and not only_cube) else None,
         ctypes.c_void_p(out.cube.ptr), bub_out.handle if bub_out is not None else None,
         1 if only_cube else 0)
    if bub_out is not None:
        if out.bubbles is not None:
            out.bubbles.close()
        out.bubbles = bub_out
    if hasattr(op, "energy"):  # Helmholtz3D_operate leaves the operator at energy 0
        op.energy = 0.0
    return out


def write_grid(grid: Grid3D, path: str):
    """Grid3D_dump (grid.F90:1128-1145): the cube grid as a .g3d file."""
    steps = [np.ascontiguousarray(grid.steps[a], dtype=np.float64) for a in range(3)]
    ncell = (ctypes.c_int * 3)(*[len(s) for s in steps])
    qmin = np.ascontiguousarray(grid.qmin, dtype=np.float64)
    ptrs = (_dp * 3)(*[s.ctypes.data_as(_dp) for s in steps])
    call("dage_g3d_write", path.encode(), grid.nlip, grid.grid_type, ncell,
         qmin.ctypes.data_as(_dp), ptrs)


def read_grid(ctx: Context, path: str) -> Grid3D:
    """Grid3D_init_read (grid.F90:1072-1126)."""
    nlip, gt = ctypes.c_int(), ctypes.c_int()
    ncell = (ctypes.c_int * 3)()
    call("dage_g3d_read", path.encode(), ctypes.byref(nlip), ctypes.byref(gt), ncell, None, None)
    steps = [np.zeros(ncell[a]) for a in range(3)]
    qmin = np.zeros(3)
    ptrs = (_dp * 3)(*[s.ctypes.data_as(_dp) for s in steps])
    call("dage_g3d_read", path.encode(), None, None, None, qmin.ctypes.data_as(_dp), ptrs)
    return Grid3D(ctx, qmin, steps, nlip.value, grid_type=gt.value)
