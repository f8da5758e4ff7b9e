"""Host-side mirror of DAGE's Operator3D API over the gfx950 C-ABI.

Names and argument meaning follow the reference:
  GaussQuad(nlin, nlog, tlin, tlog, tstart)     src/bubbles/gauss.F90:81-122
  Grid3D.step / Grid3D.equidistant              src/bubbles/grid.F90:652-751
  Coulomb3D(gridin, gridout, gauss)             src/bubbles/coulomb.F90:109-195
  Operator3D.transform_cube / .apply            src/bubbles/function3d.F90:4230-4349
  Operator3D.transform_cuda_cube                src/bubbles/function3d.F90:4354-4480
  Operator3D.set_transformation_weights         src/bubbles/function3d.F90:320-336
  Helmholtz3D.set_energy                        src/bubbles/helmholtz3d.F90:411-444
  GBFMMCoulomb3D(grid, maxlevel, lmax, gauss)   src/gbfmm/gbfmm_coulomb3d.F90:306-506
  InteractionMatrix / TranslationMatrix         src/bubbles/multipole.F90:33-560
  HelmholtzBatch (per-orbital set_energy+apply)  src/bubbles/scf_cycle.F90:2927-3054

Cubes are numpy arrays in Fortran order (x fastest) for the host path and
DeviceCube objects (the library's CudaCube, same layout) for the device path.
Errors raise DageError with the library's message, in place of the
reference's print-and-exit.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import DAGE_OP_NO_CODA, DAGE_PREC_FP64, call

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


class Context:
    """One HIP device and one stream (replaces the reference's StreamContainer)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        call("dage_ctx_create", device, ctypes.byref(h))
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            _lib.load().dage_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_ptr: int | None):
        call("dage_ctx_set_stream", self.handle, ctypes.c_void_p(stream_ptr or 0))

    @property
    def stream(self) -> int:
        s = ctypes.c_void_p()
        call("dage_ctx_get_stream", self.handle, ctypes.byref(s))
        return s.value or 0

    def synchronize(self):
        call("dage_ctx_synchronize", self.handle)

    def profile(self, enable: bool):
        call("dage_ctx_profile", self.handle, 1 if enable else 0)

    def profile_query(self, name: str):
        ms, n = ctypes.c_double(), ctypes.c_int()
        call("dage_ctx_profile_query", self.handle, name.encode(), ctypes.byref(ms), ctypes.byref(n))
        return ms.value, n.value

    def time_begin(self):
        call("dage_ctx_time_begin", self.handle)

    def time_end(self) -> float:
        ms = ctypes.c_double()
        call("dage_ctx_time_end", self.handle, ctypes.byref(ms))
        return ms.value


class DeviceCube:
    """Device-resident cube (CudaCube, src/bubbles/cube.cu:1182-1269)."""

    def __init__(self, ctx: Context, shape, host=None):
        self.shape = tuple(int(s) for s in shape)
        sh = (ctypes.c_int * 3)(*self.shape)
        h = ctypes.c_void_p()
        call("dage_cube_create", ctx.handle, sh, ctypes.byref(h))
        self.handle = h
        self.ctx = ctx
        if host is not None:
            self.upload(host)

    def close(self):
        if self.handle:
            _lib.load().dage_cube_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ptr(self) -> int:
        p = ctypes.c_void_p()
        call("dage_cube_device_pointer", self.handle, ctypes.byref(p))
        return p.value

    def upload(self, host):
        a = np.asfortranarray(host, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"host cube shape {a.shape} != {self.shape}")
        call("dage_cube_upload", self.handle, a.ctypes.data_as(_dp))

    def download(self) -> np.ndarray:
        out = np.zeros(self.shape, order="F")
        call("dage_cube_download", self.handle, out.ctypes.data_as(_dp))
        return out

    def set_to_zero(self):
        call("dage_cube_set_to_zero", self.handle)

    def copy_from(self, src: "DeviceCube"):
        call("dage_cube_copy", self.handle, src.handle)

    def view(self, offset, shape) -> "CubeView":
        """The sub-cube of `shape` starting at element `offset` (0-based),
        with this cube's leading dimensions (CudaCube_get_subcube,
        src/bubbles/cube.cu:1042-1177)."""
        return CubeView(self, offset, shape)


class CubeView:
    """A box inside a DeviceCube: device pointer of its first element and the
    parent's leading dimensions (ldx, ldy)."""

    def __init__(self, parent: DeviceCube, offset, shape):
        off = tuple(int(o) for o in offset)
        shp = tuple(int(s) for s in shape)
        if len(off) != 3 or len(shp) != 3:
            raise ValueError("offset and shape need three entries")
        for a in range(3):
            if off[a] < 0 or shp[a] < 1 or off[a] + shp[a] > parent.shape[a]:
                raise ValueError(f"box {off}+{shp} outside the cube {parent.shape}")
        self.parent, self.offset, self.shape = parent, off, shp
        self.ld = (parent.shape[0], parent.shape[1])

    @property
    def ptr(self) -> int:
        lx, ly = self.ld
        o = self.offset
        return self.parent.ptr + 8 * (o[0] + lx * (o[1] + ly * o[2]))


class GaussQuad:
    """t-quadrature: nlin Gauss-Legendre points on [tstart, tlin] plus nlog on a
    logarithmic map of [tlin, tlog].  Defaults as the reference (10, 8, 2, 1000)."""

    def __init__(self, nlin=10, nlog=8, tlin=2.0, tlog=1.0e3, tstart=0.0):
        n = nlin + nlog
        self.tp = np.zeros(n)
        self.tw = np.zeros(n)
        self.tlog = float(tlog)
        call("dage_gauss_quad", nlin, nlog, float(tstart), float(tlin), float(tlog), _ptr(self.tp),
             _ptr(self.tw))

    def get_tpoints(self):
        return self.tp.copy()

    def get_weights(self):
        return self.tw.copy()

    def get_npoints(self):
        return len(self.tp)

    def get_tlog(self):
        return self.tlog


class Grid3D:
    """Cell grid of LIP cells (nlip points per cell, shared end points)."""

    def __init__(self, ctx: Context, qmin, steps, nlip: int, grid_type: int = 1):
        steps = [_f64(s) for s in steps]
        qmin = _f64(qmin)
        ncell = (ctypes.c_int * 3)(*[len(s) for s in steps])
        h = ctypes.c_void_p()
        call("dage_grid3d_create", ctx.handle, _ptr(qmin), ncell, nlip, _ptr(steps[0]),
             _ptr(steps[1]), _ptr(steps[2]), grid_type, ctypes.byref(h))
        self.handle = h
        self.ctx = ctx
        self.nlip = nlip
        self.grid_type = grid_type
        self._steps = steps
        self.steps = steps
        self.qmin = qmin

    @classmethod
    def equidistant(cls, ctx, ranges, nlip=7, stepmax=0.2, grid_type=1):
        """Grid3D_init_equidistant: ranges = [(xmin,xmax),(ymin,ymax),(zmin,zmax)]."""
        steps, qmin = [], []
        for r0, r1 in ranges:
            nc = int(np.ceil((r1 - r0) / (stepmax * (nlip - 1))))
            steps.append(np.full(nc, (r1 - r0) / (nc * (nlip - 1))))
            qmin.append(r0)
        return cls(ctx, qmin, steps, nlip, grid_type)

    @classmethod
    def ncell(cls, ctx, ranges, ncell, nlip=7, grid_type=1):
        """Grid3D_init_equidistant_ncell (grid.F90:753-808)."""
        steps = [np.full(n, (r1 - r0) / (n * (nlip - 1))) for (r0, r1), n in zip(ranges, ncell)]
        return cls(ctx, [r[0] for r in ranges], steps, nlip, grid_type)

    def close(self):
        if self.handle:
            _lib.load().dage_grid3d_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_shape(self):
        s = (ctypes.c_int * 3)()
        call("dage_grid3d_shape", self.handle, s)
        return tuple(s)

    def get_coord(self, axis: int) -> np.ndarray:
        out = np.zeros(self.get_shape()[axis])
        call("dage_grid3d_coord", self.handle, axis, _ptr(out))
        return out

    def get_ints(self, axis: int) -> np.ndarray:
        out = np.zeros(self.get_shape()[axis])
        call("dage_grid3d_weights", self.handle, axis, _ptr(out))
        return out

    def integrate(self, f: DeviceCube, g: DeviceCube | None = None) -> float:
        """sum w_x w_y w_z f [g] over the cube (FP64 on the GPU)."""
        shape = self.get_shape()
        for c in (f, g):
            if c is not None and c.shape != shape:
                raise ValueError(f"cube shape {c.shape} != grid shape {shape}")
        r = ctypes.c_double()
        call("dage_cube_integrate_device", self.handle, ctypes.c_void_p(f.ptr),
             ctypes.c_void_p(g.ptr if g is not None else 0), ctypes.byref(r))
        return r.value


class Operator3D:
    """Tensor-quadrature operator: sum_t w_t Fx_t (x) Fy_t (x) Fz_t  [+ coda]."""

    def __init__(self, handle, ctx, gridin, gridout, ntp):
        self.handle = handle
        self.ctx = ctx
        self.gridin = gridin
        self.gridout = gridout
        self.ntp = ntp

    def close(self):
        if self.handle:
            _lib.load().dage_operator_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_dims(self):
        return self.gridin.get_shape(), self.gridout.get_shape()

    def set_transformation_weights(self, w):
        w = _f64(w)
        if len(w) != self.ntp:
            raise ValueError(f"expected {self.ntp} weights, got {len(w)}")
        call("dage_operator_set_weights", self.handle, _ptr(w))

    def get_weights(self) -> np.ndarray:
        w = np.zeros(self.ntp)
        call("dage_operator_get_weights", self.handle, _ptr(w))
        return w

    def get_matrices(self):
        (nxi, nyi, nzi), (nxo, nyo, nzo) = self.get_dims()
        T = self.ntp
        fx = np.zeros((nxo, nxi, T), order="F")
        fy = np.zeros((nyi, nyo, T), order="F")
        fz = np.zeros((nzi, nzo, T), order="F")
        call("dage_operator_get_matrices", self.handle, fx.ctypes.data_as(_dp),
             fy.ctypes.data_as(_dp), fz.ctypes.data_as(_dp))
        return fx, fy, fz

    def set_matrices(self, fx, fy, fz):
        (nxi, nyi, nzi), (nxo, nyo, nzo) = self.get_dims()
        T = self.ntp
        arrs = []
        for a, shape in ((fx, (nxo, nxi, T)), (fy, (nyi, nyo, T)), (fz, (nzi, nzo, T))):
            a = np.asfortranarray(a, dtype=np.float64)
            if a.shape != shape:
                raise ValueError(f"matrix shape {a.shape} != {shape}")
            arrs.append(a)
        call("dage_operator_set_matrices", self.handle, *[a.ctypes.data_as(_dp) for a in arrs])

    def transform_cube(self, cubein, precision: int = DAGE_PREC_FP64) -> np.ndarray:
        """Host cube in, host cube out (uploads, applies, downloads)."""
        nin, nout = self.get_dims()
        cubein = np.asfortranarray(cubein, dtype=np.float64)
        if cubein.shape != tuple(nin):
            raise ValueError(f"cube shape {cubein.shape} != grid shape {nin}")
        out = np.zeros(nout, order="F")
        call("dage_operator_apply", self.handle, cubein.ctypes.data_as(_dp),
             out.ctypes.data_as(_dp), precision)
        return out

    apply = transform_cube

    def transform_cuda_cube(self, cubein: DeviceCube, cubeout: DeviceCube,
                            precision: int = DAGE_PREC_FP64):
        """Device cubes in and out; asynchronous on the context stream."""
        nin, nout = self.get_dims()
        if cubein.shape != tuple(nin) or cubeout.shape != tuple(nout):
            raise ValueError(f"device cubes {cubein.shape}->{cubeout.shape} do not match "
                             f"the operator {tuple(nin)}->{tuple(nout)}")
        call("dage_operator_apply_device", self.handle, ctypes.c_void_p(cubein.ptr),
             ctypes.c_void_p(cubeout.ptr), precision)

    def transform_cuda_cube_strided(self, cubein, cubeout, precision: int = DAGE_PREC_FP64):
        """transform_cuda_cube on sub-cubes (CubeView or DeviceCube) with
        their own leading dimensions."""
        nin, nout = self.get_dims()
        if cubein.shape != tuple(nin) or cubeout.shape != tuple(nout):
            raise ValueError(f"device boxes {cubein.shape}->{cubeout.shape} do not match "
                             f"the operator {tuple(nin)}->{tuple(nout)}")
        lin = getattr(cubein, "ld", cubein.shape[:2])
        lout = getattr(cubeout, "ld", cubeout.shape[:2])
        call("dage_operator_apply_device_strided", self.handle, ctypes.c_void_p(cubein.ptr),
             int(lin[0]), int(lin[1]), ctypes.c_void_p(cubeout.ptr), int(lout[0]),
             int(lout[1]), precision)

    def transform_cuda_cube_noalloc(self, cubein: DeviceCube, cubeout: DeviceCube):
        """FP64 transform_cuda_cube that allocates nothing: the intermediates
        live in the library's fixed per-GPU scratch (noalloc_limit() bytes,
        shared by every such call on the GPU, so overlapping calls race)."""
        nin, nout = self.get_dims()
        if cubein.shape != tuple(nin) or cubeout.shape != tuple(nout):
            raise ValueError(f"device cubes {cubein.shape}->{cubeout.shape} do not match "
                             f"the operator {tuple(nin)}->{tuple(nout)}")
        call("dage_operator_apply_device_noalloc", self.handle, ctypes.c_void_p(cubein.ptr),
             ctypes.c_void_p(cubeout.ptr))

    @staticmethod
    def noalloc_limit() -> int:
        n = ctypes.c_size_t()
        call("dage_operator_noalloc_limit", ctypes.byref(n))
        return n.value

    def workspace_bytes(self) -> int:
        n = ctypes.c_size_t()
        call("dage_operator_workspace_bytes", self.handle, ctypes.byref(n))
        return n.value

    def tiling(self) -> dict:
        """Column tiling of the FP64 passes (x, y, z): column blocks per tile,
        tile pitch (0: 16*nb columns), whether zero K panels are skipped, the
        first t-point of the banded passes (T: none), and the number of
        low-rank t-points with the sum of their x ranks."""
        nb, pitch, skip = (ctypes.c_int * 3)(), (ctypes.c_int * 3)(), (ctypes.c_int * 3)()
        call("dage_operator_get_tiling", self.handle, nb, pitch, skip)
        t0, t1, kx = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        call("dage_operator_get_band", self.handle, ctypes.byref(t0))
        call("dage_operator_get_lowrank", self.handle, ctypes.byref(t1), ctypes.byref(kx))
        return {"nb": list(nb), "pitch": list(pitch), "skip": [bool(v) for v in skip],
                "band_t0": t0.value, "lowrank_t1": t1.value, "lowrank_kx": kx.value}


class Coulomb3D(Operator3D):
    """Coulomb operator 1/r (Coulomb3D_init_explicit_sub, coulomb.F90:109-176)."""

    def __init__(self, ctx: Context, gridin: Grid3D, gridout: Grid3D | None = None,
                 gauss: GaussQuad | None = None, coda: bool = True):
        gridout = gridout or gridin
        gauss = gauss or GaussQuad()
        h = ctypes.c_void_p()
        flags = 0 if coda else DAGE_OP_NO_CODA
        call("dage_coulomb3d_create", ctx.handle, gridin.handle, gridout.handle, _ptr(gauss.tp),
             _ptr(gauss.tw), len(gauss.tp), gauss.tlog, flags, ctypes.byref(h))
        super().__init__(h, ctx, gridin, gridout, len(gauss.tp))
        self.gaussian_quadrature = gauss


class Helmholtz3D(Coulomb3D):
    """Helmholtz operator exp(-k r)/r with k^2 = -2E (helmholtz3d.F90:367-481):
    the Coulomb operator with re-weighted t-quadrature."""

    def __init__(self, ctx, gridin, gridout=None, gauss=None, energy: float = 0.0, coda=True):
        super().__init__(ctx, gridin, gridout, gauss, coda)
        self.energy = 0.0
        self.set_energy(energy)

    def set_energy(self, energy: float):
        call("dage_operator_set_energy", self.handle, float(energy))
        self.energy = float(energy)


def radial_grid_steps(z: float, ncell: int, nlip: int = 7, r_max: float = 20.0) -> np.ndarray:
    """Grid1D_init_radial cell steps (grid.F90:569-603) for ncell cells."""
    st = np.zeros(ncell)
    call("dage_radial_grid_steps", float(z), ncell, nlip, float(r_max), _ptr(st))
    return st


class Bubbles:
    """Atom-centred radial functions x real solid harmonics (bubbles.F90:90-230).

    bf for each bubble is (npts, (lmax+1)**2) in Fortran order with lm index
    l*l+l+m; the radial part used is r**k * f_lm(r)."""

    def __init__(self, ctx: Context, nbub: int, lmax: int, k: int = 0, nlip: int = 7):
        h = ctypes.c_void_p()
        call("dage_bubbles_create", ctx.handle, nbub, lmax, k, nlip, ctypes.byref(h))
        self.handle = h
        self.ctx = ctx
        self.nbub, self.lmax, self.k, self.nlip = nbub, lmax, k, nlip

    def close(self):
        if self.handle:
            _lib.load().dage_bubbles_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_bubble(self, ibub: int, center, charge: float, steps, bf):
        steps = _f64(steps)
        npts = len(steps) * (self.nlip - 1) + 1
        bf = np.asfortranarray(bf, dtype=np.float64)
        if bf.shape != (npts, (self.lmax + 1) ** 2):
            raise ValueError(f"bf shape {bf.shape} != {(npts, (self.lmax + 1) ** 2)}")
        call("dage_bubbles_set", self.handle, ibub, _ptr(_f64(center)), float(charge), len(steps),
             _ptr(steps), bf.ctypes.data_as(_dp))

    def inject(self, grid: Grid3D, cube: DeviceCube, lmin: int = 0, multiplier: float = 1.0):
        """cube += multiplier * bubbles (l >= lmin) evaluated on grid."""
        if cube.shape != grid.get_shape():
            raise ValueError(f"cube shape {cube.shape} != grid shape {grid.get_shape()}")
        call("dage_bubbles_inject_device", self.handle, grid.handle, lmin, float(multiplier),
             ctypes.c_void_p(cube.ptr))

    def inject_strided(self, grid: Grid3D, box, lmin: int = 0, multiplier: float = 1.0):
        """inject into a sub-cube (CubeView) whose shape is the grid's."""
        if box.shape != grid.get_shape():
            raise ValueError(f"box shape {box.shape} != grid shape {grid.get_shape()}")
        ld = getattr(box, "ld", box.shape[:2])
        call("dage_bubbles_inject_device_strided", self.handle, grid.handle, lmin,
             float(multiplier), ctypes.c_void_p(box.ptr), int(ld[0]), int(ld[1]))

    def eval_3dgrid(self, grid: Grid3D, lmin: int = 0, cube: np.ndarray | None = None,
                    multiplier: float = 1.0) -> np.ndarray:
        """Bubbles_eval_3dgrid on host memory: cube (zeros when not given) +
        multiplier * bubbles (l >= lmin) on the points of grid."""
        out = np.zeros(grid.get_shape(), order="F") if cube is None else \
            np.array(cube, dtype=np.float64, order="F", copy=True)
        if out.shape != grid.get_shape():
            raise ValueError(f"cube shape {out.shape} != grid shape {grid.get_shape()}")
        call("dage_bubbles_inject", self.handle, grid.handle, lmin, float(multiplier), _ptr(out))
        return out

    @classmethod
    def like(cls, src: "Bubbles", lmax: int | None = None, k: int = 0) -> "Bubbles":
        """Bubbles(src, lmax=..., k=...) without content: same centres,
        charges and radial grids, zero radial values."""
        lmax = src.lmax if lmax is None else lmax
        h = ctypes.c_void_p()
        call("dage_bubbles_create_like", src.handle, lmax, k, ctypes.byref(h))
        obj = cls.__new__(cls)
        obj.handle, obj.ctx = h, src.ctx
        obj.nbub, obj.lmax, obj.k, obj.nlip = src.nbub, lmax, k, src.nlip
        return obj

    def npoints(self, ibub: int) -> int:
        n = ctypes.c_int()
        call("dage_bubbles_npoints", self.handle, ibub, ctypes.byref(n))
        return n.value

    def get(self, ibub: int):
        """(centre, charge, cell steps, radial values f (npts, (lmax+1)^2))."""
        npts = self.npoints(ibub)
        ncell = (npts - 1) // (self.nlip - 1)
        c, q = np.zeros(3), ctypes.c_double()
        st = np.zeros(ncell)
        bf = np.zeros((npts, (self.lmax + 1) ** 2), order="F")
        call("dage_bubbles_get", self.handle, ibub, _ptr(c), ctypes.byref(q), _ptr(st),
             bf.ctypes.data_as(_dp))
        return c, q.value, st, bf

    def axpby(self, a: float, x: "Bubbles", b: float):
        """self = a x + b self (l above x's lmax: b self)."""
        call("dage_bubbles_axpby", float(a), x.handle, float(b), self.handle)

    def integrate(self) -> float:
        """sum over bubbles of 4 pi int r^(2+k) f_00 dr (Bubbles_integrate_all)."""
        r = ctypes.c_double()
        call("dage_bubbles_integrate", self.handle, ctypes.byref(r))
        return r.value

    def multiply(self, other: "Bubbles", out: "Bubbles | None" = None) -> "Bubbles":
        """Bubble-by-bubble product (lmax and k add up)."""
        out = out if out is not None else Bubbles.like(self, lmax=self.lmax + other.lmax,
                                                       k=self.k + other.k)
        call("dage_bubbles_multiply", self.handle, other.handle, out.handle)
        return out

    def coulomb_transform(self, out: "Bubbles | None" = None) -> "Bubbles":
        """Coulomb3D_evaluate_bubbles_potential (coulomb.F90:222-336)."""
        out = out if out is not None else Bubbles.like(self, k=0)
        call("dage_bubbles_coulomb_transform", self.handle, out.handle)
        return out

    def helmholtz_transform(self, energy: float, out: "Bubbles | None" = None) -> "Bubbles":
        """Helmholtz3D_transform_bubbles_sub (helmholtz3d.F90:519-657)."""
        out = out if out is not None else Bubbles.like(self, k=0)
        call("dage_bubbles_helmholtz_transform", self.handle, float(energy), out.handle)
        return out


class GBFMMCoulomb3D:
    """Grid-based FMM Coulomb operator on one grid (gbfmm_coulomb3d.F90:306-1330).

    2**maxlevel boxes per axis, multipoles up to lmax, near field with the
    t-quadrature `gauss` (the reference's default GaussQuad(25, 25))."""

    def __init__(self, ctx: Context, grid: Grid3D, maxlevel: int = 2, lmax: int = 12,
                 gauss: "GaussQuad | None" = None):
        gauss = gauss if gauss is not None else GaussQuad(25, 25)
        h = ctypes.c_void_p()
        call("dage_gbfmm_create", ctx.handle, grid.handle, maxlevel, lmax, _ptr(gauss.tp),
             _ptr(gauss.tw), gauss.get_npoints(), ctypes.byref(h))
        self.handle = h
        self.ctx, self.grid = ctx, grid
        self.maxlevel, self.lmax = maxlevel, lmax
        self.energy = 0.0
        n = ctypes.c_int()
        call("dage_gbfmm_box_count", h, ctypes.byref(n))
        self.nbox_total = n.value

    def band_t0(self) -> int:
        """First t-point whose near-field passes run banded (T: none)."""
        t0 = ctypes.c_int()
        call("dage_gbfmm_get_band", self.handle, ctypes.byref(t0))
        return t0.value

    def lowrank_t1(self) -> int:
        """Number of leading near-field t-points applied through low-rank factors."""
        t1, kx = ctypes.c_int(), ctypes.c_int()
        call("dage_gbfmm_get_lowrank", self.handle, ctypes.byref(t1), ctypes.byref(kx))
        return t1.value

    def close(self):
        if self.handle:
            _lib.load().dage_gbfmm_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_energy(self, energy: float):
        """GBFMMHelmholtz3D_set_energy: energy < -0.005 makes this the
        Helmholtz operator (kernel exp(-kappa r)/r, kappa = sqrt(-2 energy)),
        energy >= -0.005 the Coulomb one again."""
        call("dage_gbfmm_set_energy", self.handle, float(energy))
        self.energy = float(energy)

    def apply(self, rho: np.ndarray, precision: int = DAGE_PREC_FP64) -> np.ndarray:
        """Potential of the density cube rho (host arrays, Fortran order).
        precision selects the near-field arithmetic; the far field is FP64."""
        shape = self.grid.get_shape()
        rho = np.asfortranarray(rho, dtype=np.float64)
        if rho.shape != shape:
            raise ValueError(f"density shape {rho.shape} != grid shape {shape}")
        pot = np.zeros(shape, order="F")
        call("dage_gbfmm_apply", self.handle, rho.ctypes.data_as(_dp), pot.ctypes.data_as(_dp),
             int(precision))
        return pot

    def apply_device(self, rho: "DeviceCube", pot: "DeviceCube", precision: int = DAGE_PREC_FP64):
        call("dage_gbfmm_apply_device", self.handle, ctypes.c_void_p(rho.ptr),
             ctypes.c_void_p(pot.ptr), int(precision))

    def multipole_moments(self) -> np.ndarray:
        q = np.zeros(((self.lmax + 1) ** 2, self.nbox_total), order="F")
        call("dage_gbfmm_get_multipoles", self.handle, q.ctypes.data_as(_dp))
        return q

    def local_expansions(self) -> np.ndarray:
        L = np.zeros(((self.lmax + 1) ** 2, self.nbox_total), order="F")
        call("dage_gbfmm_get_local_expansions", self.handle, L.ctypes.data_as(_dp))
        return L


def interaction_matrix(lmax: int, dR) -> np.ndarray:
    """InteractionMatrix%eval: L_target += T @ q_source, dR = source - target."""
    N = (lmax + 1) ** 2
    T = np.zeros((N, N), order="F")
    call("dage_fmm_interaction_matrix", lmax, _ptr(_f64(dR)), T.ctypes.data_as(_dp))
    return T


def translation_matrix(lmax: int, dR) -> np.ndarray:
    """TranslationMatrix%eval unpacked: q(new) = W @ q(old), dR = old - new."""
    N = (lmax + 1) ** 2
    W = np.zeros((N, N), order="F")
    call("dage_fmm_translation_matrix", lmax, _ptr(_f64(dR)), W.ctypes.data_as(_dp))
    return W


class HelmholtzBatch:
    """Helmholtz applies of many orbitals at once, one hipGraph launch per
    apply.  grids[i] is orbital i's grid (input = output), energies[i] its
    energy (Helmholtz3D_set_energy semantics); orbitals on identical grids
    share matrices."""

    def __init__(self, ctx: Context, grids, energies, gauss: GaussQuad,
                 precision: int = DAGE_PREC_FP64, coda: bool = True):
        self.ctx, self.grids = ctx, list(grids)
        n = len(self.grids)
        arr = (ctypes.c_void_p * n)(*[g.handle.value if isinstance(g.handle, ctypes.c_void_p)
                                      else g.handle for g in self.grids])
        e = _f64(energies)
        if len(e) != n:
            raise ValueError("one energy per grid")
        h = ctypes.c_void_p()
        call("dage_helmholtz_batch_create", ctx.handle, n, arr, _ptr(e), _ptr(gauss.tp),
             _ptr(gauss.tw), gauss.get_npoints(), float(gauss.get_tlog()),
             0 if coda else DAGE_OP_NO_CODA, int(precision), ctypes.byref(h))
        self.handle = h
        self.n = n
        g = ctypes.c_int()
        call("dage_batch_group_count", h, ctypes.byref(g))
        self.ngroups = g.value

    def close(self):
        if self.handle:
            _lib.load().dage_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, i: int, cube: np.ndarray):
        a = np.asfortranarray(cube, dtype=np.float64)
        if a.shape != self.grids[i].get_shape():
            raise ValueError(f"cube {i} shape {a.shape} != grid shape {self.grids[i].get_shape()}")
        call("dage_batch_upload", self.handle, i, a.ctypes.data_as(_dp))

    def download(self, i: int) -> np.ndarray:
        out = np.zeros(self.grids[i].get_shape(), order="F")
        call("dage_batch_download", self.handle, i, out.ctypes.data_as(_dp))
        return out

    def input_ptr(self, i: int) -> int:
        p = ctypes.c_void_p()
        call("dage_batch_input", self.handle, i, ctypes.byref(p))
        return p.value

    def output_ptr(self, i: int) -> int:
        p = ctypes.c_void_p()
        call("dage_batch_output", self.handle, i, ctypes.byref(p))
        return p.value

    def set_energies(self, energies):
        e = _f64(energies)
        if len(e) != self.n:
            raise ValueError("one energy per orbital")
        call("dage_batch_set_energies", self.handle, _ptr(e))

    def apply(self):
        """All orbitals: one graph launch (asynchronous on the context stream)."""
        call("dage_batch_apply", self.handle)

    def apply_eager(self):
        call("dage_batch_apply_eager", self.handle)

    def tune(self, reps: int = 2) -> np.ndarray:
        """Time every GEMM column-block count (1-8) and the operator's own
        tiling (0) per pass and grid group, keep the fastest and re-capture
        the graph; returns the choices, one row (x, y, z) per group.
        Overwrites the outputs."""
        tiles = (ctypes.c_int * (3 * self.ngroups))()
        call("dage_batch_tune", self.handle, int(reps), tiles)
        return np.array(tiles[:], dtype=int).reshape(self.ngroups, 3)
