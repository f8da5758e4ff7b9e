"""Learned orbital-update preconditioner for the Helmholtz SCF.

A 4-layer 3D convolutional network (3x3x3 kernels, 16 channels, ReLU; BF16
weights with FP32 master copies, FP32 activations and accumulation) reads
the current orbital psi and its plain Helmholtz update Delta (scaled to unit
RMS) and proposes a correction to the plain step:

    proposal = psi + Delta + s * N([psi, Delta / s]),   s = rms(Delta).

It is trained on the pairs the plain iteration generates: for every
iteration k and orbital i, the target is what the plain step misses,
(psi*_i - psi_k,i - Delta_k,i) / s, where psi* are the converged orbitals
rotated onto psi_k by the orthogonal Procrustes rotation (orbitals are only
defined up to a rotation of the occupied space).  Forward and backward run in
the library's HIP kernels (dage_conv3d_*), the optimiser is Adam
(dage_adam_step); the last layer starts at zero, so an untrained network
proposes exactly the plain step.  In the SCF the plain proposal stays the
fallback and the arbiter (HelmholtzSCF.run keeps whichever has the lower
occupied Ritz eigenvalue sum).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from ._lib import call
from .operators import Context, DeviceCube

LAYERS = ((2, 16), (16, 16), (16, 16), (16, 1))


class _Buf:
    """Raw device buffer owned through the C-ABI."""

    def __init__(self, ctx: Context, nbytes: int):
        p = ctypes.c_void_p()
        call("dage_device_alloc", ctx.handle, ctypes.c_size_t(max(8, nbytes)), ctypes.byref(p))
        self.ptr, self.ctx, self.nbytes = p.value, ctx, nbytes

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        call("dage_memcpy", self.ctx.handle, ctypes.c_void_p(self.ptr),
             a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 0)

    def download(self, dtype, count) -> np.ndarray:
        out = np.zeros(count, dtype=dtype)
        call("dage_memcpy", self.ctx.handle, out.ctypes.data_as(ctypes.c_void_p),
             ctypes.c_void_p(self.ptr), ctypes.c_size_t(out.nbytes), 1)
        return out

    def close(self):
        if self.ptr:
            call("dage_device_free", self.ctx.handle, ctypes.c_void_p(self.ptr))
            self.ptr = None


def _p(x):
    return ctypes.c_void_p(x.ptr if hasattr(x, "ptr") else x)


class ConvPreconditioner:
    def __init__(self, ctx: Context, shape, seed: int = 0, lr: float = 1e-3):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in shape)
        self.vol = int(np.prod(self.shape))
        self.lr, self.step_count = lr, 0
        rng = np.random.default_rng(seed)
        self.params = []
        for li, (ci, co) in enumerate(LAYERS):
            n = co * ci * 27
            std = math.sqrt(2.0 / (ci * 27))
            w = (rng.standard_normal(n) * std).astype(np.float32)
            if li == len(LAYERS) - 1:
                w[:] = 0.0  # untrained network == plain step
            b = np.zeros(co, np.float32)
            P = {"ci": ci, "co": co, "n": n}
            for k, nb in (("w", 4 * n), ("wbf", 2 * n), ("b", 4 * co), ("gw", 4 * n),
                          ("gb", 4 * co), ("mw", 4 * n), ("vw", 4 * n), ("mb", 4 * co),
                          ("vb", 4 * co), ("bbf", 2 * co)):
                P[k] = _Buf(ctx, nb)
            P["w"].upload(w)
            P["b"].upload(b)
            for k, cnt in (("mw", n), ("vw", n)):
                P[k].upload(np.zeros(cnt, np.float32))
            for k in ("mb", "vb"):
                P[k].upload(np.zeros(co, np.float32))
            call("dage_weights_to_bf16", ctx.handle, n, _p(P["w"]), _p(P["wbf"]))
            self.params.append(P)
        v = self.vol
        self.act = [_Buf(ctx, 4 * v * ci) for ci, _ in LAYERS] + [_Buf(ctx, 4 * v)]
        self.grad = [_Buf(ctx, 4 * v * 16) for _ in range(3)] + [_Buf(ctx, 4 * v)]
        self.target = _Buf(ctx, 4 * v)

    # ------------------------------------------------------------ parameters
    def get_weights(self):
        """[(w float32 [cout][cin][27], b float32 [cout])] of the master copies."""
        out = []
        for P in self.params:
            w = P["w"].download(np.float32, P["n"]).reshape(P["co"], P["ci"], 27)
            b = P["b"].download(np.float32, P["co"])
            out.append((w, b))
        return out

    def set_weights(self, weights):
        for P, (w, b) in zip(self.params, weights):
            P["w"].upload(np.asarray(w, np.float32).ravel())
            P["b"].upload(np.asarray(b, np.float32))
            call("dage_weights_to_bf16", self.ctx.handle, P["n"], _p(P["w"]), _p(P["wbf"]))

    # -------------------------------------------------------- forward/backward
    def forward(self):
        """act[0] (2 channels) -> act[4] (1 channel)."""
        nx, ny, nz = self.shape
        for li, P in enumerate(self.params):
            last = li == len(self.params) - 1
            call("dage_conv3d_forward", self.ctx.handle, nx, ny, nz, P["ci"], P["co"],
                 _p(self.act[li]), _p(P["wbf"]), _p(P["b"]), 0 if last else 1,
                 _p(self.act[li + 1]))

    def backward(self):
        """grad[3] = dL/d act[4]; fills every layer's gw, gb."""
        nx, ny, nz = self.shape
        for li in range(len(self.params) - 1, -1, -1):
            P = self.params[li]
            gout = self.grad[li]
            call("dage_conv3d_backward_filter", self.ctx.handle, nx, ny, nz, P["ci"], P["co"],
                 _p(self.act[li]), _p(gout), _p(P["gw"]), _p(P["gb"]))
            if li > 0:
                gin = self.grad[li - 1]
                call("dage_conv3d_backward_data", self.ctx.handle, nx, ny, nz, P["ci"], P["co"],
                     _p(gout), _p(P["wbf"]), _p(gin))
                # through the ReLU of the previous layer's output act[li]
                call("dage_relu_backward", self.ctx.handle, self.vol * P["ci"], _p(self.act[li]),
                     _p(gin))

    def adam(self, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step_count += 1
        for P in self.params:
            call("dage_adam_step", self.ctx.handle, P["n"], _p(P["w"]), _p(P["gw"]), _p(P["mw"]),
                 _p(P["vw"]), self.lr, beta1, beta2, eps, self.step_count, _p(P["wbf"]))
            call("dage_adam_step", self.ctx.handle, P["co"], _p(P["b"]), _p(P["gb"]), _p(P["mb"]),
                 _p(P["vb"]), self.lr, beta1, beta2, eps, self.step_count, _p(P["bbf"]))

    def load_input(self, psi: DeviceCube, delta: DeviceCube, scale: float):
        v = self.vol
        call("dage_cvt_f64_f32", self.ctx.handle, v, _p(psi), 1.0, _p(self.act[0]))
        call("dage_cvt_f64_f32", self.ctx.handle, v, _p(delta), 1.0 / scale,
             ctypes.c_void_p(self.act[0].ptr + 4 * v))

    def train_step(self, sample) -> float:
        """sample = (input buffer [2][vol] float32, target buffer [vol] float32)."""
        x, t = sample
        call("dage_memcpy", self.ctx.handle, _p(self.act[0]), _p(x),
             ctypes.c_size_t(8 * self.vol), 2)
        self.forward()
        loss = ctypes.c_double()
        call("dage_mse_loss", self.ctx.handle, self.vol, _p(self.act[-1]), _p(t),
             _p(self.grad[3]), ctypes.byref(loss))
        self.backward()
        self.adam()
        return loss.value

    # -------------------------------------------------------------- the SCF
    def propose(self, scf, psi, delta, out):
        """out_i = psi_i + Delta_i + s_i N([psi_i, Delta_i / s_i]) (FP64 cubes)."""
        for p, d, o in zip(psi, delta, out):
            s = math.sqrt(max(scf.integrate(d, d), 1e-300) / scf.box_volume)
            self.load_input(p, d, s)
            self.forward()
            o.copy_from(p)
            scf._axpby(1.0, d, 1.0, o)
            call("dage_axpy_f32_f64", self.ctx.handle, self.vol, s, _p(self.act[-1]), _p(o))

    def close(self):
        for P in self.params:
            for k, v in P.items():
                if isinstance(v, _Buf):
                    v.close()
        for b in self.act + self.grad + [self.target]:
            b.close()


class SampleRecorder:
    """Collects (psi_k, Delta_k) of a plain SCF run and turns them into
    training samples once the converged orbitals are known."""

    def __init__(self, scf):
        self.scf = scf
        self.records = []  # (psi copies, delta copies, scales)

    def __call__(self, psi, delta, out):
        """Used as the SCF `update`: records, then proposes the plain step."""
        scf = self.scf
        ps = [DeviceCube(scf.ctx, scf.shape) for _ in psi]
        ds = [DeviceCube(scf.ctx, scf.shape) for _ in psi]
        sc = []
        for p, d, pc, dc, o in zip(psi, delta, ps, ds, out):
            pc.copy_from(p)
            dc.copy_from(d)
            sc.append(math.sqrt(max(scf.integrate(d, d), 1e-300) / scf.box_volume))
            o.copy_from(p)
            scf._axpby(scf.w, d, 1.0, o)
        self.records.append((ps, ds, sc))

    def samples(self, model: ConvPreconditioner, final):
        """Training pairs against the converged orbitals `final` (device cubes)."""
        scf = self.scf
        v = scf.n
        out = []
        tmp = [DeviceCube(scf.ctx, scf.shape) for _ in final]
        for ps, ds, sc in self.records:
            M = scf._gram(final, ps)  # M[i, j] = <final_i | psi_j>
            U, _, Vt = np.linalg.svd(M)
            R = U @ Vt                # psi*_aligned_j = sum_i final_i R[i, j]
            scf._lincomb(final, R, tmp)
            for j, (p, d, s) in enumerate(zip(ps, ds, sc)):
                # target = (aligned - psi - Delta) / s
                scf._axpby(-1.0, p, 1.0, tmp[j])
                scf._axpby(-1.0, d, 1.0, tmp[j])
                x = _Buf(scf.ctx, 8 * v)
                t = _Buf(scf.ctx, 4 * v)
                call("dage_cvt_f64_f32", scf.ctx.handle, v, _p(p), 1.0, _p(x))
                call("dage_cvt_f64_f32", scf.ctx.handle, v, _p(d), 1.0 / s,
                     ctypes.c_void_p(x.ptr + 4 * v))
                call("dage_cvt_f64_f32", scf.ctx.handle, v, _p(tmp[j]), 1.0 / s, _p(t))
                out.append((x, t))
        for c in tmp:
            c.close()
        return out

    def close(self):
        for ps, ds, _ in self.records:
            for c in ps + ds:
                c.close()
        self.records = []


def train(model: ConvPreconditioner, samples, epochs: int = 3, seed: int = 0, log=None):
    rng = np.random.default_rng(seed)
    losses = []
    for ep in range(epochs):
        order = rng.permutation(len(samples))
        tot = 0.0
        for i in order:
            tot += model.train_step(samples[i])
        losses.append(tot / len(samples))
        if log:
            log(ep, losses[-1])
    return losses
