"""Learned preconditioner: the HIP convolution layers against the numpy
oracle (forward and every gradient), training on the GPU, and the gain over
the plain Helmholtz SCF on a small molecule.

Bars: forward of the 4-layer network within 2e-5 of max|y| of the oracle
(float64 arithmetic on the same BF16 weights; the GPU accumulates in FP32);
filter, bias and input gradients within 1e-4 of their maxima.  The oracle's
convolution is pinned by scipy.ndimage.correlate (forward) and central
finite differences (gradients) in test_oracle_learned.py.
"""
import numpy as np
import pytest

import dage_amd
import dage_amd.learned as L
import dage_amd.scf as scf
import oracle as O

pytestmark = pytest.mark.gpu


def oracle_network(weights, x):
    acts = [x]
    for li, (w, b) in enumerate(weights):
        last = li == len(weights) - 1
        acts.append(O.conv3d_forward(acts[-1], O.bf16_round(w), b.astype(np.float64), not last))
    return acts


def test_layers_match_oracle(ctx):
    shape = (9, 7, 6)  # nx, ny, nz
    model = L.ConvPreconditioner(ctx, shape, seed=3)
    rng = np.random.default_rng(4)
    # non-zero last layer so every gradient is exercised
    ws = model.get_weights()
    ws[-1] = (rng.standard_normal(ws[-1][0].shape).astype(np.float32) * 0.1,
              rng.standard_normal(1).astype(np.float32) * 0.1)
    ws[1] = (ws[1][0], rng.standard_normal(16).astype(np.float32) * 0.1)
    model.set_weights(ws)
    x = rng.standard_normal((2, 6, 7, 9)).astype(np.float32)  # [c][z][y][x]
    t = rng.standard_normal((1, 6, 7, 9)).astype(np.float32)
    xb = L._Buf(ctx, x.nbytes)
    tb = L._Buf(ctx, t.nbytes)
    xb.upload(x)
    tb.upload(t)
    loss = model.train_step((xb, tb))  # forward + backward + Adam (weights move after)
    acts = oracle_network(ws, x.astype(np.float64))
    y = acts[-1]
    got = model.act[-1].download(np.float32, y.size).reshape(y.shape)
    assert np.max(np.abs(got - y)) < 2e-5 * np.max(np.abs(y))
    assert abs(loss - np.mean((y - t) ** 2)) < 1e-5 * np.mean((y - t) ** 2)
    # oracle backprop
    g = 2.0 * (y - t) / y.size
    grads = []
    for li in range(len(ws) - 1, -1, -1):
        w, b = ws[li]
        gx, gw, gb = O.conv3d_backward(acts[li], O.bf16_round(w), g)
        grads.append((li, gw, gb))
        if li > 0:
            g = gx * (acts[li] > 0)
    for li, gw, gb in grads:
        P = model.params[li]
        gwg = P["gw"].download(np.float32, P["n"]).reshape(P["co"], P["ci"], 27)
        gbg = P["gb"].download(np.float32, P["co"])
        assert np.max(np.abs(gwg - gw)) < 1e-4 * np.max(np.abs(gw)), li
        assert np.max(np.abs(gbg - gb)) < 1e-4 * max(np.max(np.abs(gb)), 1e-12), li
    for b in (xb, tb):
        b.close()
    model.close()


@pytest.mark.parametrize("shape", [(9, 7, 6), (20, 6, 5), (40, 5, 7), (50, 9, 11)])
def test_conv16_shapes(ctx, shape):
    """16 -> 16 layers (bf16 MFMA over three-term activation splits): forward
    with bias and ReLU and backward-data, at x extents that select 1-4
    x-subtiles per block and several z-chunks, within 2e-5 of the maxima of
    the oracle's float64 results on the same BF16 weights."""
    import ctypes
    nx, ny, nz = shape
    rng = np.random.default_rng(nx * 131 + nz)
    x = rng.standard_normal((16, nz, ny, nx)).astype(np.float32)
    w = (rng.standard_normal((16, 16, 27)) * 0.1).astype(np.float32)
    b = (rng.standard_normal(16) * 0.1).astype(np.float32)
    wr = O.bf16_round(w)
    bits = (wr.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)
    bufs = [L._Buf(ctx, a.nbytes) for a in (x, bits, b, x)]
    for buf, a in zip(bufs, (x, bits, b)):
        buf.upload(a)
    dx_, dw, db, dy_ = (ctypes.c_void_p(q.ptr) for q in bufs)
    dage_amd._lib.call("dage_conv3d_forward", ctx.handle, nx, ny, nz, 16, 16, dx_, dw, db, 1, dy_)
    got = bufs[3].download(np.float32, x.size).reshape(x.shape)
    ref = O.conv3d_forward(x.astype(np.float64), wr, b.astype(np.float64), True)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    # backward data of sum(g * conv(x)): g = the forward output as a test field
    dage_amd._lib.call("dage_conv3d_backward_data", ctx.handle, nx, ny, nz, 16, 16, dy_, dw, dx_)
    gx = bufs[0].download(np.float32, x.size).reshape(x.shape)
    gref, _, _ = O.conv3d_backward(x.astype(np.float64), wr, got.astype(np.float64))
    gerr = np.max(np.abs(gx - gref)) / np.max(np.abs(gref))
    print(f"conv16 {shape}: forward {err:.2e}, backward data {gerr:.2e}")
    assert err < 2e-5 and gerr < 2e-5
    for buf in bufs:
        buf.close()


def test_training_reduces_loss(ctx):
    """Fit a fixed 3x3x3 filter of the second input channel (a target the
    network can represent) from 8 random volumes."""
    from scipy.ndimage import correlate
    shape = (12, 12, 12)
    rng = np.random.default_rng(5)
    k = rng.standard_normal((3, 3, 3)) * 0.3
    samples = []
    for _ in range(8):
        x = rng.standard_normal((2, 12, 12, 12)).astype(np.float32)
        y = correlate(x[1].astype(np.float64), k, mode="constant").astype(np.float32)
        xb = L._Buf(ctx, x.nbytes)
        xb.upload(x)
        tb = L._Buf(ctx, y.nbytes)
        tb.upload(y)
        samples.append((xb, tb))
    student = L.ConvPreconditioner(ctx, shape, seed=12, lr=3e-3)
    losses = L.train(student, samples, epochs=40)
    print("losses", [f"{l:.3e}" for l in losses[::8]], f"{losses[-1]:.3e}")
    assert losses[-1] < 0.2 * losses[0]
    for xb, tb in samples:
        xb.close()
        tb.close()
    student.close()


def test_learned_scf_converges_faster(ctx):
    ax = O.Axis.ncell_fixed(-6.0, 6.0, 5, 12)  # 49^3, h = 0.25
    g = dage_amd.Grid3D(ctx, [ax.qmin] * 3, [ax.steps] * 3, 5)
    mol = scf.Molecule(np.array([3.0, 1.0, 1.0]),
                       np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.6], [1.5, 0.0, -0.6]]))
    plain = scf.HelmholtzSCF(ctx, g, mol, nocc=2, beta=4.0)
    plain.initial_guess()
    rec = L.SampleRecorder(plain)
    r0 = plain.run(max_iter=120, update=rec)
    assert r0.converged
    model = L.ConvPreconditioner(ctx, plain.shape, seed=0, lr=2e-3)
    samples = rec.samples(model, plain.psi)
    losses = L.train(model, samples, epochs=4)
    learned = scf.HelmholtzSCF(ctx, g, mol, nocc=2, beta=4.0)
    learned.initial_guess()
    r1 = learned.run(max_iter=120, update=lambda p, d, o: model.propose(learned, p, d, o))
    print(f"plain {r0.iterations} it {r0.wall_s:.2f} s; learned {r1.iterations} it "
          f"{r1.wall_s:.2f} s (used {sum(r1.learned_used)}); losses {losses}")
    assert r1.converged
    assert abs(r1.energies[-1] - r0.energies[-1]) < 1e-5
    assert r1.iterations < r0.iterations
    for xb, tb in samples:
        xb.close()
        tb.close()
    rec.close()
    for o in (plain, learned, model):
        o.close()
