"""The oracle's 3x3x3 convolution (test infrastructure for the learned
preconditioner's HIP layers), pinned by scipy.ndimage.correlate for the
forward pass and by central finite differences for every gradient."""
import numpy as np
from scipy.ndimage import correlate

import oracle as O


def test_forward_matches_scipy():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 5, 6, 7))
    w = O.bf16_round(rng.standard_normal((3, 2, 27)))
    b = rng.standard_normal(3)
    y = O.conv3d_forward(x, w, b, relu=False)
    for o in range(3):
        ref = sum(correlate(x[c], w[o, c].reshape(3, 3, 3), mode="constant") for c in range(2))
        assert np.max(np.abs(y[o] - ref - b[o])) < 1e-12


def test_gradients_by_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 4, 5, 6))
    w = O.bf16_round(rng.standard_normal((3, 2, 27)))
    b = rng.standard_normal(3)
    g = rng.standard_normal((3, 4, 5, 6))
    f = lambda x, w, b: np.sum(g * O.conv3d_forward(x, w, b, False))
    gx, gw, gb = O.conv3d_backward(x, w, g)
    e = 1e-6
    for idx in [(0, 1, 2, 3), (1, 3, 0, 5), (0, 0, 4, 0)]:
        xp, xm = x.copy(), x.copy()
        xp[idx] += e
        xm[idx] -= e
        assert abs((f(xp, w, b) - f(xm, w, b)) / (2 * e) - gx[idx]) < 1e-6
    for idx in [(2, 1, 13), (0, 0, 0), (1, 1, 26)]:
        wp, wm = w.copy(), w.copy()
        wp[idx] += e
        wm[idx] -= e
        assert abs((f(x, wp, b) - f(x, wm, b)) / (2 * e) - gw[idx]) < 1e-6
    assert np.allclose(gb, g.sum(axis=(1, 2, 3)))


def test_bf16_rounding():
    v = O.bf16_round(np.array([1.0, 1.00390625, 1.0078125, 3.14159, -2.5e-3]))
    assert list(v[:4]) == [1.0, 1.0, 1.0078125, 3.140625]
    assert abs(v[4] + 2.5e-3) < 2.5e-3 * 2 ** -8
