"""Pins the float64 MLP reference itself: its backward pass equals central
finite differences of its loss (with the relative-L2 denominator frozen, as
the kernels treat it), on every layer."""
import numpy as np

import mlp_ref as M


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init_params(1)
    u = rng.uniform(size=(40, 9))
    t = rng.uniform(0, 2, size=(40, 3))
    g = M.backward(p, u, t, emulate=False)
    # a few parameters in every weight matrix and bias
    idx, o = [], 0
    for r, c in M.SHAPES:
        idx += list(o + rng.choice(r * c, 4, replace=False)); o += r * c
        idx += list(o + rng.choice(r, 2, replace=False)); o += r
    fd = M.loss_detached_grad_fd(p, u, t, idx=idx)
    assert np.allclose(g[idx], fd, rtol=1e-5, atol=1e-9)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -3.14159], np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == 1.0
    assert abs(r[4] + 3.140625) < 1e-12


def test_encoding_width():
    e = M.encode(np.random.default_rng(1).uniform(size=(5, 9)))
    assert e.shape == (5, 63)
    assert np.allclose(e[:, 9:18], np.sin(np.pi * e[:, :9]), atol=1e-6)
