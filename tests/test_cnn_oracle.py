"""Pins of the FP64 network restatement (oracle/cnn_oracle.py): its
backward pass against central finite differences of its forward pass, the
3x3 convolution against a direct loop, and Adam against a hand-worked step.
CPU only."""
import numpy as np
import pytest

import cnn_oracle as C


def test_conv3x3_matches_direct_loop(rng):
    x = rng.standard_normal((2, 5, 6, 3))
    w = rng.standard_normal((3, 3, 3, 4))
    b = rng.standard_normal(4)
    got = C.conv3x3(x, w, b)
    ref = np.zeros((2, 5, 6, 4))
    for n in range(2):
        for y in range(5):
            for xx in range(6):
                acc = b.copy()
                for ky in range(3):
                    for kx in range(3):
                        yy, xs = y + ky - 1, xx + kx - 1
                        if 0 <= yy < 5 and 0 <= xs < 6:
                            acc += x[n, yy, xs] @ w[ky, kx]
                ref[n, y, xx] = acc
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("levels", [1, 2, 3])
def test_backward_matches_finite_differences(levels, rng):
    base, H = 3, 8
    params = C.init_params(levels, base, seed=5)
    params = {k: (v + 0.05 * rng.standard_normal(v.shape)).astype(np.float64) for k, v in params.items()}
    x = rng.standard_normal((2, H, H, 2))
    target = rng.standard_normal((2, H, H, 1))

    def loss(p):
        y, _ = C.forward(p, x, levels)
        return C.mse(y, target)[0]

    y, act = C.forward(params, x, levels)
    _, gy = C.mse(y, target)
    grads = C.backward(params, act, gy, levels)
    assert set(grads) == set(params)
    h = 1e-6
    for name in params:
        flat = params[name].reshape(-1)
        for idx in rng.choice(flat.size, size=min(4, flat.size), replace=False):
            p1 = {k: v.copy() for k, v in params.items()}
            p2 = {k: v.copy() for k, v in params.items()}
            p1[name].reshape(-1)[idx] += h
            p2[name].reshape(-1)[idx] -= h
            fd = (loss(p1) - loss(p2)) / (2 * h)
            an = grads[name].reshape(-1)[idx]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (name, idx, fd, an)


def test_parameter_count_of_the_full_network():
    """4 levels, 32 base channels: about 1.75 million parameters."""
    shapes = C.layer_shapes(4, 32)
    n = sum(9 * ci * co + co for _, ci, co in shapes)
    assert 1.7e6 < n < 1.8e6


def test_adam_first_step_moves_by_lr():
    """The first Adam step moves each parameter by lr * sign(grad)
    (bias-corrected moments m/v^(1/2) = g/|g|)."""
    p = np.array([1.0, -2.0, 3.0])
    g = np.array([0.5, -4.0, 1e-3])
    new, m, v = C.adam_step(p, g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    np.testing.assert_allclose(new, p - 0.01 * np.sign(g), rtol=1e-6)
