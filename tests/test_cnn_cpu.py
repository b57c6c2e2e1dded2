"""CPU checks of the warm-start network's FP64 oracle (pdt_cnn_oracle.cpp):
the forward pass against closed forms (tap orientation, zero padding, ReLU),
the analytic gradient against central finite differences, and Adam against
its closed form after one and two steps."""
import numpy as np


def test_cnn_forward_closed_forms(oracle):
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 5, 6, 1))
    # centre tap = identity
    p = np.zeros(10)
    p[4] = 1.0
    assert np.array_equal(oracle.cnn_forward([1, 1], p, x), x)
    # tap (ky, kx) = (0, 1) reads the pixel above: y[i, j] = x[i - 1, j], zero on row 0
    p = np.zeros(10)
    p[1] = 1.0
    y = oracle.cnn_forward([1, 1], p, x)
    assert np.array_equal(y[:, 1:], x[:, :-1])
    assert np.array_equal(y[:, 0], np.zeros_like(y[:, 0]))
    # tap (1, 2) reads the pixel to the right, zero in the last column
    p = np.zeros(10)
    p[5] = 1.0
    y = oracle.cnn_forward([1, 1], p, x)
    assert np.array_equal(y[:, :, :-1], x[:, :, 1:])
    assert np.array_equal(y[:, :, -1], np.zeros_like(y[:, :, -1]))
    # a constant image through an all-ones kernel: 9 inside, 4 in the corners
    p = np.ones(10)
    p[9] = 0.5  # bias
    y = oracle.cnn_forward([1, 1], p, np.ones((1, 4, 4, 1)))[0, :, :, 0]
    assert y[1, 1] == 9.5 and y[0, 0] == 4.5 and y[0, 1] == 6.5
    # ReLU between layers: a negative hidden bias zeroes the second layer's input
    p = np.zeros(20)
    p[4], p[9], p[14], p[19] = 1.0, -100.0, 1.0, 0.25
    y = oracle.cnn_forward([1, 1, 1], p, x)
    assert np.all(y == 0.25)


def test_cnn_gradient_matches_finite_differences(oracle):
    ch = [2, 3, 4, 3]
    rng = np.random.default_rng(1)
    n, h, w = 2, 5, 6
    p = rng.standard_normal(oracle.cnn_num_params(ch)) * 0.5
    x = rng.standard_normal((n, h, w, 2))
    t = rng.standard_normal((n, h, w, 3))
    loss, g = oracle.cnn_loss_grad(ch, p, x, t)
    y = oracle.cnn_forward(ch, p, x)
    assert abs(loss - 0.5 * np.mean((y - t) ** 2)) <= 1e-14 * abs(loss)
    idx = np.concatenate([rng.choice(p.size, 80, replace=False), np.arange(p.size - 3, p.size)])
    eps = 1e-6
    for i in idx:
        p1, p2 = p.copy(), p.copy()
        p1[i] += eps
        p2[i] -= eps
        fd = (oracle.cnn_loss_grad(ch, p1, x, t)[0] - oracle.cnn_loss_grad(ch, p2, x, t)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, np.abs(g).max()), (i, fd, g[i])


def test_adam_closed_form(oracle):
    rng = np.random.default_rng(2)
    w0 = rng.standard_normal(50)
    g1 = rng.standard_normal(50)
    g2 = rng.standard_normal(50)
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
    w, m, v = w0.copy(), np.zeros(50), np.zeros(50)
    oracle.adam(w, g1.copy(), m, v, lr, b1, b2, eps, 1)
    # step 1: m_hat = g, v_hat = g^2
    assert np.allclose(w, w0 - lr * g1 / (np.abs(g1) + eps), rtol=0, atol=1e-15)
    oracle.adam(w, g2.copy(), m, v, lr, b1, b2, eps, 2)
    mh = (b1 * (1 - b1) * g1 + (1 - b1) * g2) / (1 - b1 ** 2)
    vh = (b2 * (1 - b2) * g1 ** 2 + (1 - b2) * g2 ** 2) / (1 - b2 ** 2)
    ref = w0 - lr * g1 / (np.abs(g1) + eps) - lr * mh / (np.sqrt(vh) + eps)
    assert np.allclose(w, ref, rtol=0, atol=1e-15)
