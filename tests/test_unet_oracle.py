"""The FP64 restatement of the correction network, pinned by mathematics:
central finite differences of its own forward pass for every parameter kind
(every layer's weights and bias, the head), closed forms for the building
blocks (a conv with a one-hot kernel is a shift, pooling and upsampling are
adjoint up to the factor 8, convolution backward is the adjoint of forward).
A wrong tap order, a transposed weight or a dropped skip connection fails one
of these."""
import numpy as np
import pytest

from oracle import unet_oracle as U


def test_conv_with_one_hot_kernel_is_a_shift():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(5, 6, 7, 2))
    w = np.zeros((2, 27, 2))
    t = (2 * 3 + 0) * 3 + 1          # dz = 2, dy = 0, dx = 1: reads (z + 1, y - 1, x)
    w[0, t, 1] = 1.0
    w[1, 13, 0] = 3.0                # centre tap
    y = U.conv3(x, w, np.array([0.5, 0.0]))
    ref0 = np.zeros((5, 6, 7))
    ref0[:-1, 1:, :] = x[1:, :-1, :, 1]
    np.testing.assert_allclose(y[..., 0], ref0 + 0.5)
    np.testing.assert_allclose(y[..., 1], 3.0 * x[..., 0])


def test_conv_backward_is_the_adjoint():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(4, 5, 6, 3))
    w = rng.normal(size=(2, 27, 3))
    gy = rng.normal(size=(4, 5, 6, 2))
    gx, gw, gb = U.conv3_backward(x, w, gy)
    # <conv(x), gy> is linear in x and in w
    lhs = np.sum(U.conv3(x, w, np.zeros(2)) * gy)
    assert abs(lhs - np.sum(gx * x)) < 1e-9 * max(1.0, abs(lhs))
    assert abs(lhs - np.sum(gw * w)) < 1e-9 * max(1.0, abs(lhs))
    np.testing.assert_allclose(gb, gy.reshape(-1, 2).sum(0))


def test_pool_and_upsample_are_adjoint():
    rng = np.random.default_rng(2)
    a = rng.normal(size=(4, 6, 8, 3))
    b = rng.normal(size=(2, 3, 4, 3))
    assert abs(np.sum(U.pool2(a) * b) - np.sum(a * U.up2(b)) / 8.0) < 1e-12
    np.testing.assert_allclose(U.pool2_backward(b), U.up2(b) / 8.0)
    np.testing.assert_allclose(U.up2_backward(U.up2(b)), 8.0 * b)


def test_parameter_count():
    # 880 + 6928 + 13856 + 27680 + 55360 + 110656 + 82976 + 27680 + 20752 + 6928 + 17
    assert U.parameter_count() == len(U.init_parameters(0)) == 353_713


@pytest.mark.parametrize("shape", [(4, 4, 4), (8, 4, 4)])
def test_gradient_matches_central_differences(shape):
    rng = np.random.default_rng(3)
    params = U.init_parameters(7).astype(np.float64)
    params[-17:-1] = rng.normal(0.0, 0.3, 16)       # a head that passes signal back
    params[-1] = 0.1
    # nonzero biases so that some ReLUs sit on either side
    at = 0
    for _, ci, co in U.LAYERS:
        at += co * 27 * ci
        params[at:at + co] = rng.normal(0.0, 0.05, co)
        at += co
    x = rng.uniform(0.0, 1.0, shape)
    d = rng.normal(0.0, 0.5, shape)
    target = x + rng.normal(0.0, 0.1, shape)
    L, g = U.loss_and_gradient(params, x, target, d)
    # two entries of every layer's weights, its bias, and the head
    picks = []
    at = 0
    for _, ci, co in U.LAYERS:
        n = co * 27 * ci
        picks += list(at + rng.choice(n, 2, replace=False)) + [at + n + int(rng.integers(co))]
        at += n + co
    picks += [at + 3, at + 16]
    h = 1e-6
    for i in picks:
        p1, p0 = params.copy(), params.copy()
        p1[i] += h
        p0[i] -= h
        fd = (U.loss_and_gradient(p1, x, target, d)[0] - U.loss_and_gradient(p0, x, target, d)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])
