"""The correction network's HIP kernels (bf16 implicit-GEMM convolutions on the
matrix cores, their backward passes, pooling, head, loss, Adam) against the
FP64 restatement in oracle/unet_oracle.py.

Tolerances are those of bf16 storage (8-bit mantissa) through ten layers:
the prediction's correction within 3% relative L2, each layer's weight and
bias gradient within 6% relative L2 (10% for layers whose level has fewer
than 100 voxels, where a gradient sums few bf16 products) and cosine
similarity >= 0.995."""
import numpy as np
import pytest

from ettention_amd import learned as Lrn
from oracle import unet_oracle as U

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def bf16_round(a):
    """Round float32 values to bf16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def test_parameter_layout_matches_the_oracle():
    assert Lrn.parameter_count() == U.parameter_count()
    assert Lrn.LAYERS == U.LAYERS


def setup_case(shape, seed):
    rng = np.random.default_rng(seed)
    params = U.init_parameters(seed).astype(np.float32)
    params[-17:-1] = rng.normal(0.0, 0.3, 16)
    params[-1] = 0.05
    at = 0
    for _, ci, co in U.LAYERS:
        at += co * 27 * ci
        params[at:at + co] = rng.normal(0.0, 0.05, co)
        at += co
    x = rng.uniform(0.0, 1.0, shape).astype(np.float32)
    d = rng.normal(0.0, 0.05, shape).astype(np.float32)
    target = (x + rng.normal(0.0, 0.1, shape)).astype(np.float32)
    return params, x, d, target


def oracle_params(params):
    """Convolution weights as the kernels use them (bf16); biases and head stay fp32."""
    p = params.astype(np.float64).copy()
    at = 0
    for _, ci, co in U.LAYERS:
        n = co * 27 * ci
        p[at:at + n] = bf16_round(params[at:at + n])
        at += n + co
    return p


G = Lrn.UPDATE_GAIN


@pytest.mark.parametrize("shape", [(16, 16, 16), (8, 12, 20)])
def test_forward_matches_oracle(shape):
    params, x, d, _ = setup_case(shape, 11)
    net = Lrn.CorrectionNet(shape, params)
    xd, dd = torch.from_numpy(x).cuda(), torch.from_numpy(d).cuda()
    y = net.predict(xd, dd, 1.0).cpu().numpy().astype(np.float64)
    ref, _ = U.forward(oracle_params(params), bf16_round(x), bf16_round(d * G))
    d_gpu, d_ref = y - x, ref - bf16_round(x)
    err = np.linalg.norm(d_gpu - d_ref) / np.linalg.norm(d_ref)
    assert err < 3e-2, err


def test_scale_only_rescales():
    shape = (8, 8, 8)
    params, x, d, _ = setup_case(shape, 12)
    net = Lrn.CorrectionNet(shape, params)
    xd = torch.from_numpy(x).cuda()
    # with scale s the network sees (x s, 0) and the correction is divided by s
    s = 4.0
    y = net.predict(xd, None, s).cpu().numpy().astype(np.float64)
    ref, _ = U.forward(oracle_params(params), bf16_round(x * s))
    d_ref = (ref - bf16_round(x * s)) / s
    err = np.linalg.norm((y - x) - d_ref) / np.linalg.norm(d_ref)
    assert err < 3e-2, err


@pytest.mark.parametrize("shape", [(16, 16, 16), (8, 12, 20)])
def test_gradient_matches_oracle(shape):
    params, x, d, target = setup_case(shape, 13)
    net = Lrn.CorrectionNet(shape, params)
    loss, g = net.gradient(torch.from_numpy(x).cuda(), torch.from_numpy(d).cuda(), torch.from_numpy(target).cuda(),
                           1.0)
    ref_loss, ref = U.loss_and_gradient(oracle_params(params), bf16_round(x), target.astype(np.float64),
                                        bf16_round(d * G))
    assert abs(loss - ref_loss) <= 2e-2 * ref_loss
    level = {"c1a": 0, "c1b": 0, "c2a": 1, "c2b": 1, "c3a": 2, "c3b": 2, "c4a": 1, "c4b": 1, "c5a": 0, "c5b": 0}
    at = 0
    for name, ci, co in U.LAYERS:
        voxels = np.prod([n >> level[name] for n in shape])
        tol = 6e-2 if voxels >= 100 else 1e-1
        for part, n in (("w", co * 27 * ci), ("b", co)):
            a, b = g[at:at + n].astype(np.float64), ref[at:at + n]
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
            assert rel < tol and cos > 0.995, (name, part, rel, cos)
            at += n
    a, b = g[at:].astype(np.float64), ref[at:]
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 3e-2


def test_gradient_is_deterministic():
    shape = (16, 16, 16)
    params, x, d, target = setup_case(shape, 14)
    net = Lrn.CorrectionNet(shape, params)
    xd, dd, td = torch.from_numpy(x).cuda(), torch.from_numpy(d).cuda(), torch.from_numpy(target).cuda()
    l1, g1 = net.gradient(xd, dd, td, 1.0)
    l2, g2 = net.gradient(xd, dd, td, 1.0)
    assert l1 == l2
    np.testing.assert_array_equal(g1, g2)


def test_adam_steps_reduce_the_loss():
    shape = (16, 16, 16)
    params, x, d, target = setup_case(shape, 15)
    net = Lrn.CorrectionNet(shape, params)
    xd, dd, td = torch.from_numpy(x).cuda(), torch.from_numpy(d).cuda(), torch.from_numpy(target).cuda()
    losses = [net.train_step(xd, dd, td, 1.0, 1e-3) for _ in range(30)]
    assert losses[-1] < 0.8 * losses[0], losses
    # the parameters moved, and the first Adam step is lr * sign(g) per weight
    p = net.parameters()
    assert np.abs(p - params).max() > 0
