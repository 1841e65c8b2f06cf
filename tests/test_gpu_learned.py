"""What the trained correction network gains over SART without it.

The network is trained by the library's own kernels on samples of seeded 64^3
phantoms -- bf16 SART iterate 3 and its last update as input, the FP32 SART
minimum-residual iterate as target -- and applied to a held-out phantom.  Plain SART is the arbiter: the jump must lower the projection
residual ||Ax - b|| (FP64, the reconstruction's own norm) below that of the
iterate it started from, and the acceptance rule must keep plain SART's
result whenever the jump loses to the iterations plain SART runs in the same
time."""
import numpy as np
import pytest

from ettention_amd import api as A
from ettention_amd import learned as L

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SHAPE = (64, 64, 64)
TILTS = 61


def norm_of(case, volume):
    r = A.BlockIterativeReconstructionOperator(algorithm="sart", **case)
    r.set_volume(volume.contiguous())
    return r.residual_norm()


@pytest.fixture(scope="module")
def trained():
    samples = []
    for i in range(8):
        case, _ = L.make_case(SHAPE, TILTS, seed=6001 + i)
        sample, _ = L.training_sample(case)
        samples.append(sample)
    net = L.CorrectionNet(SHAPE, seed=5000)
    losses = L.train(net, samples, 1.0, 300, lr=1e-3)
    return net, losses


def test_training_lowers_the_loss(trained):
    _, losses = trained
    print("loss first 8 %.4e last 20 %.4e" % (np.mean(losses[:8]), np.mean(losses[-20:])))
    assert np.mean(losses[-20:]) < 0.8 * np.mean(losses[:8]), (losses[:8], losses[-20:])


def test_jump_lowers_the_projection_residual(trained):
    net, _ = trained
    case, _ = L.make_case(SHAPE, TILTS, seed=6000)
    x2, x3 = L.sart_iterates(case, [2, 3], voxel_type="bfloat16")
    y = net.predict(x3, x3 - x2, 1.0)
    before, after = norm_of(case, x3), norm_of(case, y)
    print("residual of iterate 3 %.6e, after the jump %.6e" % (before, after))
    assert after < before, (before, after)


def test_time_to_target_and_acceptance_rule(trained):
    net, _ = trained
    case, _ = L.make_case(SHAPE, TILTS, seed=6000)
    res = L.time_to_target(case, net, 1.0, k_max=50, k0=3)
    assert 3 <= res["plain_best_iteration"] <= 50 and res["plain_s"] > 0
    assert res["accepted"] == (res["jump_norm"] < res["plain_norm_at_jump_time"])
    norms = res["plain_norms"]
    assert min(norms) == res["target_norm"] < norms[0]
    if res["reached"]:
        assert res["speedup"] > 0
