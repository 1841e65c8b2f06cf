"""Parity of the benchmarked path over the benchmark's iteration count.

The benchmark runs the fused kernels with fast arithmetic (combined bilinear
weights, hardware reciprocal, FMAs).  Here, on the benchmark's own inputs
(seeded ellipsoid phantom, forward-projected, 2 % Gaussian noise, random
projection order), fast arithmetic runs 20 iterations beside exact arithmetic,
which is bit-identical to the generic kernels (checked at the end) and through
them pinned to the oracle.  Bounds: relative L2 <= 1e-5 per volume after
iterations 1, 5, 10 and 20 (the FP32 parity bound), and per projection for the
forward projections of the final volumes (every tilt, generic kernel)."""
import numpy as np
import pytest

import ettention_amd as E
from ettention_amd import api as A
from ettention_amd import synthetic as S

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
TOL = 1e-5
CHECKPOINTS = (1, 5, 10, 20)


def rel(a, b):
    return float(torch.linalg.vector_norm((a - b).double()) / torch.linalg.vector_norm(b.double()))


@pytest.mark.parametrize("workload", [
    # SART 512^3, 121 tilts +-60 deg by 1 deg, 512^2 detector, random order (the metric)
    dict(res=(512, 512, 512), algorithm="sart", block_size=0, seed=1001, beads=0),
    # block-iterative B = 8 on 512 x 512 x 256 with 200 beads
    dict(res=(512, 512, 256), algorithm="blockiterative", block_size=8, seed=1002, beads=200),
], ids=["sart-512cubed", "block8-512x512x256"])
def test_fast_arithmetic_stays_within_fp32_bound_over_20_iterations(workload):
    nx, ny, nz = workload["res"]
    print("workload", workload, flush=True)
    vg = E.VolumeGeometry.make((nx, ny, nz))
    base = E.ScannerGeometry.mrc_base(512, 512)
    tilts = S.tilt_series(60.0, 1.0)
    phantom = S.ellipsoid_phantom((nz, ny, nx), seed=workload["seed"], beads=workload["beads"], device="cuda")
    projs = torch.zeros((len(tilts), 512, 512), dtype=torch.float32, device="cuda")
    scratch = torch.zeros((512, 512), dtype=torch.float32, device="cuda")
    for i, t in enumerate(tilts):
        scratch.zero_()
        A.forward_project(vg, phantom, A.rotate(base, float(t)), projs[i], scratch)
    projs = S.add_gaussian_noise(projs, 0.02, seed=7)
    del phantom
    torch.cuda.synchronize()
    kw = dict(volume=vg, base_geometry=base, tilt_angles=tilts, projections=projs, algorithm=workload["algorithm"],
              block_size=workload["block_size"], lam=0.3, projection_iteration="random")
    fast = A.BlockIterativeReconstructionOperator(**kw)
    exact = A.BlockIterativeReconstructionOperator(flags=E.ETN_RECON_FLAG_EXACT_ARITHMETIC, **kw)
    assert fast.uses_fused_path() and exact.uses_fused_path()
    done = 0
    for n in CHECKPOINTS:
        fast.run_iterations(n - done)
        exact.run_iterations(n - done)
        done = n
        a, b = fast.get_reconstructed_volume(), exact.get_reconstructed_volume()
        d = rel(a, b)
        print("iteration %d: fast vs exact relative L2 %.3e" % (n, d), flush=True)
        assert d <= TOL, (n, d)
    # per projection: the forward projections of the two final volumes
    pa = torch.zeros((512, 512), dtype=torch.float32, device="cuda")
    pb = torch.zeros_like(pa)
    la = torch.zeros_like(pa)
    worst = 0.0
    for t in tilts:
        g = A.rotate(base, float(t))
        for img in (pa, pb, la):
            img.zero_()
        A.forward_project(vg, a, g, pa, la)
        A.forward_project(vg, b, g, pb, la)
        worst = max(worst, rel(pa, pb))
    torch.cuda.synchronize()
    print("worst per-projection relative L2 %.3e" % worst, flush=True)
    assert worst <= TOL
    del fast
    # exact arithmetic is the generic kernels' arithmetic: bit-identical after 20 iterations
    generic = A.BlockIterativeReconstructionOperator(flags=E.ETN_RECON_FLAG_GENERIC_KERNELS, **kw)
    generic.run_iterations(CHECKPOINTS[-1])
    assert torch.equal(generic.get_reconstructed_volume(), b)
