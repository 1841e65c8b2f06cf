"""The block-iterative reconstruction on the GPU (etn_reconstruction_*) against
the oracle's loop: per volume after N iterations, relative L2 <= 1e-5 in FP32
(the FP32 parity bound), and the reference's own convergence bounds."""
import numpy as np
import pytest

import ettention_amd as E
from ettention_amd import api as A
from ettention_amd import synthetic as S
from oracle import oracle as O
from tests._util import load_golden, rel_l2

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
TOL = 1e-5

TEST_BASE = E.ScannerGeometry.make((-64, -64, -128), (-64, -64, 128), (0.5, 0, 0), (0, 0.5, 0), (256, 256))


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def og(g):
    return O.Scanner.make(g.source, g.detector, g.horizontal_pitch, g.vertical_pitch, g.resolution)


def ov(v):
    return O.Volume.make(tuple(v.resolution), tuple(v.bbox_min), tuple(v.bbox_max), v.subvolume_count)


@pytest.fixture(scope="module")
def synthetic():
    s = load_golden("synthetic_3_parallel")
    return np.ascontiguousarray(s["ground_truth"]), s["angles"], np.ascontiguousarray(s["projections"])


def gpu_recon(vg, base, tilts, projs, iterations, **kw):
    r = A.BlockIterativeReconstructionOperator(volume=vg, base_geometry=base, tilt_angles=tilts,
                                               projections=dev(projs), iterations=iterations, **kw)
    r.run()
    return r, r.get_reconstructed_volume(dtype=torch.float32).cpu().numpy()


def oracle_recon(vg, base, tilts, projs, iterations, algorithm="sart", block_size=0, lam=0.3,
                 projection_iteration="random", use_long_object_compensation=False, fp_samples=1, bp_samples=1,
                 skip_projections=(), voxel_type="float", **_):
    B = {"sart": 1, "sirt": 0xFFFFFFFF}.get(algorithm, block_size)
    vt = {"float": O.VOXEL_F32, "half": O.VOXEL_F16, "bfloat16": O.VOXEL_BF16}[voxel_type]
    spec = O.ReconSpec(ov(vg), og(base), tilts, projs, block_size=B, order=projection_iteration, lam=lam,
                       use_loc=use_long_object_compensation, fp_samples=fp_samples, bp_samples=bp_samples,
                       skip=skip_projections, voxel_type=vt)
    shape = (vg.resolution[2], vg.resolution[1], vg.resolution[0])
    vol = np.zeros(shape, np.float32 if vt == O.VOXEL_F32 else np.uint16)
    spec.run(vol, iterations)
    return spec, vol


def as_f32(raw, voxel_type):
    if voxel_type == "half":
        return raw.view(np.float16).astype(np.float32)
    if voxel_type == "bfloat16":
        return (raw.astype(np.uint32) << 16).view(np.float32)
    return raw


CASES = [
    dict(algorithm="sart", lam=0.1, projection_iteration="random"),
    dict(algorithm="sirt", lam=1.0, projection_iteration="identity"),
    dict(algorithm="blockiterative", block_size=10, lam=1.0, projection_iteration="maxangle",
         use_long_object_compensation=True),
    dict(algorithm="blockiterative", block_size=7, lam=0.5, projection_iteration="buckets", skip_projections=(3, 90)),
    dict(algorithm="sart", lam=0.3, projection_iteration="identity", fp_samples=4, bp_samples=4),
]


@pytest.mark.parametrize("case", CASES)
def test_reconstruction_matches_oracle(synthetic, case):
    gt, tilts, projs = synthetic
    vg = E.VolumeGeometry.make((64, 64, 64), subvolumes=4)
    sub = slice(None, None, 3)          # 60 of the 180 golden projections keeps the oracle fast
    r, got = gpu_recon(vg, TEST_BASE, tilts[sub], projs[sub], 2, **case)
    _, ref = oracle_recon(vg, TEST_BASE, tilts[sub], projs[sub], 2, **case)
    assert rel_l2(got, ref) <= TOL


@pytest.mark.parametrize("voxel_type", ["half", "bfloat16"])
def test_reduced_precision_storage_matches_oracle(synthetic, voxel_type):
    gt, tilts, projs = synthetic
    vg = E.VolumeGeometry.make((64, 64, 64))
    sub = slice(None, None, 6)
    r, got = gpu_recon(vg, TEST_BASE, tilts[sub], projs[sub], 2, algorithm="sart", lam=0.3,
                       projection_iteration="identity", voxel_type=voxel_type,
                       flags=E.ETN_RECON_FLAG_EXACT_ARITHMETIC)
    _, raw = oracle_recon(vg, TEST_BASE, tilts[sub], projs[sub], 2, algorithm="sart", lam=0.3,
                          projection_iteration="identity", voxel_type=voxel_type)
    # the same FP32 arithmetic and the same rounding to storage as the oracle
    assert rel_l2(got, as_f32(raw, voxel_type)) <= 1e-4
    # default arithmetic: a rounding difference can move a stored voxel by one 16-bit ulp
    _, fast = gpu_recon(vg, TEST_BASE, tilts[sub], projs[sub], 2, algorithm="sart", lam=0.3,
                        projection_iteration="identity", voxel_type=voxel_type)
    assert rel_l2(fast, as_f32(raw, voxel_type)) <= 1e-3
    spec64 = O.ReconSpec(ov(vg), og(TEST_BASE), tilts[sub], projs[sub], block_size=1, order="identity", lam=0.3)
    d = spec64.run_f64(np.zeros((64, 64, 64), np.float64), 2)
    err = rel_l2(got, d)
    assert err < (2e-3 if voxel_type == "half" else 1.5e-2), err


def test_graph_replay_equals_eager_launches(synthetic):
    gt, tilts, projs = synthetic
    vg = E.VolumeGeometry.make((64, 64, 64), subvolumes=2)
    sub = slice(None, None, 5)
    kw = dict(algorithm="blockiterative", block_size=4, lam=0.7, projection_iteration="random")
    r1, a = gpu_recon(vg, TEST_BASE, tilts[sub], projs[sub], 3, **kw)
    r2, b = gpu_recon(vg, TEST_BASE, tilts[sub], projs[sub], 3, flags=E.ETN_RECON_FLAG_NO_GRAPH, **kw)
    np.testing.assert_array_equal(a, b)
    # replaying the recorded iteration again, interleaved with host work, keeps matching
    r1.run_iterations(2)
    r2.run_iterations(2)
    np.testing.assert_array_equal(r1.get_reconstructed_volume().cpu().numpy(),
                                  r2.get_reconstructed_volume().cpu().numpy())


def test_residual_norm_matches_oracle(synthetic):
    gt, tilts, projs = synthetic
    vg = E.VolumeGeometry.make((64, 64, 64))
    sub = slice(None, None, 10)
    r, got = gpu_recon(vg, TEST_BASE, tilts[sub], projs[sub], 1, algorithm="sart", lam=0.2,
                       projection_iteration="identity", skip_projections=(2,))
    spec, ref = oracle_recon(vg, TEST_BASE, tilts[sub], projs[sub], 1, algorithm="sart", lam=0.2,
                             projection_iteration="identity", skip_projections=(2,))
    n_gpu = r.residual_norm()
    n_ref = spec.residual_norm(ref)
    assert abs(n_gpu - n_ref) <= 1e-6 * n_ref


def test_init_volume_and_single_iteration(synthetic):
    gt, tilts, projs = synthetic
    vg = E.VolumeGeometry.make((64, 64, 64))
    sub = slice(None, None, 12)
    r = A.BlockIterativeReconstructionOperator(volume=vg, base_geometry=TEST_BASE, tilt_angles=tilts[sub],
                                               projections=dev(projs[sub]), algorithm="sart", lam=0.2,
                                               projection_iteration="identity")
    r.set_volume(dev(gt))
    r.run_single_iteration()
    got = r.get_reconstructed_volume(dtype=torch.float32).cpu().numpy()
    spec = O.ReconSpec(ov(vg), og(TEST_BASE), tilts[sub], projs[sub], block_size=1, order="identity", lam=0.2)
    ref = spec.run(gt.copy(), 1)
    assert rel_l2(got, ref) <= TOL


def _max_slice_rms_z(vol, gt):
    d = vol.astype(np.float64) - gt
    return float(np.sqrt((d ** 2).mean(axis=(1, 2))).max())


@pytest.mark.parametrize("case,iterations,bound", [
    (dict(algorithm="sart", lam=0.1, fp_samples=4, bp_samples=9), 2, 0.7),     # sart.xml
    (dict(algorithm="sirt", lam=1.0), 25, 0.7),                                   # sirt.xml
    (dict(algorithm="blockiterative", block_size=10, lam=1.0), 10, 0.2),          # blockIterative.xml
])
def test_reference_convergence_bounds(synthetic, case, iterations, bound):
    # SyntheticReconstructionTest (SyntheticReconstruction.cpp:132-171): 180 projections,
    # 4 subvolumes, default random order, RMS to the ground truth.
    gt, tilts, projs = synthetic
    vg = E.VolumeGeometry.make((64, 64, 64), subvolumes=4)
    r, got = gpu_recon(vg, TEST_BASE, tilts, projs, iterations, projection_iteration="random", **case)
    assert _max_slice_rms_z(got, gt) <= bound


def test_config1_plain_sart_matches_oracle():
    # Plain SART, 128x128x64 ellipsoid phantom, 61 tilts +-60 by 2 deg, 128^2 detector,
    # Poisson noise at 1e4 counts, lambda 0.3, 10 iterations, sequential order.
    shape = (64, 128, 128)
    truth = S.ellipsoid_phantom(shape, seed=1001)
    vg = E.VolumeGeometry.make((128, 128, 64))
    base = E.ScannerGeometry.mrc_base(128, 128)
    tilts = S.tilt_series(60.0, 2.0)
    clean = np.zeros((len(tilts), 128, 128), np.float32)
    truth_t = dev(truth)
    for i, t in enumerate(tilts):
        p = torch.zeros((128, 128), dtype=torch.float32, device="cuda")
        l = torch.zeros_like(p)
        A.forward_project(vg, truth_t, A.rotate(base, float(t)), p, l)
        clean[i] = p.cpu().numpy()
    projs = S.add_poisson_noise(clean, 1e4, seed=1001)
    kw = dict(algorithm="sart", lam=0.3, projection_iteration="identity")
    r, got = gpu_recon(vg, base, tilts, projs, 10, **kw)
    spec, ref = oracle_recon(vg, base, tilts, projs, 10, **kw)
    assert rel_l2(got, ref) <= TOL
    # FP64 reference of the same loop: the FP32 path stays close
    d = O.ReconSpec(ov(vg), og(base), tilts, projs, block_size=1, order="identity", lam=0.3).run_f64(
        np.zeros(shape, np.float64), 10)
    assert rel_l2(got, d) < 1e-4
    assert r.residual_norm() < 0.5 * np.linalg.norm(projs.astype(np.float64))
