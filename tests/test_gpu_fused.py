"""The fused plane kernels (BP of block b-1 fused with the FP of block b, one
workgroup per group of volume planes) against the generic kernels and the oracle.

With ETN_RECON_FLAG_EXACT_ARITHMETIC the fused kernels repeat the generic
kernels' float operations for every ray and voxel, so the volumes must be
bit-identical.  The default (fast) arithmetic performs the same operations with
combined bilinear weights, a hardware reciprocal and fused multiply-adds; it
must stay within the FP32 parity bound, relative L2 <= 1e-5, of the generic
kernels and of the oracle (1e-3 for 16-bit storage, where a rounding
difference can flip a stored voxel by one 16-bit ulp)."""
import numpy as np
import pytest

import ettention_amd as E
from ettention_amd import api as A
from ettention_amd import synthetic as S
from oracle import oracle as O
from tests._util import load_golden, rel_l2

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TEST_BASE = E.ScannerGeometry.make((-64, -64, -128), (-64, -64, 128), (0.5, 0, 0), (0, 0.5, 0), (256, 256))


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def og(g):
    return O.Scanner.make(g.source, g.detector, g.horizontal_pitch, g.vertical_pitch, g.resolution)


def ov(v):
    return O.Volume.make(tuple(v.resolution), tuple(v.bbox_min), tuple(v.bbox_max), v.subvolume_count)


def run_pair(vg, base, tilts, projs_t, iterations, **kw):
    """(exact fused, generic, fast fused) runs: (operator, float32 volume) each."""
    out = []
    for flags in (E.ETN_RECON_FLAG_EXACT_ARITHMETIC, E.ETN_RECON_FLAG_GENERIC_KERNELS, 0):
        r = A.BlockIterativeReconstructionOperator(volume=vg, base_geometry=base, tilt_angles=tilts,
                                                   projections=projs_t, iterations=iterations, flags=flags, **kw)
        assert r.uses_fused_path() == (flags != E.ETN_RECON_FLAG_GENERIC_KERNELS)
        r.run()
        r.synchronize()
        out.append((r, r.get_reconstructed_volume(dtype=torch.float32).cpu().numpy()))
    return out


@pytest.fixture(scope="module")
def synthetic():
    s = load_golden("synthetic_3_parallel")
    return np.ascontiguousarray(s["ground_truth"]), s["angles"], np.ascontiguousarray(s["projections"])


@pytest.mark.parametrize("case", [
    dict(algorithm="sart", lam=0.1, projection_iteration="random"),
    dict(algorithm="blockiterative", block_size=10, lam=1.0, projection_iteration="maxangle",
         use_long_object_compensation=True),
    dict(algorithm="sirt", lam=1.0, projection_iteration="identity"),
    dict(algorithm="blockiterative", block_size=7, lam=0.5, projection_iteration="buckets", skip_projections=(3, 90)),
    dict(algorithm="sart", lam=0.3, projection_iteration="identity", voxel_type="half"),
    dict(algorithm="blockiterative", block_size=3, lam=0.3, projection_iteration="random", voxel_type="bfloat16"),
])
def test_fused_equals_generic_on_reference_dataset(synthetic, case):
    # two detector rows per volume plane (pitch 0.5), tilts -90..89 (rays change z sign)
    gt, tilts, projs = synthetic
    sub = slice(None, None, 3)
    vg = E.VolumeGeometry.make((64, 64, 64))
    (rf, fused), (rg, generic), (_, fast) = run_pair(vg, TEST_BASE, tilts[sub], dev(projs[sub]), 2, **case)
    np.testing.assert_array_equal(fused, generic)
    assert rf.residual_norm() == rg.residual_norm()
    assert rel_l2(fast, generic) <= (1e-5 if case.get("voxel_type", "float") == "float" else 1e-3)
    if case.get("voxel_type", "float") == "float":
        B = {"sart": 1, "sirt": 0xFFFFFFFF}.get(case["algorithm"], case.get("block_size", 1))
        spec = O.ReconSpec(ov(vg), og(TEST_BASE), tilts[sub], projs[sub], block_size=B,
                           order=case["projection_iteration"], lam=case["lam"],
                           use_loc=case.get("use_long_object_compensation", False),
                           skip=case.get("skip_projections", ()))
        ref = spec.run(np.zeros((64, 64, 64), np.float32), 2)
        assert rel_l2(fused, ref) <= 1e-5
        assert rel_l2(fast, ref) <= 1e-5


@pytest.mark.parametrize("shape,det", [((64, 48, 40), (64, 48)), ((96, 80, 72), (96, 80)), ((130, 66, 33), (130, 66))])
def test_fused_equals_generic_on_mrc_geometry(shape, det):
    # MRC stack geometry (unit pitch, one detector row per plane), odd sizes
    nx, ny, nz = shape
    truth = S.ellipsoid_phantom((nz, ny, nx), seed=77)
    vg = E.VolumeGeometry.make(shape)
    base = E.ScannerGeometry.mrc_base(*det)
    tilts = S.tilt_series(60.0, 4.0)
    truth_t = dev(truth)
    projs = torch.zeros((len(tilts), det[1], det[0]), dtype=torch.float32, device="cuda")
    scratch = torch.zeros((det[1], det[0]), dtype=torch.float32, device="cuda")
    for i, t in enumerate(tilts):
        A.forward_project(vg, truth_t, A.rotate(base, float(t)), projs[i], scratch)
    (rf, fused), (rg, generic), (_, fast) = run_pair(vg, base, tilts, projs, 3, algorithm="sart", lam=0.3,
                                                     projection_iteration="random")
    np.testing.assert_array_equal(fused, generic)
    spec = O.ReconSpec(ov(vg), og(base), tilts, projs.cpu().numpy(), block_size=1, order="random", lam=0.3)
    ref = spec.run(np.zeros((nz, ny, nx), np.float32), 3)
    assert rel_l2(fused, ref) <= 1e-5
    assert rel_l2(fast, ref) <= 1e-5


def test_rounds_of_plane_groups(synthetic):
    # 1024 one-plane groups per launch, more than the device holds at once:
    # the rounds are dispatched interleaved, which must not change a voxel
    nx, ny, nz = 512, 1024, 8
    truth = S.ellipsoid_phantom((nz, ny, nx), seed=78)
    vg = E.VolumeGeometry.make((nx, ny, nz))
    base = E.ScannerGeometry.mrc_base(nx, ny)
    tilts = S.tilt_series(60.0, 15.0)
    truth_t = dev(truth)
    projs = torch.zeros((len(tilts), ny, nx), dtype=torch.float32, device="cuda")
    scratch = torch.zeros((ny, nx), dtype=torch.float32, device="cuda")
    for i, t in enumerate(tilts):
        A.forward_project(vg, truth_t, A.rotate(base, float(t)), projs[i], scratch)
    vols = {}
    for name, flags, tile in (("exact", E.ETN_RECON_FLAG_EXACT_ARITHMETIC, (1, 8)),
                              ("generic", E.ETN_RECON_FLAG_GENERIC_KERNELS, None),
                              ("fast", 0, (1, 8)), ("fast_two_planes", 0, (2, 4))):
        r = A.BlockIterativeReconstructionOperator(volume=vg, base_geometry=base, tilt_angles=tilts, projections=projs,
                                                   iterations=2, algorithm="sart", lam=0.3,
                                                   projection_iteration="random",
                                                   flags=flags | E.ETN_RECON_FLAG_NO_AUTOTUNE)
        if tile is not None:
            r.set_workgroup_threads(512)
            r.set_tile(*tile)
            assert r.tile() == tile
        r.run()
        r.synchronize()
        vols[name] = r.get_reconstructed_volume(dtype=torch.float32).cpu().numpy()
    np.testing.assert_array_equal(vols["exact"], vols["generic"])
    np.testing.assert_array_equal(vols["fast"], vols["fast_two_planes"])
    assert rel_l2(vols["fast"], vols["generic"]) <= 1e-5


@pytest.mark.parametrize("kind", ["f32_mask", "exact", "bf16", "block4"])
def test_wide_planes_take_the_large_lds_band(kind):
    # 1536 rays per plane: four per thread, one 512-thread workgroup per CU,
    # so the launches take the 160 KiB array with taller bands (the LDS back
    # projection); against the generic kernels
    nx, ny, nz = 1536, 4, 96
    truth = S.ellipsoid_phantom((nz, ny, nx), seed=80)
    vg = E.VolumeGeometry.make((nx, ny, nz))
    base = E.ScannerGeometry.mrc_base(nx, ny)
    tilts = S.tilt_series(60.0, 10.0)
    truth_t = dev(truth)
    projs = torch.zeros((len(tilts), ny, nx), dtype=torch.float32, device="cuda")
    scratch = torch.zeros((ny, nx), dtype=torch.float32, device="cuda")
    for i, t in enumerate(tilts):
        A.forward_project(vg, truth_t, A.rotate(base, float(t)), projs[i], scratch)
    kw = dict(volume=vg, base_geometry=base, tilt_angles=tilts, projections=projs, iterations=2, lam=0.3,
              algorithm="sart", projection_iteration="random")
    flags = 0
    if kind == "f32_mask":
        kw["mask"] = dev(seeded_mask((nx, ny, nz), 81))
    if kind == "exact":   # (with a mask, exact records of 1536 columns leave no room for a band row)
        flags = E.ETN_RECON_FLAG_EXACT_ARITHMETIC
    if kind == "bf16":
        kw["voxel_type"] = "bfloat16"
    if kind == "block4":
        kw.update(algorithm="blockiterative", block_size=4)
    vols = {}
    for name, f in (("fused", flags), ("generic", E.ETN_RECON_FLAG_GENERIC_KERNELS)):
        r = A.BlockIterativeReconstructionOperator(flags=f, **kw)
        assert r.uses_fused_path() == (name == "fused")
        r.run()
        r.synchronize()
        vols[name] = r.get_reconstructed_volume(dtype=torch.float32).cpu().numpy()
    if kind == "exact":
        np.testing.assert_array_equal(vols["fused"], vols["generic"])
    else:
        assert rel_l2(vols["fused"], vols["generic"]) <= (1e-3 if kind == "bf16" else 1e-5)


def test_ineligible_configurations_use_generic_kernels(synthetic):
    gt, tilts, projs = synthetic
    sub = slice(None, None, 20)
    p = dev(projs[sub])
    for kw in (dict(fp_samples=5), dict(bp_samples=25),
               dict(fp_samples=2, flags=E.ETN_RECON_FLAG_EXACT_ARITHMETIC,
                    volume=E.VolumeGeometry.make((64, 64, 64), subvolumes=2)),
               dict(fp_samples=2, volume=E.VolumeGeometry.make((64, 64, 64), subvolumes=4)),
               dict(volume=E.VolumeGeometry.make((64, 64, 64), subvolumes=2), flags=E.ETN_RECON_FLAG_EXACT_ARITHMETIC),
               dict(x_axis_tilt=5.0),
               # a mask one byte off 4-byte alignment (the fused kernels read 4 bytes at a time)
               dict(mask=torch.full((64 * 64 * 64 + 4,), 255, dtype=torch.uint8, device="cuda")[1:1 + 64 ** 3]
                    .view(64, 64, 64))):
        args = dict(volume=E.VolumeGeometry.make((64, 64, 64)), base_geometry=TEST_BASE, tilt_angles=tilts[sub],
                    projections=p, algorithm="sart")
        args.update(kw)
        r = A.BlockIterativeReconstructionOperator(**args)
        assert not r.uses_fused_path(), kw


@pytest.mark.parametrize("workload", [
    dict(res=(512, 512, 512), algorithm="sart", block_size=0),
    dict(res=(512, 512, 256), algorithm="blockiterative", block_size=8),
])
def test_full_size_fused_equals_generic(workload):
    # the benchmark shapes: 512^3 SART and 512x512x256 block-iterative B=8, 121 tilts
    nx, ny, nz = workload["res"]
    vg = E.VolumeGeometry.make((nx, ny, nz))
    base = E.ScannerGeometry.mrc_base(512, 512)
    tilts = S.tilt_series(60.0, 1.0)
    phantom = S.ellipsoid_phantom((nz, ny, nx), seed=1001, device="cuda")
    projs = torch.zeros((len(tilts), 512, 512), dtype=torch.float32, device="cuda")
    scratch = torch.zeros((512, 512), dtype=torch.float32, device="cuda")
    for i, t in enumerate(tilts):
        A.forward_project(vg, phantom, A.rotate(base, float(t)), projs[i], scratch)
    del phantom
    vols = []
    for flags in (E.ETN_RECON_FLAG_EXACT_ARITHMETIC, E.ETN_RECON_FLAG_GENERIC_KERNELS, 0):
        r = A.BlockIterativeReconstructionOperator(volume=vg, base_geometry=base, tilt_angles=tilts,
                                                   projections=projs, algorithm=workload["algorithm"],
                                                   block_size=workload["block_size"], lam=0.3,
                                                   projection_iteration="random", flags=flags)
        assert r.uses_fused_path() == (flags != E.ETN_RECON_FLAG_GENERIC_KERNELS)
        r.run_single_iteration()
        r.synchronize()
        vols.append(r.get_reconstructed_volume())
        del r
    assert torch.equal(vols[0], vols[1])
    diff = float(torch.linalg.vector_norm((vols[2] - vols[1]).double()) / torch.linalg.vector_norm(vols[1].double()))
    assert diff <= 1e-5, diff


def test_subvolumes_run_fused_with_fast_arithmetic(synthetic):
    """hardware.subVolumeCount > 1: the fused kernels sweep the whole volume in
    one pass, which regroups the slab-wise sums of the forward projection; the
    result stays within 1e-5 of the slab-wise generic kernels and the oracle."""
    gt, tilts, projs = synthetic
    sub = slice(None, None, 4)
    p = dev(projs[sub])
    vg = E.VolumeGeometry.make((64, 64, 64), subvolumes=4)
    kw = dict(volume=vg, base_geometry=TEST_BASE, tilt_angles=tilts[sub], projections=p, algorithm="blockiterative",
              block_size=5, lam=0.8, projection_iteration="identity", use_long_object_compensation=True)
    fast = A.BlockIterativeReconstructionOperator(iterations=2, **kw)
    assert fast.uses_fused_path()
    fast.run()
    generic = A.BlockIterativeReconstructionOperator(iterations=2, flags=E.ETN_RECON_FLAG_GENERIC_KERNELS, **kw)
    generic.run()
    a = fast.get_reconstructed_volume().cpu().numpy()
    b = generic.get_reconstructed_volume().cpu().numpy()
    assert rel_l2(a, b) <= 1e-5
    spec = O.ReconSpec(ov(vg), og(TEST_BASE), tilts[sub], projs[sub], block_size=5, order="identity", lam=0.8,
                       use_loc=True)
    ref = spec.run(np.zeros((64, 64, 64), np.float32), 2)
    assert rel_l2(a, ref) <= 1e-5


def seeded_mask(shape, seed):
    """A prior-knowledge mask: 255 inside an off-centre ellipsoid, 0 far outside,
    seeded random weights 1..254 in a shell between, (nz, ny, nx) uint8."""
    nx, ny, nz = shape
    z, y, x = np.meshgrid(np.linspace(-1, 1, nz), np.linspace(-1, 1, ny), np.linspace(-1, 1, nx), indexing="ij")
    r = np.sqrt(((x - 0.1) / 0.8) ** 2 + (y / 0.9) ** 2 + ((z + 0.05) / 0.7) ** 2)
    rng = np.random.default_rng(seed)
    m = np.where(r < 0.8, 255, np.where(r < 1.0, rng.integers(1, 255, r.shape), 0))
    return np.ascontiguousarray(m.astype(np.uint8))


@pytest.mark.parametrize("case", [
    dict(shape=(64, 64, 64), geometry="test", algorithm="blockiterative", block_size=5, lam=0.8,
         projection_iteration="maxangle", use_long_object_compensation=True),
    dict(shape=(66, 48, 40), geometry="mrc", algorithm="sart", lam=0.3, projection_iteration="random"),
    dict(shape=(64, 64, 64), geometry="test", algorithm="sirt", lam=1.0, projection_iteration="identity"),
    dict(shape=(64, 64, 64), geometry="test", algorithm="sart", lam=0.3, projection_iteration="identity",
         voxel_type="bfloat16"),
])
def test_mask_runs_fused_with_fast_arithmetic(synthetic, case):
    """USE_PRIOR_KNOWLEDGE_MASK on the fused kernels (default arithmetic): the
    ray length sums dt * weight and a voxel moves by weight * update.  Within
    1e-5 of the masked generic kernels and the masked oracle.  With bf16
    storage a rounding difference flips a stored voxel by one bf16 ulp (2^-8
    relative) and later updates carry it on; the bound 2^-9 allows such flips
    in up to a quarter of the voxels.  Odd nx (66) takes the scalar band path."""
    case = dict(case)
    shape, geometry = case.pop("shape"), case.pop("geometry")
    nx, ny, nz = shape
    vg = E.VolumeGeometry.make(shape)
    if geometry == "test":
        _, tilts, projs = synthetic
        tilts, projs = tilts[::4], np.ascontiguousarray(projs[::4])
        base = TEST_BASE
    else:
        base = E.ScannerGeometry.mrc_base(nx, ny)
        tilts = S.tilt_series(60.0, 4.0)
        truth_t = dev(S.ellipsoid_phantom((nz, ny, nx), seed=78))
        pt = torch.zeros((len(tilts), ny, nx), dtype=torch.float32, device="cuda")
        scratch = torch.zeros((ny, nx), dtype=torch.float32, device="cuda")
        for i, t in enumerate(tilts):
            scratch.zero_()
            A.forward_project(vg, truth_t, A.rotate(base, float(t)), pt[i], scratch)
        projs = pt.cpu().numpy()
    mask = seeded_mask(shape, 5)
    mask_t = dev(mask)
    vols = []
    for flags in (0, E.ETN_RECON_FLAG_GENERIC_KERNELS):
        r = A.BlockIterativeReconstructionOperator(volume=vg, base_geometry=base, tilt_angles=tilts,
                                                   projections=dev(projs), iterations=2, flags=flags, mask=mask_t,
                                                   **case)
        assert r.uses_fused_path() == (flags == 0)
        r.run()
        r.synchronize()
        vols.append(r.get_reconstructed_volume(dtype=torch.float32).cpu().numpy())
    fast, generic = vols
    bound = 1e-5 if case.get("voxel_type", "float") == "float" else 2.0 ** -9
    assert rel_l2(fast, generic) <= bound
    assert np.all(fast[mask == 0] == 0.0)       # weight 0: never updated from the initial 0
    if case.get("voxel_type", "float") == "float":
        B = {"sart": 1, "sirt": 0xFFFFFFFF}.get(case["algorithm"], case.get("block_size", 1))
        spec = O.ReconSpec(ov(vg), og(base), tilts, projs, block_size=B, order=case["projection_iteration"],
                           lam=case["lam"], use_loc=case.get("use_long_object_compensation", False), mask=mask)
        ref = spec.run(np.zeros((nz, ny, nx), np.float32), 2)
        assert rel_l2(fast, ref) <= 1e-5


def test_full_weight_mask_equals_no_mask_on_fused_path(synthetic):
    # weight 1 everywhere: fma(1, x, y) rounds like y + x, so the masked fused
    # kernels give the unmasked fused kernels' volume bit for bit
    _, tilts, projs = synthetic
    kw = dict(volume=E.VolumeGeometry.make((64, 64, 64)), base_geometry=TEST_BASE, tilt_angles=tilts[::6],
              projections=dev(projs[::6]), algorithm="blockiterative", block_size=3, lam=0.5,
              projection_iteration="random", use_long_object_compensation=True, iterations=2)
    full = torch.full((64, 64, 64), 255, dtype=torch.uint8, device="cuda")
    a = A.BlockIterativeReconstructionOperator(mask=full, **kw)
    b = A.BlockIterativeReconstructionOperator(**kw)
    assert a.uses_fused_path() and b.uses_fused_path()
    a.run()
    b.run()
    assert torch.equal(a.get_reconstructed_volume(), b.get_reconstructed_volume())
    assert a.residual_norm() == b.residual_norm()


@pytest.mark.parametrize("case", [
    # the reference's synthetic_3_parallel/sart.xml without its slab split: FP samples 4, BP samples 9, lambda 0.1
    dict(shape=(64, 64, 64), geometry="test", algorithm="sart", lam=0.1, fp_samples=4, bp_samples=9,
         projection_iteration="random"),
    # BP oversampling over 4 slabs (one sweep regroups nothing in the BP)
    dict(shape=(64, 64, 64), geometry="test", subvolumes=4, algorithm="sart", lam=0.3, bp_samples=4,
         projection_iteration="identity"),
    # the masked_reconstruction files: FP and BP samples 4 with a mask
    dict(shape=(64, 64, 64), geometry="test", algorithm="sart", lam=0.3, fp_samples=4, bp_samples=4, mask=True,
         projection_iteration="identity"),
    dict(shape=(66, 48, 40), geometry="mrc", algorithm="blockiterative", block_size=3, lam=0.5, fp_samples=2,
         bp_samples=4, use_long_object_compensation=True, projection_iteration="maxangle"),
    dict(shape=(64, 64, 64), geometry="test", algorithm="sirt", lam=1.0, fp_samples=3, bp_samples=1,
         projection_iteration="identity"),
    dict(shape=(64, 64, 64), geometry="test", algorithm="sart", lam=0.3, fp_samples=1, bp_samples=16,
         projection_iteration="identity", voxel_type="bfloat16"),
])
def test_oversampling_runs_fused_with_fast_arithmetic(synthetic, case):
    """FP sub-rays (fp_n x fp_n per pixel, averaged) and the BP sample pattern on
    the fused kernels (default arithmetic).  Within 1e-5 of the oversampled
    generic kernels and the oversampled oracle (2^-9 for bf16 storage, as for
    the mask)."""
    case = dict(case)
    shape, geometry = case.pop("shape"), case.pop("geometry")
    subvolumes = case.pop("subvolumes", 1)
    use_mask = case.pop("mask", False)
    nx, ny, nz = shape
    vg = E.VolumeGeometry.make(shape, subvolumes=subvolumes)
    if geometry == "test":
        _, tilts, projs = synthetic
        tilts, projs = tilts[::6], np.ascontiguousarray(projs[::6])
        base = TEST_BASE
    else:
        base = E.ScannerGeometry.mrc_base(nx, ny)
        tilts = S.tilt_series(60.0, 6.0)
        truth_t = dev(S.ellipsoid_phantom((nz, ny, nx), seed=79))
        pt = torch.zeros((len(tilts), ny, nx), dtype=torch.float32, device="cuda")
        scratch = torch.zeros((ny, nx), dtype=torch.float32, device="cuda")
        for i, t in enumerate(tilts):
            scratch.zero_()
            A.forward_project(E.VolumeGeometry.make(shape), truth_t, A.rotate(base, float(t)), pt[i], scratch)
        projs = pt.cpu().numpy()
    mask = seeded_mask(shape, 6) if use_mask else None
    vols = []
    for flags in (0, E.ETN_RECON_FLAG_GENERIC_KERNELS):
        r = A.BlockIterativeReconstructionOperator(volume=vg, base_geometry=base, tilt_angles=tilts,
                                                   projections=dev(projs), iterations=2, flags=flags,
                                                   mask=None if mask is None else dev(mask), **case)
        assert r.uses_fused_path() == (flags == 0)
        r.run()
        r.synchronize()
        vols.append(r.get_reconstructed_volume(dtype=torch.float32).cpu().numpy())
    fast, generic = vols
    f32 = case.get("voxel_type", "float") == "float"
    assert rel_l2(fast, generic) <= (1e-5 if f32 else 2.0 ** -9)
    if f32:
        B = {"sart": 1, "sirt": 0xFFFFFFFF}.get(case["algorithm"], case.get("block_size", 1))
        spec = O.ReconSpec(ov(vg), og(base), tilts, projs, block_size=B, order=case["projection_iteration"],
                           lam=case["lam"], use_loc=case.get("use_long_object_compensation", False),
                           fp_samples=case.get("fp_samples", 1), bp_samples=case.get("bp_samples", 1), mask=mask)
        ref = spec.run(np.zeros((nz, ny, nx), np.float32), 2)
        assert rel_l2(fast, ref) <= 1e-5


def test_oversampled_members_share_their_own_engine():
    """A batch mixing oversampled and single-sample members: each volume equals
    the member run alone on the fused path, bit for bit."""
    base = TEST_BASE
    s = load_golden("synthetic_3_parallel")
    tilts, projs = s["angles"][::9], dev(np.ascontiguousarray(s["projections"][::9]))
    common = dict(volume=E.VolumeGeometry.make((64, 64, 64)), base_geometry=base, tilt_angles=tilts,
                  projections=projs, algorithm="sart", lam=0.2, projection_iteration="identity")
    members = [dict(common, fp_samples=2, bp_samples=4), dict(common), dict(common, fp_samples=4, bp_samples=9)]
    b = A.BatchReconstruction(members)
    assert all(b.uses_fused_path(i) for i in range(3))
    b.run_iterations(2)
    b.synchronize()
    for i, m in enumerate(members):
        r = A.BlockIterativeReconstructionOperator(**m)
        r.run_iterations(2)
        r.synchronize()
        np.testing.assert_array_equal(b.get_reconstructed_volume(i, dtype=torch.float32).cpu().numpy(),
                                      r.get_reconstructed_volume(dtype=torch.float32).cpu().numpy())
    o = A.BlockIterativeReconstructionOperator(**members[0])
    with pytest.raises(E.InvalidArgument):
        o.set_workgroup_threads(256)     # the oversampled kernels run 512-thread workgroups


@pytest.mark.parametrize("kw", [
    dict(),
    dict(mask=True),
    dict(fp_samples=2, bp_samples=4),
    dict(skip_projections=(1, 5)),
    dict(algorithm="blockiterative", block_size=4, use_long_object_compensation=True),
])
def test_residual_norm_on_fused_kernels(synthetic, kw):
    """||Ax - b|| with the fused kernels' forward-only passes: bit-identical to
    the generic kernels' norm with exact arithmetic, within 1e-6 with the
    default arithmetic (masks and oversampling included)."""
    kw = dict(kw)
    _, tilts, projs = synthetic
    mask = dev(seeded_mask((64, 64, 64), 7)) if kw.pop("mask", False) else None
    args = dict(volume=E.VolumeGeometry.make((64, 64, 64)), base_geometry=TEST_BASE, tilt_angles=tilts[::5],
                projections=dev(projs[::5]), lam=0.3, projection_iteration="identity", mask=mask)
    args.setdefault("algorithm", "sart")
    args.update(kw)
    norms = {}
    for name, flags in (("fast", 0), ("exact", E.ETN_RECON_FLAG_EXACT_ARITHMETIC),
                        ("generic", E.ETN_RECON_FLAG_GENERIC_KERNELS)):
        r = A.BlockIterativeReconstructionOperator(flags=flags, **args)
        r.run_iterations(1)
        # the same volume for every path: the generic one's after one iteration
        norms[name] = (r, r.uses_fused_path())
    ref_op = norms["generic"][0]
    vol = ref_op.get_reconstructed_volume()
    n_generic = ref_op.residual_norm()
    for name in ("fast", "exact"):
        r, fused = norms[name]
        r.set_volume(vol.clone())
        if name == "fast":
            assert fused
            assert abs(r.residual_norm() - n_generic) <= 1e-6 * n_generic
        elif fused:
            assert r.residual_norm() == n_generic


@pytest.mark.parametrize("case", [
    dict(shape=(64, 64, 64), geometry="test", algorithm="blockiterative", block_size=5, lam=0.8,
         projection_iteration="maxangle", use_long_object_compensation=True),
    dict(shape=(66, 48, 40), geometry="mrc", algorithm="sart", lam=0.3, projection_iteration="random"),
    dict(shape=(64, 64, 64), geometry="test", algorithm="sart", lam=0.3, projection_iteration="identity",
         voxel_type="half"),
])
def test_mask_exact_fused_equals_generic(synthetic, case):
    """With exact arithmetic the masked fused kernels perform the generic
    kernels' operations (dt * weight summed into the ray length, weight * sum
    added to the voxel), so volumes and residual norms are bit-identical."""
    case = dict(case)
    shape, geometry = case.pop("shape"), case.pop("geometry")
    nx, ny, nz = shape
    vg = E.VolumeGeometry.make(shape)
    if geometry == "test":
        _, tilts, projs = synthetic
        tilts, projs = tilts[::5], np.ascontiguousarray(projs[::5])
        base = TEST_BASE
    else:
        base = E.ScannerGeometry.mrc_base(nx, ny)
        tilts = S.tilt_series(60.0, 5.0)
        truth_t = dev(S.ellipsoid_phantom((nz, ny, nx), seed=80))
        pt = torch.zeros((len(tilts), ny, nx), dtype=torch.float32, device="cuda")
        scratch = torch.zeros((ny, nx), dtype=torch.float32, device="cuda")
        for i, t in enumerate(tilts):
            scratch.zero_()
            A.forward_project(vg, truth_t, A.rotate(base, float(t)), pt[i], scratch)
        projs = pt.cpu().numpy()
    mask_t = dev(seeded_mask(shape, 8))
    out = []
    for flags in (E.ETN_RECON_FLAG_EXACT_ARITHMETIC, E.ETN_RECON_FLAG_GENERIC_KERNELS):
        r = A.BlockIterativeReconstructionOperator(volume=vg, base_geometry=base, tilt_angles=tilts,
                                                   projections=dev(projs), iterations=2, flags=flags, mask=mask_t,
                                                   **case)
        assert r.uses_fused_path() == (flags == E.ETN_RECON_FLAG_EXACT_ARITHMETIC)
        r.run()
        r.synchronize()
        out.append((r.get_reconstructed_volume(dtype=torch.float32).cpu().numpy(), r.residual_norm()))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


@pytest.mark.parametrize("case", [
    dict(fp_samples=2, bp_samples=4, algorithm="sart", lam=0.3, projection_iteration="random"),
    dict(fp_samples=4, bp_samples=9, algorithm="blockiterative", block_size=4, lam=0.5,
         projection_iteration="maxangle", use_long_object_compensation=True),
    dict(fp_samples=3, bp_samples=1, algorithm="sirt", lam=1.0, projection_iteration="identity", mask=True),
    dict(fp_samples=1, bp_samples=16, algorithm="sart", lam=0.3, projection_iteration="identity",
         voxel_type="half"),
])
def test_oversampling_exact_fused_equals_generic(synthetic, case):
    """Exact arithmetic with oversampling: the fused kernels trace each sub-ray
    column once and add its sums in the generic (si, sj) order, and run the
    generic sample loop of the BP per voxel, so volumes and residual norms are
    bit-identical to the generic kernels'."""
    case = dict(case)
    mask = dev(seeded_mask((64, 64, 64), 9)) if case.pop("mask", False) else None
    _, tilts, projs = synthetic
    out = []
    for flags in (E.ETN_RECON_FLAG_EXACT_ARITHMETIC, E.ETN_RECON_FLAG_GENERIC_KERNELS):
        r = A.BlockIterativeReconstructionOperator(volume=E.VolumeGeometry.make((64, 64, 64)), base_geometry=TEST_BASE,
                                                   tilt_angles=tilts[::6], projections=dev(projs[::6]),
                                                   iterations=2, flags=flags, mask=mask, **case)
        assert r.uses_fused_path() == (flags == E.ETN_RECON_FLAG_EXACT_ARITHMETIC)
        r.run()
        r.synchronize()
        out.append((r.get_reconstructed_volume(dtype=torch.float32).cpu().numpy(), r.residual_norm()))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]
