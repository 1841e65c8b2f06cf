"""Edge cases of the reference's path on the GPU, against the oracle.

* The DDA step limit (DDARaytracer.hl:85-108, at most 2048 steps per ray): a
  thin 1152 x 2 x 1152 slab whose oblique rays cross more than 2048 voxel
  borders, so the fused kernels run their step-limited variant.  Generic forward projections against the oracle within 1e-6; SART
  through the generic kernels against the oracle within 1e-5; exact fused
  arithmetic bit-identical to the generic kernels; fast arithmetic within 1e-5.
* A 1536 x 2 x 600 volume at tilts up to 60 degrees: nx + nz > 2049, but no
  ray can reach the limit, so the fused kernels run without it; exact fused
  arithmetic bit-identical to the generic kernels (which keep it), fast
  within 1e-5, the generic kernels within 1e-5 of the oracle.  The same at
  1024 x 2 x 1024 with tilts up to 44 degrees, where a ray's bound (2015
  voxels) sits just below the 2040 at which the limited kernels take over.
* x-axis tilt (SingleAxisTiltRotator.cpp:62-78): rays leave the y planes, so
  the generic kernels run; SART, SIRT and weighted back projection against the
  oracle within 1e-5."""
import numpy as np
import pytest

import ettention_amd as E
from ettention_amd import api as A
from oracle import oracle as O
from oracle import wbp_oracle as W
from tests._util import load_golden, rel_l2

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

CAP_SHAPE = (1152, 2, 1152)      # nx, ny, nz
CAP_TILTS = np.array([-60, -45, -40, -20, 0, 20, 40, 45, 60], np.float32)
TEST_BASE = E.ScannerGeometry.make((-64, -64, -128), (-64, -64, 128), (0.5, 0, 0), (0, 0.5, 0), (256, 256))


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def og(g):
    return O.Scanner.make(g.source, g.detector, g.horizontal_pitch, g.vertical_pitch, g.resolution)


@pytest.fixture(scope="module")
def cap_case():
    nx, ny, nz = CAP_SHAPE
    rng = np.random.default_rng(21)
    truth = rng.uniform(0.0, 1.0, (nz, ny, nx)).astype(np.float32)
    ov = O.Volume.make(CAP_SHAPE)
    obase = O.mrc_base_scanner(nx, ny)
    projs = np.stack([O.forward_project(ov, truth, O.rotate(obase, float(t)))[0] for t in CAP_TILTS])
    return truth, projs


def test_step_limit_forward_projection_matches_oracle(cap_case):
    truth, _ = cap_case
    nx, ny, nz = CAP_SHAPE
    print("step limit FP", CAP_SHAPE, flush=True)
    vg = E.VolumeGeometry.make(CAP_SHAPE)
    base = E.ScannerGeometry.mrc_base(nx, ny)
    ov = O.Volume.make(CAP_SHAPE)
    obase = O.mrc_base_scanner(nx, ny)
    vol = dev(truth)
    for t in (40.0, -45.0, 60.0):
        p = torch.zeros((ny, nx), dtype=torch.float32, device="cuda")
        ln = torch.zeros_like(p)
        A.forward_project(vg, vol, A.rotate(base, t), p, ln)
        rp, rl = O.forward_project(ov, truth, O.rotate(obase, t))
        assert rel_l2(p.cpu().numpy(), rp) <= 1e-6, t
        assert rel_l2(ln.cpu().numpy(), rl) <= 1e-6, t


def test_step_limit_reconstruction(cap_case):
    _, projs = cap_case
    nx, ny, nz = CAP_SHAPE
    print("step limit SART", CAP_SHAPE, flush=True)
    vg = E.VolumeGeometry.make(CAP_SHAPE)
    base = E.ScannerGeometry.mrc_base(nx, ny)
    kw = dict(volume=vg, base_geometry=base, tilt_angles=CAP_TILTS, projections=dev(projs), algorithm="sart",
              lam=0.3, projection_iteration="random")
    vols = {}
    for name, flags in (("generic", E.ETN_RECON_FLAG_GENERIC_KERNELS), ("exact", E.ETN_RECON_FLAG_EXACT_ARITHMETIC),
                        ("fast", 0)):
        r = A.BlockIterativeReconstructionOperator(flags=flags, **kw)
        assert r.uses_fused_path() == (name != "generic"), name
        r.run_iterations(2)
        vols[name] = r.get_reconstructed_volume().cpu().numpy()
    spec = O.ReconSpec(O.Volume.make(CAP_SHAPE), O.mrc_base_scanner(nx, ny), CAP_TILTS, projs, block_size=1,
                       order="random", lam=0.3)
    ref = spec.run(np.zeros((nz, ny, nx), np.float32), 2)
    assert rel_l2(vols["generic"], ref) <= 1e-5
    assert np.array_equal(vols["exact"], vols["generic"])
    assert rel_l2(vols["fast"], vols["generic"]) <= 1e-5


def test_wide_volume_below_the_step_limit():
    # nx + nz > 2049, but at tilts up to 60 degrees a ray visits at most
    # floor(600 * tan(60)) + 600 + 3 = 1642 voxels: the fused kernels run
    # without the step limit and must still match the generic kernels, which
    # keep it
    shape = (1536, 2, 600)
    nx, ny, nz = shape
    tilts = np.array([-60, -35, -10, 0, 15, 30, 50, 60], np.float32)
    rng = np.random.default_rng(22)
    truth = rng.uniform(0.0, 1.0, (nz, ny, nx)).astype(np.float32)
    ov = O.Volume.make(shape)
    obase = O.mrc_base_scanner(nx, ny)
    projs = np.stack([O.forward_project(ov, truth, O.rotate(obase, float(t)))[0] for t in tilts])
    print("wide volume SART", shape, flush=True)
    kw = dict(volume=E.VolumeGeometry.make(shape), base_geometry=E.ScannerGeometry.mrc_base(nx, ny),
              tilt_angles=tilts, projections=dev(projs), algorithm="sart", lam=0.3, projection_iteration="random")
    vols = {}
    for name, flags in (("generic", E.ETN_RECON_FLAG_GENERIC_KERNELS), ("exact", E.ETN_RECON_FLAG_EXACT_ARITHMETIC),
                        ("fast", 0)):
        r = A.BlockIterativeReconstructionOperator(flags=flags, **kw)
        assert r.uses_fused_path() == (name != "generic"), name
        r.run_iterations(2)
        vols[name] = r.get_reconstructed_volume().cpu().numpy()
    spec = O.ReconSpec(ov, obase, tilts, projs, block_size=1, order="random", lam=0.3)
    ref = spec.run(np.zeros((nz, ny, nx), np.float32), 2)
    assert rel_l2(vols["generic"], ref) <= 1e-5
    assert np.array_equal(vols["exact"], vols["generic"])
    assert rel_l2(vols["fast"], vols["generic"]) <= 1e-5


def test_step_limit_bound_near_the_limit():
    # 1024 x 2 x 1024 at tilts up to 44 degrees: a ray can visit up to
    # floor(1024 tan 44) + 1024 + 3 = 2015 voxels (the actual longest about
    # 2013), just below the 2040 bound, so the fused kernels run without the
    # limit; exact fused arithmetic must still equal the generic kernels
    # (which keep it) bit for bit
    shape = (1024, 2, 1024)
    nx, ny, nz = shape
    tilts = np.array([-44, -30, -5, 0, 20, 44], np.float32)
    rng = np.random.default_rng(23)
    truth = rng.uniform(0.0, 1.0, (nz, ny, nx)).astype(np.float32)
    ov = O.Volume.make(shape)
    obase = O.mrc_base_scanner(nx, ny)
    projs = np.stack([O.forward_project(ov, truth, O.rotate(obase, float(t)))[0] for t in tilts])
    print("near-limit SART", shape, flush=True)
    kw = dict(volume=E.VolumeGeometry.make(shape), base_geometry=E.ScannerGeometry.mrc_base(nx, ny),
              tilt_angles=tilts, projections=dev(projs), algorithm="sart", lam=0.3, projection_iteration="random")
    vols = {}
    for name, flags in (("generic", E.ETN_RECON_FLAG_GENERIC_KERNELS), ("exact", E.ETN_RECON_FLAG_EXACT_ARITHMETIC),
                        ("fast", 0)):
        r = A.BlockIterativeReconstructionOperator(flags=flags, **kw)
        assert r.uses_fused_path() == (name != "generic"), name
        r.run_iterations(2)
        vols[name] = r.get_reconstructed_volume().cpu().numpy()
    assert np.array_equal(vols["exact"], vols["generic"])
    assert rel_l2(vols["fast"], vols["generic"]) <= 1e-5


@pytest.fixture(scope="module")
def synthetic():
    s = load_golden("synthetic_3_parallel")
    return s["angles"], np.ascontiguousarray(s["projections"])


@pytest.mark.parametrize("case", [
    dict(x_axis_tilt=5.0, algorithm="sart", lam=0.3, projection_iteration="random", block=1),
    dict(x_axis_tilt=-8.0, algorithm="sirt", lam=1.0, projection_iteration="identity", block=0xFFFFFFFF),
    dict(x_axis_tilt=3.0, algorithm="blockiterative", block_size=6, lam=0.5, projection_iteration="maxangle",
         use_long_object_compensation=True, block=6),
])
def test_x_axis_tilt_reconstruction_matches_oracle(synthetic, case):
    tilts, projs = synthetic
    case = dict(case)
    block = case.pop("block")
    print("x-axis tilt", case, flush=True)
    sub = slice(None, None, 4)
    t, p = tilts[sub], np.ascontiguousarray(projs[sub])
    vg = E.VolumeGeometry.make((64, 64, 64))
    r = A.BlockIterativeReconstructionOperator(volume=vg, base_geometry=TEST_BASE, tilt_angles=t,
                                               projections=dev(p), iterations=2, **case)
    assert not r.uses_fused_path()
    r.run()
    got = r.get_reconstructed_volume().cpu().numpy()
    spec = O.ReconSpec(O.Volume.make((64, 64, 64)), og(TEST_BASE), t, p, block_size=block,
                       order=case["projection_iteration"], lam=case["lam"],
                       use_loc=case.get("use_long_object_compensation", False), xtilt=case["x_axis_tilt"])
    ref = spec.run(np.zeros((64, 64, 64), np.float32), 2)
    assert rel_l2(got, ref) <= 1e-5
    # the tilt moved the result: the same run without it differs
    spec0 = O.ReconSpec(O.Volume.make((64, 64, 64)), og(TEST_BASE), t, p, block_size=block,
                        order=case["projection_iteration"], lam=case["lam"],
                        use_loc=case.get("use_long_object_compensation", False))
    assert rel_l2(spec0.run(np.zeros((64, 64, 64), np.float32), 2), ref) > 1e-2


def test_x_axis_tilt_wbp_matches_oracle(synthetic):
    tilts, projs = synthetic
    sub = slice(0, 180, 5)
    t, p = tilts[sub], np.ascontiguousarray(projs[sub])
    vg = E.VolumeGeometry.make((64, 64, 64))
    r = A.WeightedBackProjection(volume=vg, base_geometry=TEST_BASE, tilt_angles=t, projections=dev(p), lam=1.0,
                                 x_axis_tilt=6.0)
    r.run()
    got = r.get_reconstructed_volume().cpu().numpy()
    ref = W.wbp(O.Volume.make((64, 64, 64)), og(TEST_BASE), t, p, lam=1.0, xtilt=6.0)
    assert rel_l2(got, ref) < 1e-5


def test_per_projection_geometries_equal_rotated_base(synthetic):
    """etn_reconstruction_desc.geometries (what ImageStackDatasource::
    createScannerGeometry returns per projection) gives the same volume, bit for
    bit, as the base geometry rotated by the tilt angles, on the fused and the
    generic kernels and with an x-axis tilt."""
    tilts, projs = synthetic
    sub = slice(None, None, 6)
    t, p = tilts[sub], dev(projs[sub])
    vg = E.VolumeGeometry.make((64, 64, 64))
    for flags, xt in ((0, 0.0), (E.ETN_RECON_FLAG_GENERIC_KERNELS, 0.0), (0, 4.0)):
        kw = dict(volume=vg, base_geometry=TEST_BASE, tilt_angles=t, projections=p, algorithm="sart", lam=0.3,
                  projection_iteration="random", flags=flags, x_axis_tilt=xt)
        a = A.BlockIterativeReconstructionOperator(**kw)
        a.run_iterations(2)
        geoms = [A.rotate(TEST_BASE, float(x), xt) for x in t]
        kw["x_axis_tilt"] = 0.0           # carried by the geometries
        b = A.BlockIterativeReconstructionOperator(geometries=geoms, **kw)
        assert a.uses_fused_path() == b.uses_fused_path()
        b.run_iterations(2)
        assert torch.equal(a.get_reconstructed_volume(), b.get_reconstructed_volume()), (flags, xt)
