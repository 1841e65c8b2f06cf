"""The fused kernels' register back projection against their LDS back
projection (ETN_RECON_FLAG_LDS_BACK_PROJECTION).

With fast arithmetic, nx a multiple of 4 and no oversampling, the fused
kernels apply a pass's back projections to each group of 4 voxels in
registers as the band arrives from HBM, all of them in block order, and stage
the detector records of every back projection of the pass once per launch
(in a 160 KiB LDS array when they do not fit the workgroup shape's own).  The
LDS path applies one back projection after another to the band in LDS and
restages the records per projection and band.  Every voxel sees the same
chain of updates and roundings on both, so the volumes must agree bit for
bit: single projections (SART), blocks of 4 and 8 (up to 8 back projections
per pass), long-object compensation, masks, 16-bit storage, both workgroup
shapes, several tiles, and the weighted back projection's many-projection
passes."""
import numpy as np
import pytest

import ettention_amd as E
from ettention_amd import api as A
from ettention_amd import synthetic as S

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

LDS = E.ETN_RECON_FLAG_LDS_BACK_PROJECTION


def inputs(shape, ntilts, seed, **kw):
    nx, ny, nz = shape
    vg = E.VolumeGeometry.make((nx, ny, nz))
    base = E.ScannerGeometry.mrc_base(nx, ny)
    tilts = np.linspace(-60.0, 60.0, ntilts).astype(np.float32)
    ph = S.ellipsoid_phantom((nz, ny, nx), seed=seed, device=torch.device("cuda"))
    projs = torch.zeros((ntilts, ny, nx), dtype=torch.float32, device="cuda")
    scratch = torch.zeros((ny, nx), dtype=torch.float32, device="cuda")
    for i, t in enumerate(tilts):
        scratch.zero_()
        A.forward_project(vg, ph, A.rotate(base, float(t)), projs[i], scratch)
    torch.cuda.synchronize()
    d = dict(volume=vg, base_geometry=base, tilt_angles=tilts, projections=S.add_gaussian_noise(projs, 0.02, seed=seed),
             algorithm="sart", lam=0.3, projection_iteration="random")
    d.update(kw)
    return d


def run(case, flags=0, threads=None, tile=None, iterations=2):
    r = A.BlockIterativeReconstructionOperator(flags=flags, **case)
    assert r.uses_fused_path()
    if threads is not None:
        r.set_workgroup_threads(threads)
    if tile is not None:
        r.set_tile(*tile)
    r.run_iterations(iterations)
    r.synchronize()
    return r.get_reconstructed_volume(dtype=torch.float32).cpu().numpy(), r.residual_norm()


CASES = {
    "sart": ((64, 64, 64), 41, {}),
    "block4": ((48, 48, 40), 61, dict(algorithm="blockiterative", block_size=4, projection_iteration="maxangle")),
    "block8": ((128, 64, 64), 61, dict(algorithm="blockiterative", block_size=8)),
    "sirt_loc": ((64, 32, 48), 21, dict(algorithm="sirt", use_long_object_compensation=True)),
    "block8_bf16": ((64, 64, 32), 41, dict(algorithm="blockiterative", block_size=8, voxel_type="bfloat16")),
    "sart_half": ((64, 48, 40), 31, dict(voxel_type="half")),
    # 512 wide: 8 record sets of 512-voxel rows no longer fit the workgroup
    # shape's LDS, so one-plane tiles take the 80 KiB register path (bands of
    # at most 7 rows) and wider tiles the 160 KiB one
    "block8_wide": ((512, 8, 24), 61, dict(algorithm="blockiterative", block_size=8)),
    "block8_wide_bf16": ((512, 8, 24), 41, dict(algorithm="blockiterative", block_size=8, voxel_type="bfloat16")),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("threads", [512, 256])
def test_register_back_projection_equals_lds_back_projection(name, threads):
    shape, ntilts, kw = CASES[name]
    case = inputs(shape, ntilts, 7100 + len(name), **kw)
    print("register bp case", name, threads, flush=True)
    reg, reg_norm = run(case, threads=threads)
    lds, lds_norm = run(case, flags=LDS, threads=threads)
    np.testing.assert_array_equal(reg, lds)
    assert reg_norm == lds_norm


@pytest.mark.parametrize("name", ["sart", "block8", "block8_wide"])
def test_register_back_projection_with_a_mask(name):
    shape, ntilts, kw = CASES[name]
    nx, ny, nz = shape
    rng = np.random.default_rng(71)
    mask = torch.from_numpy(rng.integers(0, 256, (nz, ny, nx), dtype=np.uint8)).cuda()
    case = inputs(shape, ntilts, 7200, mask=mask, **kw)
    for threads in (512, 256):
        print("register bp mask case", name, threads, flush=True)
        reg, _ = run(case, threads=threads)
        lds, _ = run(case, flags=LDS, threads=threads)
        np.testing.assert_array_equal(reg, lds)


@pytest.mark.parametrize("tile", [(1, 1), (1, 5), (2, 3), (4, 2), (8, 1)])
def test_register_back_projection_every_tile(tile):
    # planes x band rows; every tile cuts the same work, on both paths
    shape, ntilts, kw = CASES["block4"]
    case = inputs(shape, ntilts, 7300, **kw)
    print("register bp tile", tile, flush=True)
    ref, _ = run(case, flags=LDS | E.ETN_RECON_FLAG_NO_AUTOTUNE)
    got, _ = run(case, flags=E.ETN_RECON_FLAG_NO_AUTOTUNE, tile=tile)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("tile", [(1, 1), (1, 3), (1, 12), (2, 2), (4, 1)])
def test_register_back_projection_wide_tiles(tile):
    # 512 wide, 8 back projections per pass: one plane per tile takes the
    # 80 KiB register path with the band cut to what fits beside the records,
    # two planes the 160 KiB one, four (whose records exceed even that) the
    # LDS path
    shape, ntilts, kw = CASES["block8_wide"]
    case = inputs(shape, ntilts, 7500, **kw)
    print("register bp wide tile", tile, flush=True)
    ref, _ = run(case, flags=LDS | E.ETN_RECON_FLAG_NO_AUTOTUNE, threads=512)
    got, _ = run(case, flags=E.ETN_RECON_FLAG_NO_AUTOTUNE, threads=512, tile=tile)
    np.testing.assert_array_equal(got, ref)


def test_wbp_register_back_projection_equals_lds():
    # the weighted back projection's passes carry 8 projections each
    n = 64
    vg = E.VolumeGeometry.make((n, n, n))
    base = E.ScannerGeometry.mrc_base(n, n)
    tilts = np.linspace(-60.0, 60.0, 37).astype(np.float32)
    ph = S.ellipsoid_phantom((n, n, n), seed=7400, device=torch.device("cuda"))
    projs = torch.zeros((len(tilts), n, n), dtype=torch.float32, device="cuda")
    scratch = torch.zeros((n, n), dtype=torch.float32, device="cuda")
    for i, t in enumerate(tilts):
        scratch.zero_()
        A.forward_project(vg, ph, A.rotate(base, float(t)), projs[i], scratch)
    vols = []
    for flags in (0, LDS):
        for kw in (dict(), dict(bp_samples=9)):   # 9 samples: oversampled, the LDS path either way
            r = A.WeightedBackProjection(volume=vg, base_geometry=base, tilt_angles=tilts, projections=projs,
                                         lam=1.0, flags=flags, **kw)
            assert r.uses_fused_path()
            r.run()
            vols.append(r.get_reconstructed_volume().cpu().numpy())
    np.testing.assert_array_equal(vols[0], vols[2])
    np.testing.assert_array_equal(vols[1], vols[3])
