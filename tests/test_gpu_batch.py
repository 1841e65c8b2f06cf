"""Batched reconstructions and the fused kernels' tiles.

A batch advances independent reconstructions of different sizes together,
each fused launch carrying one pass of every member; with exact arithmetic a
member's volume must be bit-identical to the same reconstruction run alone
with the generic kernels, and with the default arithmetic bit-identical to it
run alone on the fused path.  The tile (planes per workgroup, z rows per LDS
band) only changes how the work is cut, so every tile, and whatever the
autotuner picks, must give bit-identical volumes too."""
import numpy as np
import pytest

import ettention_amd as E
from ettention_amd import api as A
from ettention_amd import synthetic as S

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def member(shape, ntilts, seed, **kw):
    """SART-style inputs on the MRC geometry: a seeded phantom projected by the
    generic forward projector at ntilts angles over +-60 degrees."""
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
    d = dict(volume=vg, base_geometry=base, tilt_angles=tilts, projections=projs, algorithm="sart", lam=0.3,
             projection_iteration="random")
    d.update(kw)
    return d


def alone(m, iterations, flags=E.ETN_RECON_FLAG_GENERIC_KERNELS):
    r = A.BlockIterativeReconstructionOperator(flags=flags | m.get("flags", 0),
                                               **{k: v for k, v in m.items() if k != "flags"})
    r.run_iterations(iterations)
    r.synchronize()
    return r, r.get_reconstructed_volume(dtype=torch.float32).cpu().numpy()


@pytest.fixture(scope="module")
def members():
    return [
        member((64, 64, 64), 41, 4000),
        member((48, 48, 40), 61, 4001, algorithm="blockiterative", block_size=4, projection_iteration="maxangle"),
        member((32, 32, 32), 25, 4002, use_long_object_compensation=True),
        member((96, 64, 32), 33, 4003),
        member((30, 30, 30), 21, 4004),                          # nx % 4 != 0: its own engine
        member((40, 40, 24), 19, 4005, voxel_type="bfloat16"),  # another storage type: its own engine
        member((32, 32, 32), 17, 4006, fp_samples=2),            # oversampled: its own engine
        member((64, 64, 64), 29, 4007, flags=E.ETN_RECON_FLAG_GENERIC_KERNELS),
    ]


EXACT = E.ETN_RECON_FLAG_EXACT_ARITHMETIC


def test_batch_members_equal_standalone_generic(members):
    b = A.BatchReconstruction(members, flags=EXACT)
    fused = [b.uses_fused_path(i) for i in range(len(b))]
    assert fused == [True, True, True, True, True, True, True, False]
    b.run_iterations(2)
    b.synchronize()
    for i, m in enumerate(members):
        r, ref = alone(m, 2)
        got = b.get_reconstructed_volume(i, dtype=torch.float32).cpu().numpy()
        np.testing.assert_array_equal(got, ref, err_msg="member %d" % i)
        assert b.residual_norm(i) == r.residual_norm()
        if fused[i]:
            planes, band = b.tile(i)
            assert planes in (1, 2, 4, 8) and m["volume"].resolution[1] % planes == 0 and band >= 1
    ms, n = b.time_iteration()
    assert ms > 0.0 and n >= max(len(m["tilt_angles"]) for m in members[:6])


def test_batch_members_equal_standalone_fused(members):
    # default arithmetic: each member as it runs alone on the fused path
    sel = members[:4]
    b = A.BatchReconstruction(sel)
    b.run_iterations(2)
    for i, m in enumerate(sel):
        _, ref = alone(m, 2, flags=0)
        np.testing.assert_array_equal(b.get_reconstructed_volume(i, dtype=torch.float32).cpu().numpy(), ref)


def test_batch_graph_replays_equal_eager(members):
    sel = members[:4]
    outs = []
    for flags in (0, E.ETN_RECON_FLAG_NO_GRAPH | E.ETN_RECON_FLAG_NO_AUTOTUNE):
        b = A.BatchReconstruction(sel, flags=flags)
        b.run_iterations(1)
        b.run_iterations(2)
        b.synchronize()
        outs.append([b.get_reconstructed_volume(i, dtype=torch.float32).cpu().numpy() for i in range(len(sel))])
    for a, c in zip(*outs):
        np.testing.assert_array_equal(a, c)


def test_batch_with_a_member_in_rounds():
    # a 512 x 1024 x 8 member has 1024 one-plane groups, more than the device
    # holds at once: its groups are dispatched interleaved, the small member's
    # in order; with exact arithmetic each equals its generic run alone
    sel = [member((512, 1024, 8), 13, 4010), member((64, 48, 40), 21, 4011)]
    b = A.BatchReconstruction(sel, flags=EXACT)
    assert all(b.uses_fused_path(i) for i in range(len(b)))
    b.run_iterations(2)
    b.synchronize()
    for i, m in enumerate(sel):
        r, ref = alone(m, 2)
        np.testing.assert_array_equal(b.get_reconstructed_volume(i, dtype=torch.float32).cpu().numpy(), ref,
                                      err_msg="member %d" % i)


def test_batch_all_generic_flag(members):
    sel = members[2:4]
    b = A.BatchReconstruction(sel, flags=E.ETN_RECON_FLAG_GENERIC_KERNELS)
    assert not any(b.uses_fused_path(i) for i in range(len(b)))
    b.run_iterations(1)
    for i, m in enumerate(sel):
        _, ref = alone(m, 1)
        np.testing.assert_array_equal(b.get_reconstructed_volume(i, dtype=torch.float32).cpu().numpy(), ref)


@pytest.mark.parametrize("tile", [(1, 1), (1, 7), (1, 21), (2, 5), (4, 3), (8, 4), (8, 1)])
def test_every_tile_gives_the_same_volume(members, tile):
    m = members[0]   # 64^3, 41 tilts
    _, ref = alone(m, 2)
    r = A.BlockIterativeReconstructionOperator(flags=EXACT, **m)
    r.set_tile(*tile)
    assert r.tile() == tile
    r.run_iterations(2)
    r.synchronize()
    np.testing.assert_array_equal(r.get_reconstructed_volume(dtype=torch.float32).cpu().numpy(), ref)


def test_invalid_tiles_are_rejected(members):
    r = A.BlockIterativeReconstructionOperator(**members[3])   # 96 x 64 x 32
    for planes, band in [(3, 4), (16, 1), (1, 0), (1, 1000), (8, 100)]:
        with pytest.raises(E.InvalidArgument):
            r.set_tile(planes, band)
    g = A.BlockIterativeReconstructionOperator(flags=E.ETN_RECON_FLAG_GENERIC_KERNELS, **members[3])
    with pytest.raises(E.Unsupported):
        g.set_tile(1, 1)


def test_autotuned_equals_default_tile(members):
    m = members[1]
    outs = []
    for flags in (EXACT, EXACT | E.ETN_RECON_FLAG_NO_AUTOTUNE):
        r = A.BlockIterativeReconstructionOperator(flags=flags, **m)
        r.run_single_iteration()     # the autotuner runs first and restores the volume
        r.run_single_iteration()
        r.synchronize()
        outs.append(r.get_reconstructed_volume(dtype=torch.float32).cpu().numpy())
    np.testing.assert_array_equal(outs[0], outs[1])
    _, ref = alone(m, 2)
    np.testing.assert_array_equal(outs[0], ref)


def test_batch_with_masks_equals_each_member_alone(members):
    """Masked members share their own fused engine (the mask kernels); members
    without a mask keep theirs.  Each volume equals the member run alone on the
    fused path, and every tile of a masked member gives the same volume."""
    def mask_for(m, seed):
        nx, ny, nz = m["volume"].resolution
        rng = np.random.default_rng(seed)
        return torch.from_numpy(rng.integers(0, 256, (nz, ny, nx), dtype=np.uint8)).cuda()

    sel = [dict(members[0], mask=mask_for(members[0], 1)), members[1],
           dict(members[3], mask=mask_for(members[3], 2)), members[2]]
    b = A.BatchReconstruction(sel)
    assert all(b.uses_fused_path(i) for i in range(len(sel)))
    b.run_iterations(2)
    b.synchronize()
    for i, m in enumerate(sel):
        _, ref = alone(m, 2, flags=0)
        np.testing.assert_array_equal(b.get_reconstructed_volume(i, dtype=torch.float32).cpu().numpy(), ref,
                                      err_msg="member %d" % i)
    _, ref = alone(sel[0], 2, flags=0)
    for tile in [(1, 1), (2, 5), (8, 1)]:
        r = A.BlockIterativeReconstructionOperator(**sel[0])
        r.set_tile(*tile)
        r.run_iterations(2)
        r.synchronize()
        np.testing.assert_array_equal(r.get_reconstructed_volume(dtype=torch.float32).cpu().numpy(), ref)


@pytest.mark.parametrize("tile", [(1, 1), (1, 6), (2, 5), (4, 2), (8, 1)])
def test_256_thread_workgroups_give_the_same_volume(members, tile):
    # the 256-thread shape (four workgroups per CU, 40 KiB LDS) with exact
    # arithmetic: bit-identical to the generic kernels for every tile
    m = members[0]   # 64^3, 41 tilts
    _, ref = alone(m, 2)
    r = A.BlockIterativeReconstructionOperator(flags=EXACT, **m)
    r.set_workgroup_threads(256)
    r.set_tile(*tile)
    assert r.workgroup_threads() == 256 and r.tile() == tile
    r.run_iterations(2)
    r.synchronize()
    np.testing.assert_array_equal(r.get_reconstructed_volume(dtype=torch.float32).cpu().numpy(), ref)


@pytest.mark.parametrize("which", [1, 2, 4, 5])
def test_workgroup_shapes_agree_with_fast_arithmetic(members, which):
    # default arithmetic, block / LOC / odd nx / bf16 members, with and without
    # a mask: 256 and 512 threads per workgroup give the same volume bit for bit
    m = members[which]
    nx, ny, nz = m["volume"].resolution
    rng = np.random.default_rng(which)
    for mask in (None, torch.from_numpy(rng.integers(0, 256, (nz, ny, nx), dtype=np.uint8)).cuda()):
        vols = []
        for threads in (512, 256):
            r = A.BlockIterativeReconstructionOperator(mask=mask, **m)
            r.set_workgroup_threads(threads)
            r.run_iterations(2)
            r.synchronize()
            vols.append(r.get_reconstructed_volume(dtype=torch.float32).cpu().numpy())
        np.testing.assert_array_equal(vols[0], vols[1])


def test_workgroup_threads_are_validated(members):
    r = A.BlockIterativeReconstructionOperator(**members[3])
    for threads in (0, 128, 384, 1024):
        with pytest.raises(E.InvalidArgument):
            r.set_workgroup_threads(threads)
    r.set_workgroup_threads(256)
    assert r.workgroup_threads() == 256
    with pytest.raises(E.InvalidArgument):
        r.set_tile(8, 100)   # beyond the 40 KiB array
    g = A.BlockIterativeReconstructionOperator(flags=E.ETN_RECON_FLAG_GENERIC_KERNELS, **members[3])
    with pytest.raises(E.Unsupported):
        g.set_workgroup_threads(256)
    b = A.BatchReconstruction(members[:2])
    b.run_iterations(1)
    assert b.workgroup_threads(0) in (256, 512) and b.workgroup_threads(1) in (256, 512)
