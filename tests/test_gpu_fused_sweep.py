"""A seeded sweep over combinations of the fused kernels' options: volume
shape (nx a multiple of 4 or not, odd sizes), detector pitch, algorithm and
block size, projection order, long-object compensation, z-slabs, mask,
forward and back-projection oversampling, storage type, workgroup shape and
tile; and wide planes of 512 to 2048 columns.  With the default arithmetic and f32 storage every combination must
stay within the FP32 parity bound (relative L2 1e-5) of the same
reconstruction on the generic kernels, whichever path the options allow.
With bf16 storage the two paths round differently and their volumes drift
apart by several bf16 ulps (up to 5e-3 with lambda = 1), each about 8e-3 from
the f32 volume; there the fused volume must be as close to the f32 generic
volume as the bf16 generic volume is (within 10% and 1e-4)."""
import numpy as np
import pytest

import ettention_amd as E
from ettention_amd import api as A
from ettention_amd import synthetic as S
from tests._util import rel_l2

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def config(seed):
    rng = np.random.default_rng(seed)
    shape = tuple(int(v) for v in (rng.choice([32, 40, 44, 48, 52, 64]), rng.choice([16, 24, 32, 33]),
                                    rng.choice([17, 24, 32, 40])))
    if rng.random() < 0.3:
        shape = (shape[0] + 2, shape[1], shape[2])        # nx not a multiple of 4: scalar band path
    alg = rng.choice(["sart", "sirt", "blockiterative"])
    c = dict(shape=shape, algorithm=str(alg), lam=float(rng.choice([0.2, 0.5, 1.0])),
             projection_iteration=str(rng.choice(["identity", "random", "maxangle", "buckets"])),
             tilt_step=float(rng.choice([4.0, 6.0, 10.0])), pitch=float(rng.choice([1.0, 0.5])))
    if alg == "blockiterative":
        c["block_size"] = int(rng.integers(2, 6))
    if rng.random() < 0.3:
        c["use_long_object_compensation"] = True
    if rng.random() < 0.25:
        c["subvolumes"] = int(rng.choice([2, 3]))
    if rng.random() < 0.3:
        c["mask"] = True
    if rng.random() < 0.35:
        c["fp_samples"] = int(rng.choice([2, 3]))
    if rng.random() < 0.35:
        c["bp_samples"] = int(rng.choice([4, 9]))
    if rng.random() < 0.2:
        c["voxel_type"] = "bfloat16"
    c["threads"] = int(rng.choice([512, 256]))
    c["seed"] = seed
    return c


def wide_config(seed):
    """Wide planes (512 to 2048 columns): four or more rays per thread, so the
    launches on the LDS path take the 160 KiB array, and 2048 columns put
    nx + nz past 2049 with rays that cannot reach the step limit."""
    rng = np.random.default_rng(seed)
    nx = int(rng.choice([512, 1024, 1536, 2048]))
    c = dict(shape=(nx, int(rng.choice([2, 4, 6])), int(rng.choice([24, 48, 96]))),
             algorithm=str(rng.choice(["sart", "sirt", "blockiterative"])), lam=float(rng.choice([0.2, 0.5])),
             projection_iteration=str(rng.choice(["identity", "random", "maxangle"])),
             tilt_step=float(rng.choice([6.0, 10.0])), pitch=1.0 if nx > 1024 else float(rng.choice([1.0, 0.5])))
    if c["algorithm"] == "blockiterative":
        c["block_size"] = int(rng.integers(2, 9))
    if rng.random() < 0.3:
        c["mask"] = True
    if rng.random() < 0.2:
        c["use_long_object_compensation"] = True
    if rng.random() < 0.2:
        c["voxel_type"] = "bfloat16"
    c["threads"] = 512
    c["seed"] = seed
    return c


@pytest.mark.parametrize("seed", list(range(100, 140)) + ["wide%d" % k for k in range(12)])
def test_fused_option_sweep_matches_generic(seed):
    c = wide_config(200 + int(seed[4:])) if isinstance(seed, str) else config(seed)
    seed = c["seed"]
    print("sweep case", c, flush=True)
    nx, ny, nz = c["shape"]
    vg = E.VolumeGeometry.make((nx, ny, nz), subvolumes=c.get("subvolumes", 1))
    pitch = c["pitch"]
    nu, nv = int(round(nx / pitch)), int(round(ny / pitch))
    base = E.ScannerGeometry.make((-nx / 2, -ny / 2, -2.0 * nz), (-nx / 2, -ny / 2, 2.0 * nz), (pitch, 0, 0),
                                  (0, pitch, 0), (nu, nv))
    tilts = S.tilt_series(60.0, c["tilt_step"])
    truth = torch.from_numpy(S.ellipsoid_phantom((nz, ny, nx), seed=seed)).cuda()
    projs = torch.zeros((len(tilts), nv, nu), dtype=torch.float32, device="cuda")
    scratch = torch.zeros((nv, nu), dtype=torch.float32, device="cuda")
    whole = E.VolumeGeometry.make((nx, ny, nz))
    for i, t in enumerate(tilts):
        scratch.zero_()
        A.forward_project(whole, truth, A.rotate(base, float(t)), projs[i], scratch)
    mask = None
    if c.get("mask"):
        rng = np.random.default_rng(seed + 1)
        mask = torch.from_numpy(rng.integers(0, 256, (nz, ny, nx), dtype=np.uint8)).cuda()
    kw = dict(volume=vg, base_geometry=base, tilt_angles=tilts, projections=projs, algorithm=c["algorithm"],
              lam=c["lam"], projection_iteration=c["projection_iteration"], iterations=2, mask=mask,
              block_size=c.get("block_size", 0),
              use_long_object_compensation=c.get("use_long_object_compensation", False),
              fp_samples=c.get("fp_samples", 1), bp_samples=c.get("bp_samples", 1),
              voxel_type=c.get("voxel_type", "float"))
    fast = A.BlockIterativeReconstructionOperator(**kw)
    if fast.uses_fused_path():
        try:
            fast.set_workgroup_threads(c["threads"])
        except E.InvalidArgument:
            pass                       # the shape does not fit this volume (or oversampling): keep 512
    fast.run()
    fast.synchronize()
    generic = A.BlockIterativeReconstructionOperator(flags=E.ETN_RECON_FLAG_GENERIC_KERNELS, **kw)
    generic.run()
    generic.synchronize()
    a = fast.get_reconstructed_volume(dtype=torch.float32).cpu().numpy()
    b = generic.get_reconstructed_volume(dtype=torch.float32).cpu().numpy()
    if c.get("voxel_type", "float") == "float":
        assert rel_l2(a, b) <= 1e-5, (c, fast.uses_fused_path(), rel_l2(a, b))
    else:
        ref = A.BlockIterativeReconstructionOperator(flags=E.ETN_RECON_FLAG_GENERIC_KERNELS,
                                                     **dict(kw, voxel_type="float"))
        ref.run()
        r32 = ref.get_reconstructed_volume().cpu().numpy()
        e_fast, e_generic = rel_l2(a, r32), rel_l2(b, r32)
        assert e_fast <= 1.1 * e_generic + 1e-4, (c, e_fast, e_generic)
    na, nb = fast.residual_norm(), generic.residual_norm()
    assert abs(na - nb) <= (1e-5 if c.get("voxel_type", "float") == "float" else 1e-3) * nb
