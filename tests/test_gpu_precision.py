"""16-bit volume storage with stochastic rounding (ETN_RECON_FLAG_STOCHASTIC_ROUNDING)
against round to nearest even and FP32 storage.

Stochastic rounding sends a value to the upper of its two enclosing 16-bit
neighbours with probability equal to the dropped fraction, so the stored
value equals the unrounded one in expectation.  The tests check the property
that makes it worth having: SART updates smaller than half a unit in the
last place, which round to nearest drops every time (the volume stays at its
initial value bit for bit), survive on average."""
import numpy as np
import pytest

import ettention_amd as E
from ettention_amd import api as A
from ettention_amd import synthetic as S

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def small_update_runs(voxel_type, eps, lam=0.3, init=1.0, flags=0, n=64):
    """One SART iteration from a volume of `init` towards projections of a
    volume of init (1 + eps): each update is about lam * eps * init."""
    vg = E.VolumeGeometry.make((n, n, n))
    base = E.ScannerGeometry.mrc_base(n, n)
    tilts = S.tilt_series(60.0, 4.0)
    target = torch.full((n, n, n), init * (1.0 + eps), dtype=torch.float32, device="cuda")
    projs = torch.zeros((len(tilts), n, n), dtype=torch.float32, device="cuda")
    scratch = torch.zeros((n, n), dtype=torch.float32, device="cuda")
    for i, t in enumerate(tilts):
        scratch.zero_()
        A.forward_project(vg, target, A.rotate(base, float(t)), projs[i], scratch)
    r = A.BlockIterativeReconstructionOperator(volume=vg, base_geometry=base, tilt_angles=tilts, projections=projs,
                                               algorithm="sart", lam=lam, projection_iteration="random",
                                               voxel_type=voxel_type, init_value=init, flags=flags)
    assert r.uses_fused_path()
    r.run_iterations(1)
    return r.get_reconstructed_volume(dtype=torch.float32).double().cpu().numpy()


@pytest.mark.parametrize("voxel_type,eps", [("bfloat16", 4e-3), ("half", 5e-4)])
def test_stochastic_rounding_keeps_sub_ulp_updates_on_average(voxel_type, eps):
    # bf16 ulp at 1 is 2^-7, f16's 2^-10: every update (about 0.3 eps) is
    # below half an ulp
    f32 = small_update_runs("float", eps)
    rne = small_update_runs(voxel_type, eps)
    sr = small_update_runs(voxel_type, eps, flags=E.ETN_RECON_FLAG_STOCHASTIC_ROUNDING)
    gain = f32.mean() - 1.0
    assert gain > 0.5 * eps                  # the f32 volume moves towards 1 + eps
    assert np.all(rne == 1.0)                # round to nearest drops every update
    assert np.any(sr != 1.0)
    assert abs((sr.mean() - 1.0) - gain) <= 0.05 * gain, (sr.mean() - 1.0, gain)


def test_stochastic_rounding_is_reproducible():
    a = small_update_runs("bfloat16", 4e-3, flags=E.ETN_RECON_FLAG_STOCHASTIC_ROUNDING)
    b = small_update_runs("bfloat16", 4e-3, flags=E.ETN_RECON_FLAG_STOCHASTIC_ROUNDING)
    np.testing.assert_array_equal(a, b)


def test_stochastic_rounding_needs_the_fast_16bit_kernels():
    vg = E.VolumeGeometry.make((32, 32, 32))
    base = E.ScannerGeometry.mrc_base(32, 32)
    tilts = S.tilt_series(60.0, 10.0)
    projs = torch.zeros((len(tilts), 32, 32), dtype=torch.float32, device="cuda")
    for vt, extra in (("float", 0), ("bfloat16", E.ETN_RECON_FLAG_EXACT_ARITHMETIC),
                      ("bfloat16", E.ETN_RECON_FLAG_GENERIC_KERNELS)):
        with pytest.raises(E.EtnError):
            A.BlockIterativeReconstructionOperator(volume=vg, base_geometry=base, tilt_angles=tilts,
                                                   projections=projs, voxel_type=vt,
                                                   flags=E.ETN_RECON_FLAG_STOCHASTIC_ROUNDING | extra)


def _case(n=64, seed=21):
    vg = E.VolumeGeometry.make((n, n, n))
    base = E.ScannerGeometry.mrc_base(n, n)
    tilts = S.tilt_series(60.0, 4.0)
    ph = S.ellipsoid_phantom((n, n, n), seed=seed, device=torch.device("cuda"))
    projs = torch.zeros((len(tilts), n, n), dtype=torch.float32, device="cuda")
    scratch = torch.zeros((n, n), dtype=torch.float32, device="cuda")
    for i, t in enumerate(tilts):
        scratch.zero_()
        A.forward_project(vg, ph, A.rotate(base, float(t)), projs[i], scratch)
    return vg, base, tilts, S.add_gaussian_noise(projs, 0.02, seed=seed)


@pytest.mark.parametrize("voxel_type", ["bfloat16", "half"])
@pytest.mark.parametrize("flags", [0, E.ETN_RECON_FLAG_EXACT_ARITHMETIC, E.ETN_RECON_FLAG_GENERIC_KERNELS])
def test_16bit_projections_equal_f32_projections_of_the_same_values(voxel_type, flags):
    """Measured projections stored in the volume's 16-bit type give the same
    volume and residual norm, bit for bit, as f32 projections holding the same
    (rounded) values: only the storage differs."""
    vg, base, tilts, projs = _case()
    p16 = projs.to(torch.bfloat16 if voxel_type == "bfloat16" else torch.float16).contiguous()
    p32 = p16.float().contiguous()
    out = []
    for p in (p16, p32):
        r = A.BlockIterativeReconstructionOperator(volume=vg, base_geometry=base, tilt_angles=tilts, projections=p,
                                                   algorithm="sart", lam=0.3, projection_iteration="random",
                                                   voxel_type=voxel_type, flags=flags)
        assert r.uses_fused_path() == (flags != E.ETN_RECON_FLAG_GENERIC_KERNELS)
        r.run_iterations(2)
        out.append((r.get_reconstructed_volume(dtype=torch.float32), r.residual_norm()))
    assert torch.equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


def test_16bit_projections_need_16bit_storage():
    vg, base, tilts, projs = _case(32)
    with pytest.raises(ValueError):
        A.BlockIterativeReconstructionOperator(volume=vg, base_geometry=base, tilt_angles=tilts,
                                               projections=projs.to(torch.bfloat16), voxel_type="float")
    with pytest.raises(ValueError):   # the projections take the volume's 16-bit type
        A.BlockIterativeReconstructionOperator(volume=vg, base_geometry=base, tilt_angles=tilts,
                                               projections=projs.to(torch.float16), voxel_type="bfloat16")
