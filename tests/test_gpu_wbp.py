"""Weighted back projection on the GPU (algorithm "wbp", etn_ramp_filter)
against the oracle (oracle/wbp_oracle.py: the ramp filter in float64, the C
oracle's ray lengths and back projections).  Bounds: the filter within 1e-6
relative L2 (float sums of 128 taps against float64), the volume within 1e-5
relative L2."""
import os

import numpy as np
import pytest

import ettention_amd as E
from ettention_amd import api as A
from ettention_amd import io as EI
from ettention_amd._lib import Unsupported
from oracle import oracle as O
from oracle import wbp_oracle as W
from tests._util import load_golden, rel_l2

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TEST_BASE = E.ScannerGeometry.make((-64, -64, -128), (-64, -64, 128), (0.5, 0, 0), (0, 0.5, 0), (256, 256))


@pytest.fixture(scope="module")
def synthetic():
    s = load_golden("synthetic_3_parallel")
    return s["angles"], np.ascontiguousarray(s["projections"])


@pytest.mark.parametrize("shape", [(256, 256), (64, 100), (37, 101), (1, 200), (300, 520)])
def test_ramp_filter_matches_oracle(shape):
    print("ramp filter", shape, flush=True)
    img = np.random.default_rng(sum(shape)).normal(0, 1, shape).astype(np.float32)
    src = torch.from_numpy(img).cuda()
    out = torch.empty_like(src)
    A.ramp_filter(src, out)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), W.ramp_filter(img)) < 1e-6


def test_wbp_matches_oracle(synthetic):
    tilts, projs = synthetic
    sub = slice(0, 180, 3)
    t, p = tilts[sub], projs[sub]
    vg = E.VolumeGeometry.make((64, 64, 64))
    r = A.WeightedBackProjection(volume=vg, base_geometry=TEST_BASE, tilt_angles=t,
                                 projections=torch.from_numpy(np.ascontiguousarray(p)).cuda(), lam=1.0)
    r.run()
    got = r.get_reconstructed_volume().cpu().numpy()
    ov = O.Volume.make((64, 64, 64))
    ob = O.Scanner.make(TEST_BASE.source, TEST_BASE.detector, TEST_BASE.horizontal_pitch, TEST_BASE.vertical_pitch,
                        TEST_BASE.resolution)
    ref = W.wbp(ov, ob, t, p, lam=1.0)
    assert rel_l2(got, ref) < 1e-5
    with pytest.raises(Unsupported):
        r.residual_norm()


def test_wbp_partial_runs_and_options(synthetic):
    tilts, projs = synthetic
    sub = slice(1, 180, 9)
    t, p = tilts[sub], np.ascontiguousarray(projs[sub])
    vg = E.VolumeGeometry.make((48, 40, 24), subvolumes=3)
    dp = torch.from_numpy(p).cuda()
    r = A.WeightedBackProjection(volume=vg, base_geometry=TEST_BASE, tilt_angles=t, projections=dp, lam=0.5,
                                 bp_samples=4, init_value=0.25)
    r.run_iterations(5)          # WBP::runNiterations: projections 0..4 only
    got = r.get_reconstructed_volume().cpu().numpy()
    ov = O.Volume.make((48, 40, 24))
    ob = O.Scanner.make(TEST_BASE.source, TEST_BASE.detector, TEST_BASE.horizontal_pitch, TEST_BASE.vertical_pitch,
                        TEST_BASE.resolution)
    ref = np.full((24, 40, 48), 0.25, np.float32)
    n = len(t)
    for i in range(5):
        g = O.rotate(ob, float(t[i]))
        f = W.ramp_filter(p[i]).astype(np.float32)
        O.back_project(ov, ref, g, f, O.ray_length(ov, g), lam=0.5, block_size=n, samples=4)
    assert rel_l2(got, ref) < 1e-5


def test_wbp_parameter_file(tmp_path, synthetic):
    """The WBP plugin test's parameter file layout (WBPPluginTest.cpp, wbp.xml):
    an unsigned8 MRC stack with a tilt file, volume 128 x 128 x 20, x-y-z output."""
    tilts, projs = synthetic
    sub = slice(30, 150, 3)
    t = tilts[sub]
    stack = np.clip(projs[sub][:, ::2, ::2] * 4.0, 0, 255).astype(np.uint8)       # 128 x 128 bytes
    h = np.zeros(256, np.int32)
    h[:4] = (128, 128, len(t), 0)
    with open(tmp_path / "stack.mrc", "wb") as f:
        f.write(h.tobytes() + stack.tobytes())
    (tmp_path / "stack.tlt").write_text("\n".join("%g" % a for a in t))
    (tmp_path / "wbp.xml").write_text(
        '<algorithm value="wbp"><numberOfIterations value="1"/><lambda value="1.0"/></algorithm>'
        '<input><projections value="stack.mrc"/><tiltAngles value="stack.tlt"/></input>'
        '<output><filename value="out.mrc"/><options><voxelType value="float32"/>'
        '<orientation value="xyz_order"/></options></output>'
        '<volume><dimension value="128 128 20"/><initValue value="0.0"/></volume>'
        '<wbp><filterResolution value="128 1 1"/></wbp>')
    print("etn --config wbp.xml", flush=True)
    assert EI.main(["--config", str(tmp_path / "wbp.xml")]) == 0
    out = EI.read_mrc_volume(str(tmp_path / "out.mrc"), "xyz_order")
    r = A.WeightedBackProjection(volume=E.VolumeGeometry.make((128, 128, 20)),
                                 base_geometry=E.ScannerGeometry.mrc_base(128, 128),
                                 tilt_angles=np.array(["%g" % a for a in t], np.float32),
                                 projections=torch.from_numpy(stack.astype(np.float32)).cuda(), lam=1.0)
    r.run()
    np.testing.assert_array_equal(out, r.get_reconstructed_volume().cpu().numpy())


@pytest.mark.parametrize("kw", [dict(), dict(bp_samples=9), dict(volume=E.VolumeGeometry.make((64, 64, 64), subvolumes=4)),
                                dict(init_value=0.5, lam=0.7)])
def test_wbp_fused_back_projections_match_generic(synthetic, kw):
    """Single-axis WBP back-projects on the fused kernels, eight projections
    per volume sweep (default arithmetic): within 1e-6 of the generic kernels'
    one-launch-per-projection WBP; the generic kernels and exact arithmetic
    keep the unfused path."""
    tilts, projs = synthetic
    sub = slice(0, 180, 4)
    args = dict(volume=E.VolumeGeometry.make((64, 64, 64)), base_geometry=TEST_BASE, tilt_angles=tilts[sub],
                projections=torch.from_numpy(np.ascontiguousarray(projs[sub])).cuda(), lam=1.0)
    args.update(kw)
    vols = {}
    for name, flags in (("fused", 0), ("generic", E.ETN_RECON_FLAG_GENERIC_KERNELS),
                        ("exact", E.ETN_RECON_FLAG_EXACT_ARITHMETIC)):
        r = A.WeightedBackProjection(flags=flags, **args)
        assert r.uses_fused_path() == (name == "fused")
        r.run()
        vols[name] = r.get_reconstructed_volume().cpu().numpy()
    assert rel_l2(vols["fused"], vols["generic"]) < 1e-6
    np.testing.assert_array_equal(vols["exact"], vols["generic"])
    xt = A.WeightedBackProjection(x_axis_tilt=3.0, **args)
    assert not xt.uses_fused_path()


def test_wbp_rounds_of_plane_groups():
    # 1024 one-plane groups per launch, more than the device holds at once
    # (the rounds are dispatched interleaved): within 1e-6 of the generic WBP
    nx, ny, nz = 512, 1024, 8
    vg = E.VolumeGeometry.make((nx, ny, nz))
    base = E.ScannerGeometry.mrc_base(nx, ny)
    tilts = np.linspace(-60.0, 60.0, 17).astype(np.float32)
    rng = np.random.default_rng(79)
    projs = torch.from_numpy(rng.standard_normal((len(tilts), ny, nx)).astype(np.float32)).cuda()
    vols = {}
    for name, flags in (("fused", 0), ("generic", E.ETN_RECON_FLAG_GENERIC_KERNELS)):
        r = A.WeightedBackProjection(volume=vg, base_geometry=base, tilt_angles=tilts, projections=projs, lam=1.0,
                                     flags=flags)
        assert r.uses_fused_path() == (name == "fused")
        r.run()
        vols[name] = r.get_reconstructed_volume().cpu().numpy()
    assert rel_l2(vols["fused"], vols["generic"]) < 1e-6
