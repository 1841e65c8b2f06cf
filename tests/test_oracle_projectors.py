"""The oracle's forward/back projection, residual, long-object compensation
and ray length against the reference's golden fixtures and against closed
forms (CPU only)."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests._util import (bp_fixture, image_comparator_ok, load_golden, max_slice_rms,
                         projection_test_phantom, rel_l2)

V64 = O.Volume.make((64, 64, 64))


@pytest.fixture(scope="module")
def fp_phantom():
    return projection_test_phantom(1.0)


def test_fp_45_matches_reference_fixture(fp_phantom):
    # ForwardProjectionTest::Unit_ForwardProjection (ForwardProjectionTest.cpp:51-77)
    ref = load_golden("forward_projection")["projection_45"]
    p, _ = O.forward_project(V64, fp_phantom, O.rotate(O.test_base_scanner(), 45.0))
    assert image_comparator_ok(p, ref)
    assert rel_l2(p, ref) < 2e-6


def test_fp_opposite_ray_sign(fp_phantom):
    # Unit_ForwardProjection_Raysign: 225 deg equals the 45 deg fixture (:79-95)
    ref = load_golden("forward_projection")["projection_45"]
    p, _ = O.forward_project(V64, fp_phantom, O.rotate(O.test_base_scanner(), 225.0))
    assert image_comparator_ok(p, ref)
    assert rel_l2(p, ref) < 3e-6


def test_fp_zero_degrees_fixture(fp_phantom):
    ref = load_golden("forward_projection")["reference_0"]
    p, _ = O.forward_project(V64, fp_phantom, O.rotate(O.test_base_scanner(), 0.0))
    np.testing.assert_array_equal(p, ref)


def test_traversal_length_and_ray_length_fixture(fp_phantom):
    # Unit_ComputeRayLength (RMS < 0.1) and the FP traversal length output
    ref = load_golden("forward_projection")["raylength_reference"]
    g = O.rotate(O.test_base_scanner(), 45.0)
    rl = O.ray_length(V64, g)
    assert float(np.sqrt(np.mean((rl - ref) ** 2))) < 0.1
    assert rel_l2(rl, ref) < 1e-6
    _, length = O.forward_project(V64, fp_phantom, g)
    assert rel_l2(length, ref) < 2e-6


def test_fp_synthetic_dataset_goldens():
    # 180 golden FP outputs of the reference on ground_truth.mrc (SyntheticReconstruction.cpp:66-107)
    s = load_golden("synthetic_3_parallel")
    gt = np.ascontiguousarray(s["ground_truth"])
    base = O.test_base_scanner()
    worst = 0.0
    for a, ref in zip(s["angles"][::7], s["projections"][::7]):
        p, _ = O.forward_project(V64, gt, O.rotate(base, float(a)))
        worst = max(worst, rel_l2(p, ref))
    assert worst < 2e-6


def test_fp_dda_equals_exact_line_integral():
    # Closed form: the DDA with nearest sampling integrates the piecewise-constant
    # field exactly; compare with an independent Siddon integral in double.
    rng = np.random.default_rng(7)
    v = O.Volume.make((12, 10, 8), (-6, -5, -4), (6, 5, 4))
    vol = rng.uniform(0, 1, (8, 10, 12)).astype(np.float32)
    base = O.Scanner.make((-8, -8, 20), (-8, -8, -20), (0.7, 0, 0), (0, 0.9, 0), (24, 18))
    for t in (0.0, 12.0, -37.0, 61.0, 90.0):
        g = O.rotate(base, t)
        p, length = O.forward_project(v, vol, g)
        r = O.ray_direction(g).astype(np.float64)
        for (u, w) in [(3, 4), (11, 9), (17, 13), (9, 2), (20, 16)]:
            o = (np.array(g.source, np.float64) + np.array(g.hpitch, np.float64) * (u + 0.5)
                 + np.array(g.vpitch, np.float64) * (w + 0.5))
            val, chord = O.siddon(v, vol.astype(np.float64), o, r)
            assert abs(p[w, u] - val) <= 2e-5 * max(1.0, abs(val))
            assert abs(length[w, u] - chord) <= 2e-5 * max(1.0, chord)


def test_fp_dda_world_space_border_quirk():
    # DDARaytracer.hl:35 measures the distance to the next voxel border with
    # floor(pos / vs) in world space, so with bbox faces that are not multiples of the
    # voxel size the traversal deviates from the exact integral (reference quirk).
    v = O.Volume.make((4, 4, 4), (-2, -2, -1.5), (2, 2, 2.5))
    vol = np.ones((4, 4, 4), np.float32)
    vol[1] = 5.0
    g = O.rotate(O.Scanner.make((-2, -2, 10), (-2, -2, -10), (0.5, 0, 0), (0, 0.5, 0), (8, 8)), 30.0)
    p, _ = O.forward_project(v, vol, g)
    r = O.ray_direction(g).astype(np.float64)
    o = np.array(g.source, np.float64) + np.array(g.hpitch, np.float64) * 4.5 + np.array(g.vpitch, np.float64) * 4.5
    exact, _ = O.siddon(v, vol.astype(np.float64), o, r)
    assert abs(p[4, 4] - exact) > 0.05


def test_fp_constant_volume_equals_traversal_length_with_subsamples():
    # PerspectiveIntersectionLengthAccumulation logic (ForwardProjectionTest.cpp:163-183):
    # for a volume of ones, the projection equals the traversal length.
    v = O.Volume.make((16, 16, 16))
    vol = np.ones((16, 16, 16), np.float32)
    g = O.rotate(O.Scanner.make((-16, -16, 40), (-16, -16, -40), (0.5, 0, 0), (0, 0.5, 0), (64, 64)), 23.0)
    for n in (1, 2, 4):
        p, length = O.forward_project(v, vol, g, samples=n)
        np.testing.assert_allclose(p, length, rtol=2e-6, atol=1e-5)
    # samples=2 is the mean of the four sub-ray chords
    p2, _ = O.forward_project(v, vol, g, samples=2)
    r = O.ray_direction(g).astype(np.float64)
    u, w = 30, 33
    acc = 0.0
    for i in range(2):
        for j in range(2):
            o = (np.array(g.source, np.float64) + np.array(g.hpitch, np.float64) * (u + 0.25 + 0.5 * i)
                 + np.array(g.vpitch, np.float64) * (w + 0.25 + 0.5 * j))
            acc += O.siddon(v, vol.astype(np.float64), o, r)[0]
    assert abs(p2[w, u] - acc / 4) < 1e-4


def test_fp_slabs_accumulate_to_full_volume(fp_phantom):
    g = O.rotate(O.test_base_scanner(), 30.0)
    full, lfull = O.forward_project(V64, fp_phantom, g)
    v4 = O.Volume.make((64, 64, 64), slabs=4)
    p = np.zeros((256, 256), np.float32)
    length = np.zeros((256, 256), np.float32)
    for s in range(4):
        O.forward_project(v4, fp_phantom, g, slab=s, proj=p, length=length)
    assert rel_l2(p, full) < 1e-6
    assert rel_l2(length, lfull) < 1e-6


def test_fp_mask_scales_traversal_length(fp_phantom):
    g = O.rotate(O.test_base_scanner(), 10.0)
    p0, l0 = O.forward_project(V64, fp_phantom, g)
    p1, l1 = O.forward_project(V64, fp_phantom, g, mask=np.full((64, 64, 64), 255, np.uint8))
    np.testing.assert_array_equal(p0, p1)
    np.testing.assert_array_equal(l0, l1)
    p2, l2 = O.forward_project(V64, fp_phantom, g, mask=np.zeros((64, 64, 64), np.uint8))
    np.testing.assert_array_equal(p0, p2)
    assert float(np.abs(l2).max()) == 0.0


def test_fp32_and_fp64_forward_projection_agree(fp_phantom):
    base = O.test_base_scanner()
    for t in (-60.0, 5.0, 44.0):
        p, _ = O.forward_project(V64, fp_phantom, O.rotate(base, t))
        pd, _ = O.forward_project_f64(V64, fp_phantom.astype(np.float64), O.rotate_f64(base, t), (256, 256))
        assert rel_l2(p, pd) < 2e-6


@pytest.mark.parametrize("tag,angle", [("z", 0.0), ("y", 90.0), ("45", 45.0)])
def test_bp_matches_reference_fixtures(tag, angle):
    # BackprojectionTest ZDirection / YDirection / 45Direction (BackprojectionTest.cpp:33-132)
    proj, ref = bp_fixture(tag)
    g = O.rotate(O.test_base_scanner(), angle)
    rl = O.ray_length(V64, g)
    out = np.zeros((64, 64, 64), np.float32)
    O.back_project(V64, out, g, proj, rl)
    assert max_slice_rms(out, ref) <= 0.011
    assert float(np.sqrt(np.mean((out - ref) ** 2))) < 1e-5
    assert int((out != 0).sum()) == int((ref != 0).sum())


def test_bp_subvolumes_match_reference_fixture():
    # Subvolume_BackProjectionTest: 4 z-slabs against the 45 deg reference (:48-70)
    proj, ref = bp_fixture("45")
    g = O.rotate(O.test_base_scanner(), 45.0)
    v4 = O.Volume.make((64, 64, 64), slabs=4)
    rl = O.ray_length(v4, g)
    out = np.zeros((64, 64, 64), np.float32)
    for s in (3, 2, 1, 0):
        O.back_project(v4, out, g, proj, rl, slab=s)
    assert float(np.sqrt(np.mean((out - ref) ** 2))) < 1e-5


def test_bp_half_storage_within_reference_tolerance():
    # BackprojectionTest_HalfFloat16 compares FP16 storage with the FP32 reference (RMS <= 0.011)
    proj, ref = bp_fixture("45")
    g = O.rotate(O.test_base_scanner(), 45.0)
    rl = O.ray_length(V64, g)
    out = np.zeros((64, 64, 64), np.uint16)
    O.back_project(V64, out, g, proj, rl, voxel_type=O.VOXEL_F16)
    as_f32 = out.view(np.float16).astype(np.float32)
    assert max_slice_rms(as_f32, ref) <= 0.011
    assert float(np.abs(as_f32 - ref).max()) < 1e-3


def test_bp_lambda_over_block_size_and_mask():
    # lambda is divided by the block size at bind time (BackProjectionKernel.cpp:94); a
    # prior-knowledge mask of m/255 scales the update (ParallelBeamsBackProjection.cl:47-51).
    proj, _ = bp_fixture("45")
    g = O.rotate(O.test_base_scanner(), 45.0)
    rl = O.ray_length(V64, g)
    a = O.back_project(V64, np.zeros((64, 64, 64), np.float32), g, proj, rl, lam=1.0, block_size=1)
    b = O.back_project(V64, np.zeros((64, 64, 64), np.float32), g, proj, rl, lam=1.0, block_size=4)
    np.testing.assert_allclose(b, a / 4, rtol=1e-6, atol=1e-9)
    m = np.full((64, 64, 64), 51, np.uint8)   # 0.2
    c = O.back_project(V64, np.zeros((64, 64, 64), np.float32), g, proj, rl, mask=m)
    np.testing.assert_allclose(c, a * np.float32(51.0 / 255.0), rtol=1e-6, atol=1e-9)


def test_bp_single_sample_closed_form_against_numpy():
    # Closed form of one voxel update at 0 deg: u = 2x+128.5 -> bilinear mixes pixels
    # floor(u-0.5) and +1 with weight frac(u-0.5) = 0.5 (the +0.5 pixel quirk).
    g = O.rotate(O.test_base_scanner(), 0.0)
    rng = np.random.default_rng(11)
    res = rng.uniform(-1, 1, (256, 256)).astype(np.float32)
    rl = O.ray_length(V64, g)
    out = O.back_project(V64, np.zeros((64, 64, 64), np.float32), g, res, rl)
    for (x, y, z) in [(0, 0, 0), (10, 20, 30), (63, 63, 63), (31, 5, 60)]:
        u = 2 * (x - 32 + 0.5) + 128.5
        vv = 2 * (y - 32 + 0.5) + 128.5
        i0, j0 = int(math.floor(u - 0.5)), int(math.floor(vv - 0.5))
        tx, ty = (u - 0.5) - i0, (vv - 0.5) - j0
        R = (1 - ty) * ((1 - tx) * res[j0, i0] + tx * res[j0, i0 + 1]) + ty * ((1 - tx) * res[j0 + 1, i0] + tx * res[j0 + 1, i0 + 1])
        expected = R * 1.0 / 64.0
        assert abs(out[z, y, x] - expected) < 1e-6


def test_bp_samples_average_used_samples():
    # nSamples = 4 averages over samples whose ray length exceeds EPSILON
    g = O.rotate(O.test_base_scanner(), 30.0)
    rl = O.ray_length(V64, g)
    res = np.ones((256, 256), np.float32)
    out = O.back_project(V64, np.zeros((64, 64, 64), np.float32), g, res, rl, samples=4)
    pd = O.back_project_f64(V64, np.zeros((64, 64, 64), np.float64), O.rotate_f64(O.test_base_scanner(), 30.0),
                            (256, 256), res.astype(np.float64), rl.astype(np.float64), samples=4)
    assert rel_l2(out, pd) < 1e-5


def test_compare_and_long_object_compensation_fixture():
    # Compare.cl: r = b - p ; LongObjectCompensationTest::Unit_Long_Object_Compensation
    rng = np.random.default_rng(5)
    b = rng.normal(size=1000).astype(np.float32)
    p = rng.normal(size=1000).astype(np.float32)
    np.testing.assert_array_equal(O.compare(b, p), b - p)
    f = load_golden("long_object_compensation")
    proj = f["input_virtual_projection"].copy()
    length = f["input_traversal_length"].copy()
    base = O.Scanner.make((-16, -16, -256), (-16, -16, 256), (0.125, 0, 0), (0, 0.125, 0), (256, 256))
    c = O.traversal_length_bbox(O.rotate(base, 80.0), (-16, -16, -16), (16, 16, 16))
    O.long_object_compensate(proj, length, c)
    ref = f["virtual_projection_reference"]
    assert float(np.sqrt(np.mean((proj - ref) ** 2))) < 0.1
    assert np.all(length == np.float32(c))
    # traversal length below EPSILON zeroes the projection
    pz = np.array([5.0, 5.0], np.float32)
    lz = np.array([1e-5, 2.0], np.float32)
    O.long_object_compensate(pz, lz, 3.0)
    np.testing.assert_array_equal(pz, [0.0, 7.5])


def test_long_object_inputs_regenerate_from_fp():
    # The LOC inputs were written by the reference FP with 16 sub-samples at 80 deg on a
    # 32^3 volume of ones (LongObjectCompensationTest.cpp:83-101).
    f = load_golden("long_object_compensation")
    base = O.Scanner.make((-16, -16, -256), (-16, -16, 256), (0.125, 0, 0), (0, 0.125, 0), (256, 256))
    v = O.Volume.make((32, 32, 32))
    p, length = O.forward_project(v, np.ones((32, 32, 32), np.float32), O.rotate(base, 80.0), samples=16)
    assert rel_l2(p, f["input_virtual_projection"]) < 1e-5
    assert rel_l2(length, f["input_traversal_length"]) < 1e-5
