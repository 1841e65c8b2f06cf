"""The oracle's geometry, storage and projection-order functions against the
reference's own unit-test expectations and closed forms (CPU only)."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def test_rotation_zero_is_identity_and_tilt_sign():
    base = O.test_base_scanner()
    g = O.rotate(base, 0.0)
    assert g.as_tuple() == base.as_tuple()
    # SingleAxisTiltRotator rotates by -tilt about y (SingleAxisTiltRotator.cpp:46):
    # a +90 deg tilt maps the base ray (0,0,1) to (1,0,0) up to rounding.
    r = O.ray_direction(O.rotate(base, 90.0))
    np.testing.assert_allclose(r, [1.0, 0.0, 0.0], atol=1e-6)
    r = O.ray_direction(O.rotate(base, 45.0))
    np.testing.assert_allclose(r, [math.sqrt(0.5), 0.0, math.sqrt(0.5)], atol=1e-6)


def test_rotation_f64_agrees_with_f32():
    base = O.mrc_base_scanner(512, 512)
    for t in (-60.0, -13.0, 0.0, 27.5, 60.0):
        g = O.rotate(base, t)
        gd = O.rotate_f64(base, t)
        f = np.concatenate([g.source, g.detector, g.hpitch, g.vpitch])
        np.testing.assert_allclose(f, gd, rtol=0, atol=1e-3 * 512 * 2e-3 + 1e-5)


def test_projection_matrix_maps_points_to_detector_coordinates():
    # u = a + nu/2 + 1/2 where p = a*h + b*v + c*r (ParallelScannerGeometry.cpp:49-62);
    # checked against a double-precision solve for random points and tilts.
    rng = np.random.default_rng(1)
    base = O.test_base_scanner()
    for t in (0.0, 17.0, 45.0, -60.0, 90.0):
        g = O.rotate(base, t)
        M = O.projection_matrix(g).reshape(4, 4).T   # values[4*col+row] -> row-major
        r = O.ray_direction(g).astype(np.float64)
        A = np.stack([np.array(g.hpitch, np.float64), np.array(g.vpitch, np.float64), r], axis=1)
        for _ in range(5):
            p = rng.uniform(-40, 40, 3)
            a, b, _c = np.linalg.solve(A, p)
            uvw = M @ np.append(p, 1.0)
            assert abs(uvw[0] / uvw[3] - (a + 128.5)) < 2e-3
            assert abs(uvw[1] / uvw[3] - (b + 128.5)) < 2e-3


def test_projection_matrix_zero_degrees_closed_form():
    g = O.rotate(O.test_base_scanner(), 0.0)
    M = O.projection_matrix(g).reshape(4, 4).T
    # h = (0.5,0,0) -> u = 2x + 128.5 ; v = 2y + 128.5
    np.testing.assert_allclose(M[0], [2.0, 0.0, 0.0, 128.5], atol=1e-6)
    np.testing.assert_allclose(M[1], [0.0, 2.0, 0.0, 128.5], atol=1e-6)


def test_bp_sample_pattern_closed_forms():
    # BackprojectionTest.cpp:48-102
    base = O.test_base_scanner()
    vs = (1.0, 1.0, 1.0)
    g0 = O.rotate(base, 0.0)
    s1 = O.bp_sample_pattern(g0, vs, 1)
    np.testing.assert_array_equal(s1, [[0.0, 0.0, 0.0, 1.0]])
    s4 = O.bp_sample_pattern(g0, vs, 4)
    np.testing.assert_allclose(s4, [[-0.25, -0.25, 0, 1], [0.25, -0.25, 0, 1],
                                    [-0.25, 0.25, 0, 1], [0.25, 0.25, 0, 1]], atol=1e-7)
    g45 = O.rotate(base, 45.0)
    s = O.bp_sample_pattern(g45, vs, 1)
    np.testing.assert_allclose(s[0, :3], 0.0, atol=1e-7)
    assert abs(s[0, 3] - math.sqrt(2.0)) < 4 * np.finfo(np.float32).eps * math.sqrt(2.0)
    for bad in (0, 3, 5):
        with pytest.raises(ValueError):
            O.bp_sample_pattern(g0, vs, bad)


def test_traversal_length_long_object_constant():
    # LongObjectCompensationTest.cpp:27-33,129-141: c = 32 / cos(80 deg), |err| < 0.01
    base = O.Scanner.make((-16, -16, -256), (-16, -16, 256), (0.125, 0, 0), (0, 0.125, 0), (256, 256))
    g = O.rotate(base, 80.0)
    c = O.traversal_length_bbox(g, (-16, -16, -16), (16, 16, 16))
    assert abs(c - 32.0 / math.cos(math.radians(80.0))) < 0.01


def test_slabs_split_like_volume_properties():
    # VolumeProperties::setSubVolumeCount (VolumeProperties.cpp:74-98)
    v = O.Volume.make((8, 6, 10), (-4, -3, -5), (4, 3, 5), slabs=4)
    sizes, z0s, prev_max = [], [], -5.0
    for i in range(4):
        sres, z0, smin, smax = O.slab(v, i)
        sizes.append(int(sres[2]))
        z0s.append(z0)
        assert smin[2] == prev_max
        prev_max = smax[2]
    assert sizes == [3, 3, 2, 2]
    assert z0s == [0, 3, 6, 8]
    assert prev_max == 5.0
    with pytest.raises(ValueError):
        O.slab(O.Volume.make((4, 4, 3), slabs=4), 0)


def test_mt19937_known_answer():
    # C++ standard [rand.predef]: 10000th output of default-seeded mt19937.
    assert O.mt19937_nth(5489, 10000) == 4123659995


def test_projection_orders_closed_forms():
    # ProjectionSetTest.cpp:30-58 and the reorder() bodies
    np.testing.assert_array_equal(O.projection_order("identity", 5), [0, 1, 2, 3, 4])
    np.testing.assert_array_equal(O.projection_order("identity", 5, skip=[0, 2, 4]), [1, 3])
    np.testing.assert_array_equal(O.projection_order("maxangle", 5), [2, 0, 3, 1, 4])
    np.testing.assert_array_equal(O.projection_order("maxangle", 6), [3, 0, 4, 1, 5, 2])
    np.testing.assert_array_equal(O.projection_order("buckets", 8), [0, 2, 4, 6, 1, 3, 5, 7])
    with pytest.raises(ValueError):
        O.projection_order("nonsense", 3)


def test_buckets_six_by_hand():
    # current_idx = {0, 1, 3, 4} for n = 6 (i*6/4 in int math); segments visited round
    # robin, so for n not divisible by 4 the reference repeats indices (reference quirk).
    np.testing.assert_array_equal(O.projection_order("buckets", 6), [0, 1, 3, 4, 1, 2])


def test_random_order_is_a_permutation_and_ignores_skip_list():
    # ProjectionSetTest.cpp:67-83 (permutation); ProjectionSetRandom.cpp:33-36 resets
    # the list to 0..n-1, so skipped indices reappear (reference quirk).
    for n in (1, 2, 5, 61, 121):
        p = O.projection_order("random", n)
        assert sorted(p.tolist()) == list(range(n))
    p = O.projection_order("random", 6, skip=[0, 5])
    assert sorted(p.tolist()) == [0, 1, 2, 3]
    # the first swap uses uniform_int(0,1) = first mt19937 output / bucket(2^31) -> its top bit
    first = O.mt19937_nth(5489, 1)
    p2 = O.projection_order("random", 2)
    assert p2.tolist() == ([1, 0] if (first >> 31) == 0 else [0, 1])


def test_half_and_bfloat_conversions():
    lib = O.lib()
    # round to nearest even at the half-way point between 1 and 1+2^-10
    assert lib.orc_f32_to_f16_rte(1.0 + 2.0 ** -11) == 0x3C00
    assert lib.orc_f32_to_f16_rte(1.0 + 3 * 2.0 ** -11) == 0x3C02
    assert lib.orc_f32_to_f16_trunc(1.0 + 3 * 2.0 ** -11) == 0x3C01
    assert lib.orc_f32_to_f16_rte(65519.0) == 0x7BFF and lib.orc_f32_to_f16_rte(65520.0) == 0x7C00
    assert lib.orc_f32_to_f16_rte(2.0 ** -24) == 0x0001
    rng = np.random.default_rng(3)
    for x in rng.uniform(-1000, 1000, 200).astype(np.float32):
        h = lib.orc_f32_to_f16_rte(float(x))
        assert np.float16(x).view(np.uint16) == h
        assert lib.orc_f16_to_f32(h) == np.float32(np.float16(x))
        b = lib.orc_f32_to_bf16_rne(float(x))
        # bf16 RNE: |x - bf16(x)| <= half an ulp of bf16
        back = lib.orc_bf16_to_f32(b)
        assert abs(back - x) <= 2.0 ** (math.floor(math.log2(abs(x))) - 8) + 1e-12
