"""The oracle's x-axis tilt rotation and the DDA step limit, pinned by closed
forms (CPU only).

* SingleAxisTiltRotator::doRotate (ettention/libReconstruction/src/model/
  geometry/SingleAxisTiltRotator.cpp:62-78).  With xAxisTilt = 0 it turns a
  point by t = -tilt about y as x' = x cos t - z sin t, z' = x sin t + z cos t.
  Otherwise it first turns the point about x by the x tilt a
  (y' = y cos a + z sin a, z' = -y sin a + z cos a) and then applies the
  Rodrigues rotation by t about the tilted axis R_x(a) e_y.  Since
  Rot(R_x e_y, t) = R_x Rot(e_y, t) R_x^-1, the composite is R_x(a) Rot(e_y, t):
  a rotation about y in the Rodrigues sense (x' = x cos t + z sin t, the
  opposite sense of the a = 0 branch) followed by the turn about x.  The
  float64 matrices below are built from that product, not from the
  reference's expression.
* DDARaytracer.hl:85-108: the loop counter is incremented in the condition and
  at the end of the body and stops at 4096, so a ray takes at most 2048 steps.
  On a constant-1 volume the traced length is then the time of the ray's
  2048th voxel-border crossing minus its entry time, which the tests compute in
  float64 from the sorted crossing times of the x and z grid planes.
"""
import math

import numpy as np

from oracle import oracle as O


def r_x(a):
    c, s = math.cos(a), math.sin(a)
    # y' = y c + z s, z' = -y s + z c
    return np.array([[1, 0, 0], [0, c, s], [0, -s, c]], np.float64)


def r_y_rodrigues(t):
    # rotation by t about +y: x' = x cos t + z sin t, z' = -x sin t + z cos t
    c, s = math.cos(t), math.sin(t)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]], np.float64)


def fields(g):
    return [np.array(g.source, np.float64), np.array(g.detector, np.float64),
            np.array(g.hpitch, np.float64), np.array(g.vpitch, np.float64)]


def test_x_axis_tilt_is_a_turn_about_y_then_about_x():
    base = O.mrc_base_scanner(256, 192)
    for xt in (5.0, -10.0, 30.0, 0.5):
        for tilt in (-60.0, -13.0, 0.0, 27.5, 60.0):
            t = -np.float32(tilt) / np.float32(180.0) * np.float32(math.pi)
            a = np.float32(xt) / np.float32(180.0) * np.float32(math.pi)
            R = r_x(float(a)) @ r_y_rodrigues(float(t))
            g = O.rotate(base, tilt, xt)
            for got, b in zip(fields(g), fields(base)):
                want = R @ b
                scale = max(1.0, float(np.abs(b).max()))
                np.testing.assert_allclose(got, want, rtol=0, atol=4e-6 * scale, err_msg=str((xt, tilt)))


def test_x_axis_tilt_preserves_lengths_and_angles():
    base = O.mrc_base_scanner(256, 256)
    b = fields(base)
    for xt in (7.0, -25.0):
        for tilt in (-45.0, 10.0, 60.0):
            f = fields(O.rotate(base, tilt, xt))
            for i in range(4):
                for k in range(4):
                    scale = max(1.0, np.linalg.norm(b[i]) * np.linalg.norm(b[k]))
                    assert abs(f[i] @ f[k] - b[i] @ b[k]) <= 2e-6 * scale


def test_zero_x_tilt_branch_turns_the_other_way():
    # with a = 0 the reference uses its first branch: x' = x cos t - z sin t
    base = O.mrc_base_scanner(128, 128)
    for tilt in (-30.0, 20.0):
        t = -np.float32(tilt) / np.float32(180.0) * np.float32(math.pi)
        R = r_y_rodrigues(-float(t))
        for got, b in zip(fields(O.rotate(base, tilt, 0.0)), fields(base)):
            np.testing.assert_allclose(got, R @ b, rtol=0, atol=4e-6 * max(1.0, float(np.abs(b).max())))


def crossing_length(src, d, bmin, bmax, steps):
    """float64 traversal length of a ray from src along unit d (d.y = 0) through
    the unit grid of the box, cut after `steps` voxel-border crossings."""
    t0, t1 = -np.inf, np.inf
    for k in (0, 2):
        if d[k] != 0:
            a, b = (bmin[k] - src[k]) / d[k], (bmax[k] - src[k]) / d[k]
            t0, t1 = max(t0, min(a, b)), min(t1, max(a, b))
        elif not (bmin[k] <= src[k] <= bmax[k]):
            return 0.0
    if t1 - t0 <= 1e-4:
        return 0.0
    times = []
    for k in (0, 2):
        if d[k] == 0:
            continue
        planes = np.arange(math.ceil(bmin[k]), math.floor(bmax[k]) + 1, dtype=np.float64)
        tk = (planes - src[k]) / d[k]
        times.append(tk[(tk > t0 + 1e-9) & (tk < t1 - 1e-9)])
    times = np.sort(np.concatenate(times + [np.array([t1])]))
    end = times[steps - 1] if len(times) >= steps else t1
    return float(end - t0)


def test_step_limit_is_2048_crossings():
    # a thin slab (1152 x 2 x 1152): rays at 40 deg cross up to ~2100 borders
    nx, ny, nz = 1152, 2, 1152
    v = O.Volume.make((nx, ny, nz))
    base = O.mrc_base_scanner(nx, ny)
    ones = np.ones((nz, ny, nx), np.float32)
    capped_rays = 0
    for tilt in (40.0, -40.0):
        g = O.rotate(base, tilt)
        _, length = O.forward_project(v, ones, g)
        d = np.array(O.ray_direction(g), np.float64)
        bmin = np.array([-nx / 2, -ny / 2, -nz / 2], np.float64)
        bmax = -bmin
        src0 = np.array(g.source, np.float64)
        h = np.array(g.hpitch, np.float64)
        vv = np.array(g.vpitch, np.float64)
        for u in range(0, nx, 7):
            src = src0 + h * (u + 0.5) + vv * 0.5
            full = crossing_length(src, d, bmin, bmax, 10 ** 9)
            cut = crossing_length(src, d, bmin, bmax, 2048)
            got = float(length[0, u])
            assert abs(got - cut) <= 2e-4 * max(cut, 1.0), (tilt, u, got, cut, full)
            if full - cut > 1.0:
                capped_rays += 1
    # the limit binds on a good share of the rays
    assert capped_rays > 40
