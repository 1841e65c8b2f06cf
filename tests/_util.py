"""Helpers shared by the test modules (phantoms of the reference's unit tests,
fixture decoding, error measures)."""
from __future__ import annotations

import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def projection_test_phantom(background=1.0):
    """ProjectionTestBase::allocateVolume: 64^3 volume, radius-1 blobs of 16 at
    {16,32,48}^3 and radius-3 blobs on the diagonal (ProjectionTestBase.cpp:70-103);
    ForwardProjectionTest uses background 1 (ForwardProjectionTest.cpp:46-49).
    Layout [z][y][x]."""
    vol = np.full((64, 64, 64), background, np.float32)

    def blob(p, r, val):
        px, py, pz = (int(c) for c in p)
        ri = int(r)
        for z in range(pz - ri, pz + ri + 1):
            for y in range(py - ri, py + ri + 1):
                for x in range(px - ri, px + ri + 1):
                    d = np.sqrt(np.float64(np.float32(x - p[0]) ** 2 + np.float32(y - p[1]) ** 2
                                           + np.float32(z - p[2]) ** 2))
                    if d < r:
                        vol[z, y, x] = val

    for x in (16.0, 32.0, 48.0):
        for y in (16.0, 32.0, 48.0):
            for z in (16.0, 32.0, 48.0):
                blob((x, y, z), 1, 16.0)
    for c in (16.0, 32.0, 48.0):
        blob((c, c, c), 3, 16.0)
    return vol


def bp_fixture(tag):
    """Input residual image and expected 64^3 volume of BackprojectionTest
    (projection_<tag>.tif, voxel_by_voxel_reference_<tag>.mrc).  The MRC is
    written in XZY order (file[y][z][x]); returned as [z][y][x]."""
    b = load_golden("back_projection")
    proj = np.zeros(256 * 256, np.float32)
    proj[b["proj_%s_idx" % tag]] = b["proj_%s_val" % tag]
    ref = np.zeros(64 ** 3, np.float32)
    ref[b["vol_%s_idx" % tag]] = b["vol_%s_val" % tag]
    return proj.reshape(256, 256), np.ascontiguousarray(ref.reshape(64, 64, 64).transpose(1, 0, 2))


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


def image_comparator_ok(a, b):
    """ImageComparator::assertImagesAreEqual: RMS <= 0.1 and max squared
    difference <= 1e-5 * max intensity (ImageComparator.cpp:109-121)."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    rms = float(np.sqrt(np.mean((a - b).astype(np.float64) ** 2)))
    mx = max(float(a.max()), float(b.max()))
    return rms <= 0.1 and float(((a - b) ** 2).max()) <= mx * 1e-5


def max_slice_rms(vol_zyx, ref_zyx):
    """StackComparator::getMaxRMS over the file slices (y planes for XZY files),
    bound 0.011 (StackComparator.cpp:12,55-98)."""
    d = (np.asarray(vol_zyx, np.float64) - np.asarray(ref_zyx, np.float64))
    return float(np.sqrt((d ** 2).mean(axis=(0, 2))).max())


def ellipsoid_phantom(shape_zyx, n=20, seed=1000, beads=0, bead_density=5.0):
    """Seeded random-ellipsoid phantom (additive densities U[0.1, 1]); optional
    radius-1.5 beads.  numpy reference used by the tests to check the GPU
    phantom generator."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape_zyx
    z, y, x = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    vol = np.zeros(shape_zyx, np.float32)
    for _ in range(n):
        c = rng.uniform(0.1, 0.9, 3) * np.array([nx, ny, nz])
        ax = rng.uniform(0.05, 0.35, 3) * np.array([nx, ny, nz])
        d = rng.uniform(0.1, 1.0)
        inside = ((x - c[0]) / ax[0]) ** 2 + ((y - c[1]) / ax[1]) ** 2 + ((z - c[2]) / ax[2]) ** 2 <= 1.0
        vol[inside] += d
    for _ in range(beads):
        c = rng.uniform(0.1, 0.9, 3) * np.array([nx, ny, nz])
        inside = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= 1.5 ** 2
        vol[inside] = bead_density
    return vol
