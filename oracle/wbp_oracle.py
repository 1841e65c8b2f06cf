"""TEST INFRASTRUCTURE ONLY -- restatement of weighted back projection (the
reference's WBP plugin) for checking libetn.so: the ramp filter in float64
numpy, the ray lengths and back projections through the C oracle.

Only tests/ may import it.

WBP (ettention/plugins/WBPPlugin/src/reconstruction/WBP.cpp:13-84), per
projection i in stack order:
  1. filtered = ConvolutionOperator(RamLak kernel of 128 taps, projection i),
     the Fourier-space implementation
     (libReconstruction/src/algorithm/imagemanipulation/convolution/
     ConvolutionFourierSpaceImplementation.cpp:50-83);
  2. L = ComputeRayLengthKernel: pixel-centre chord through the volume box;
  3. BackProjectionOperator with (filtered, L, block size N): lambda / N.

The Fourier-space convolution, written out: the image is mirror-padded
(PaddingForFFT.cl:computePaddedCopyForFourierSpace) to U = powers of two of
max(filter, image); the 128 x 1 filter is zero-padded centred in U
(computeZeroPaddedCopyForFourierSpace) and its origin moved to the corner
(MoveOriginToCenter.cl); the product of the 2-D FFTs is transformed back and
cropped from the centre (CroppingForFFT.cl).  For even U that is the circular
convolution
    C[y][x] = sum_t k[t] * P[(y + 1) mod Uy][(x - (t - 64)) mod Ux]
-- the filter row lands on padded row -1, hence the y + 1 -- cropped at
(Ux/2 - nu/2, Uy/2 - nv/2).  tests/test_wbp_oracle.py checks this closed
form against the padded-FFT pipeline itself.
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as O

FILTER_SIZE = 128


def ramlak_taps(resolution=FILTER_SIZE):
    """GenerateRamLakKernel.cl: position p = i - resolution / 2; 1/4 at 0,
    0 at even p, -1 / (p^2 (float)(pi_f * pi_f)) at odd p, in float."""
    pi = np.float32(np.pi)
    pi2 = np.float32(pi * pi)
    k = np.zeros(resolution, np.float32)
    for i in range(resolution):
        p = i - resolution // 2
        if p == 0:
            k[i] = np.float32(0.25)
        elif p % 2 == 0:
            k[i] = np.float32(0.0)
        else:
            k[i] = np.float32(-1.0) / (np.float32(p * p) * pi2)
    return k


def next_pow2(v):
    """NumericalAlgorithms::nearestPowerOfTwo (NumericalAlgorithms.cpp:20-31)."""
    v = int(v) - 1
    for s in (1, 2, 4, 8, 16):
        v |= v >> s
    return v + 1


def padded_size(nu, nv, taps=FILTER_SIZE):
    return next_pow2(max(taps, nu)), next_pow2(max(1, nv))


def mirror_index(x, begin, end, size):
    """computePaddedCopyForFourierSpace, one axis (clamped where the reference
    would read outside the image)."""
    if begin <= x < end:
        m = x - begin
    elif x < end:
        m = begin - x
    else:
        m = (size + end - 1) - x
    return min(max(m, 0), size - 1)


def mirror_pad(img, ux, uy):
    nv, nu = img.shape
    bx, by = (ux - nu) // 2, (uy - nv) // 2
    ex, ey = (ux + nu) // 2, (uy + nv) // 2
    ix = [mirror_index(x, bx, ex, nu) for x in range(ux)]
    iy = [mirror_index(y, by, ey, nv) for y in range(uy)]
    return np.asarray(img, np.float64)[np.ix_(iy, ix)]


def ramp_filter(img, taps=None):
    """The closed form above, float64."""
    k = ramlak_taps() if taps is None else np.asarray(taps)
    nv, nu = img.shape
    ux, uy = padded_size(nu, nv, len(k))
    P = mirror_pad(img, ux, uy)
    P = np.roll(P, -1, axis=0)                     # row y reads padded row y + 1
    half = len(k) // 2
    C = np.zeros_like(P)
    for t in range(len(k)):
        if k[t] != 0.0:
            C += float(k[t]) * np.roll(P, t - half, axis=1)   # P[x - (t - half)]
    cx, cy = ux // 2 - nu // 2, uy // 2 - nv // 2
    return C[cy:cy + nv, cx:cx + nu]


def wbp(vol: "O.Volume", base: "O.Scanner", tilts, projections, lam=1.0, samples=1, xtilt=0.0, volume=None):
    """WBP::run on the CPU: float32 volume [z][y][x] (C oracle ray lengths and
    back projections, the filter from ramp_filter rounded to float)."""
    n = len(tilts)
    if volume is None:
        volume = np.zeros((vol.res[2], vol.res[1], vol.res[0]), np.float32)
    for i in range(n):
        g = O.rotate(base, float(tilts[i]), float(xtilt))
        L = O.ray_length(vol, g)
        f = ramp_filter(projections[i]).astype(np.float32)
        O.back_project(vol, volume, g, f, L, lam=lam, block_size=n, samples=samples)
    return volume
