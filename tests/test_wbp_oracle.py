"""The weighted-back-projection oracle (oracle/wbp_oracle.py) on the CPU.

Pins:
* the closed-form circular convolution equals the reference's Fourier-space
  pipeline step by step (zero-padded centred filter, origin moved to the
  corner, 2-D FFT product, centred crop; ConvolutionFourierSpaceImplementation.cpp,
  PaddingForFFT.cl, MoveOriginToCenter.cl, CroppingForFFT.cl), for power-of-two,
  even, odd and single-row images;
* a delta image gives the taps back, centred, one row up;
* the taps are GenerateRamLakKernel.cl's closed form and sum to
  1/4 - (2 / pi^2) sum_{k<32} 1 / (2k + 1)^2;
* on the reference's own WBP test data, the reference's own bound
  (WBPPluginTest.cpp:35-41, RMS < 0.05 to ground_truth.mrc) -- a weak bound:
  the data's RMS is 0.013, so it mostly checks that nothing blows up.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from oracle import wbp_oracle as W


def fft_pipeline(img, k):
    nv, nu = img.shape
    ux, uy = W.padded_size(nu, nv, len(k))
    # computeZeroPaddedCopyForFourierSpace of the taps x 1 filter image
    F = np.zeros((uy, ux))
    bx, by = (ux - len(k)) // 2, (uy - 1) // 2
    F[by, bx:bx + len(k)] = k
    # computeShiftedToCenterOrigin: result[y][x] = src[(y + uy/2) % uy][(x + ux/2) % ux]
    ys = (np.arange(uy) + uy // 2) % uy
    xs = (np.arange(ux) + ux // 2) % ux
    Fs = F[np.ix_(ys, xs)]
    P = W.mirror_pad(img, ux, uy)
    C = np.fft.ifft2(np.fft.fft2(Fs) * np.fft.fft2(P)).real
    cx, cy = ux // 2 - nu // 2, uy // 2 - nv // 2          # computeCropFromFourierSpace
    return C[cy:cy + nv, cx:cx + nu]


@pytest.mark.parametrize("shape", [(128, 128), (64, 100), (37, 101), (1, 200), (96, 256), (5, 130)])
def test_closed_form_equals_fourier_pipeline(shape):
    img = np.random.default_rng(sum(shape)).normal(0, 1, shape)
    k = W.ramlak_taps().astype(np.float64)
    np.testing.assert_allclose(W.ramp_filter(img, k), fft_pipeline(img, k), atol=1e-11)


def test_delta_gives_the_taps_one_row_up():
    nu = nv = 128
    k = W.ramlak_taps()
    img = np.zeros((nv, nu))
    img[41, 70] = 1.0
    out = W.ramp_filter(img)
    # output row 40 holds the taps centred on column 70; every other row is zero
    assert np.count_nonzero(out[np.arange(nv) != 40]) == 0
    expect = np.zeros(nu)
    for t in range(128):
        expect[(70 + t - 64) % nu] = k[t]
    np.testing.assert_allclose(out[40], expect, rtol=1e-12, atol=0)


def test_ramlak_taps_closed_form():
    k = W.ramlak_taps()
    assert k[64] == np.float32(0.25) and np.all(k[::2][np.arange(64) != 32] == 0)
    for p in (1, 3, 63, -1, -63):
        np.testing.assert_allclose(k[64 + p], -1.0 / (p * p * np.pi ** 2), rtol=2e-7)
    s = 0.25 - 2.0 / np.pi ** 2 * sum(1.0 / (2 * j + 1) ** 2 for j in range(32))
    np.testing.assert_allclose(k.astype(np.float64).sum(), s, rtol=0, atol=1e-7)   # float taps
    # a constant row is scaled by the sum of the taps (no padding at nu = 128)
    img = np.full((4, 128), 3.0)
    np.testing.assert_allclose(W.ramp_filter(img), 3.0 * k.astype(np.float64).sum(), rtol=1e-9)


REF = "/root/reference/ettention/plugins/WBPPlugin/test/data/wbp_shepp_logan"


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference test data not present")
def test_reference_wbp_bound():
    raw = open(os.path.join(REF, "empiar_10009_128.mrc"), "rb").read()
    projs = np.frombuffer(raw[1024:], np.uint8).reshape(51, 128, 128).astype(np.float32)   # mode 0
    tilts = np.array(open(os.path.join(REF, "empiar_10009_128.tlt")).read().split(), np.float32)
    gt = np.frombuffer(open(os.path.join(REF, "ground_truth.mrc"), "rb").read()[1024:], np.float32).reshape(20, 128, 128)
    v = O.Volume.make((128, 128, 20))
    vol = W.wbp(v, O.mrc_base_scanner(128, 128), tilts, projs, lam=1.0)
    d = vol.astype(np.float64) - gt        # both x-y-z files: the reference compares file slices
    assert np.sqrt((d ** 2).mean(axis=(1, 2))).max() < 0.05
