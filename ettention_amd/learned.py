"""Learned correction of early SART iterates (no reference counterpart).

A 3-level 3-D U-Net (16/32/64 channels, 3x3x3 convolutions, bf16 on the
matrix cores, libetn.so's own forward and backward kernels) is trained on
pairs (SART iterate k0, SART iterate k1) that the reconstruction itself
produces, and predicts iterate k1 from iterate k0 as x + delta(x).

Plain SART stays the judge: `LearnedSart.run` returns the corrected volume
only if its projection residual ||Ax - b|| is below what plain SART reaches
in the same wall-clock time, and plain SART's volume otherwise.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import api as A
from ._lib import call, lib


def _torch():
    import torch
    return torch


def parameter_count():
    n = C.c_size_t()
    call("etn_unet_parameter_count", C.byref(n))
    return n.value


# (name, input channels, output channels) in parameter order; c4a / c5a read
# concat(upsample(lower level), skip)
LAYERS = [("c1a", 1, 16), ("c1b", 16, 16), ("c2a", 16, 32), ("c2b", 32, 32), ("c3a", 32, 64),
          ("c3b", 64, 64), ("c4a", 96, 32), ("c4b", 32, 32), ("c5a", 48, 16), ("c5b", 16, 16)]


def init_parameters(seed):
    """He-normal convolution weights (numpy PCG64 stream of `seed`), zero
    biases, a small head so that the initial correction is close to zero."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = np.zeros(parameter_count(), np.float32)
    at = 0
    for _, ci, co in LAYERS:
        nw = co * 27 * ci
        out[at:at + nw] = rng.normal(0.0, np.sqrt(2.0 / (27 * ci)), nw)
        at += nw + co
    out[at:at + 16] = rng.normal(0.0, 1e-3, 16)
    return out


class CorrectionNet:
    """etn_unet_*: the network at a fixed working size (volumes [nz][ny][nx])."""

    def __init__(self, shape_zyx, params=None, seed=5000, stream=None):
        nz, ny, nx = shape_zyx
        self.shape = (nz, ny, nx)
        h = C.c_void_p()
        call("etn_unet_create", nx, ny, nz, C.c_void_p(stream) if stream else None, C.byref(h))
        self._h = h
        self.set_parameters(init_parameters(seed) if params is None else params)

    def close(self):
        if getattr(self, "_h", None):
            lib().etn_unet_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_parameters(self, p):
        p = np.ascontiguousarray(p, np.float32)
        call("etn_unet_set_parameters", self._h, p.ctypes.data_as(C.c_void_p), p.size)

    def parameters(self):
        p = np.zeros(parameter_count(), np.float32)
        call("etn_unet_get_parameters", self._h, p.ctypes.data_as(C.c_void_p), p.size)
        return p

    def predict(self, x, scale, out=None):
        torch = _torch()
        out = torch.empty_like(x) if out is None else out
        call("etn_unet_predict", self._h, C.c_void_p(x.data_ptr()), float(scale), C.c_void_p(out.data_ptr()))
        return out

    def gradient(self, x, target, scale):
        g = np.zeros(parameter_count(), np.float32)
        loss = C.c_double()
        call("etn_unet_gradient", self._h, C.c_void_p(x.data_ptr()), C.c_void_p(target.data_ptr()), float(scale),
             g.ctypes.data_as(C.c_void_p), C.byref(loss))
        return loss.value, g

    def train_step(self, x, target, scale, lr):
        loss = C.c_double()
        call("etn_unet_train_step", self._h, C.c_void_p(x.data_ptr()), C.c_void_p(target.data_ptr()), float(scale),
             float(lr), C.byref(loss))
        return loss.value

    def time_forward(self, x, scale):
        ms, flops = C.c_double(), C.c_double()
        call("etn_unet_time_forward", self._h, C.c_void_p(x.data_ptr()), float(scale), C.byref(ms), C.byref(flops))
        return ms.value, flops.value


def sart_pairs(members, k0, k1, flags=0):
    """Runs the batch of reconstructions (keyword dicts as for
    BlockIterativeReconstructionOperator) and returns float32 device volumes
    (iterate k0, iterate k1) per member."""
    b = A.BatchReconstruction(members, flags=flags)
    b.run_iterations(k0)
    early = [b.get_reconstructed_volume(i, dtype=_torch().float32) for i in range(len(members))]
    b.run_iterations(k1 - k0)
    late = [b.get_reconstructed_volume(i, dtype=_torch().float32) for i in range(len(members))]
    b.close()
    return early, late


def train(net, pairs, scale, steps, lr=1e-3, log=None):
    """Adam over the (early, late) pairs in a fixed cyclic order; returns the loss per step."""
    losses = []
    for s in range(steps):
        x, t = pairs[s % len(pairs)]
        losses.append(net.train_step(x, t, scale, lr))
        if log and (s % 50 == 0 or s == steps - 1):
            log("step %d loss %.6e" % (s, losses[-1]))
    return losses


class LearnedSart:
    """SART with a learned jump from iterate k0: run k0 iterations, apply the
    network, and keep the result only if ||Ax - b|| is below what plain SART
    reaches in the same wall-clock time (the reconstruction's own FP64 norm)."""

    def __init__(self, net, scale, k0=3):
        self.net, self.scale, self.k0 = net, scale, k0

    def run(self, make_reconstruction, plain_iteration_ms=None):
        torch = _torch()
        r = make_reconstruction()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r.run_iterations(self.k0)
        x = r.get_reconstructed_volume(dtype=torch.float32)
        y = self.net.predict(x, self.scale)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        r.set_volume(y.to(_voxel_dtype(r)))
        corrected = r.residual_norm()
        # plain SART given the same time: whole iterations that fit
        p = make_reconstruction()
        if plain_iteration_ms is None:
            p.run_iterations(1)
            p.synchronize()
            t1 = time.perf_counter()
            p.run_iterations(1)
            p.synchronize()
            plain_iteration_ms = 1e3 * (time.perf_counter() - t1)
            p = make_reconstruction()
        k = max(self.k0, int(elapsed * 1e3 / plain_iteration_ms))
        p.run_iterations(k)
        plain = p.residual_norm()
        accepted = corrected < plain
        return {"accepted": accepted, "corrected_norm": corrected, "plain_norm": plain, "plain_iterations": k,
                "elapsed_ms": 1e3 * elapsed, "volume": y if accepted else p.get_reconstructed_volume(torch.float32)}


def _voxel_dtype(r):
    torch = _torch()
    from ._lib import ETN_VOXEL_BFLOAT16, ETN_VOXEL_FLOAT16
    return {ETN_VOXEL_FLOAT16: torch.float16, ETN_VOXEL_BFLOAT16: torch.bfloat16}.get(r.voxel_type, torch.float32)
