"""Learned correction of early SART iterates (no reference counterpart).

A 3-level 3-D U-Net (16/32/64 channels, 3x3x3 convolutions, bf16 on the
matrix cores, libetn.so's own forward and backward kernels) is trained on
samples that the reconstruction itself produces: the SART iterate x_k0, its
last update d = x_k0 - x_(k0-1), and a late iterate as the target.  It
predicts the late iterate as x_k0 + delta(x_k0, d).

SART as the reference implements it semi-converges: ||Ax - b|| falls for the
first iterations and then grows (the voxel-driven bilinear back projection,
with its half-pixel offset, is not the forward projector's adjoint), on
noiseless data as well.  The late iterate is therefore the one with the
lowest residual within k_max iterations, not iterate k_max itself.

Plain SART stays the arbiter: a jump is accepted only if its projection
residual is below what plain SART reaches in the same wall-clock time.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import api as A
from ._lib import call, lib

# the network sees (x scale, d scale UPDATE_GAIN): updates are about an
# order of magnitude smaller than the iterate
UPDATE_GAIN = 16.0


def _torch():
    import torch
    return torch


def parameter_count():
    n = C.c_size_t()
    call("etn_unet_parameter_count", C.byref(n))
    return n.value


# (name, input channels, output channels) in parameter order; c1a reads (x, d),
# c4a / c5a read concat(upsample(lower level), skip)
LAYERS = [("c1a", 2, 16), ("c1b", 16, 16), ("c2a", 16, 32), ("c2b", 32, 32), ("c3a", 32, 64),
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


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class CorrectionNet:
    """etn_unet_*: the network at a fixed working size (volumes [nz][ny][nx],
    float32 device tensors)."""

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

    def predict(self, x, d, scale, update_scale=None, out=None):
        """x + delta(x scale, d update_scale) / scale; d may be None (zeros)."""
        torch = _torch()
        us = scale * UPDATE_GAIN if update_scale is None else update_scale
        out = torch.empty_like(x) if out is None else out
        call("etn_unet_predict", self._h, _p(x), _p(d), float(scale), float(us), _p(out))
        return out

    def gradient(self, x, d, target, scale, update_scale=None):
        us = scale * UPDATE_GAIN if update_scale is None else update_scale
        g = np.zeros(parameter_count(), np.float32)
        loss = C.c_double()
        call("etn_unet_gradient", self._h, _p(x), _p(d), _p(target), float(scale), float(us),
             g.ctypes.data_as(C.c_void_p), C.byref(loss))
        return loss.value, g

    def train_step(self, x, d, target, scale, lr, update_scale=None):
        us = scale * UPDATE_GAIN if update_scale is None else update_scale
        loss = C.c_double()
        call("etn_unet_train_step", self._h, _p(x), _p(d), _p(target), float(scale), float(us), float(lr),
             C.byref(loss))
        return loss.value

    def time_forward(self, x, scale):
        ms, flops = C.c_double(), C.c_double()
        call("etn_unet_time_forward", self._h, _p(x), float(scale), C.byref(ms), C.byref(flops))
        return ms.value, flops.value


def make_case(shape_zyx, ntilts, seed, counts=1e4, max_angle=60.0, device="cuda"):
    """A seeded ellipsoid phantom projected by the library's forward projector
    on the MRC geometry (detector nx x ny, tilts evenly over +-max_angle) with
    Poisson transmission noise at `counts` per pixel.  Keyword arguments of
    BlockIterativeReconstructionOperator, plus the phantom."""
    import ettention_amd as E
    from . import synthetic as S
    torch = _torch()
    nz, ny, nx = shape_zyx
    vg = E.VolumeGeometry.make((nx, ny, nz))
    base = E.ScannerGeometry.mrc_base(nx, ny)
    tilts = np.linspace(-max_angle, max_angle, ntilts).astype(np.float32)
    ph = S.ellipsoid_phantom((nz, ny, nx), seed=seed, device=torch.device(device))
    projs = torch.zeros((ntilts, ny, nx), dtype=torch.float32, device=device)
    scratch = torch.zeros((ny, nx), dtype=torch.float32, device=device)
    for i, t in enumerate(tilts):
        scratch.zero_()
        A.forward_project(vg, ph, A.rotate(base, float(t)), projs[i], scratch)
    projs = S.add_poisson_noise(projs, counts, seed + 100000)
    torch.cuda.synchronize()
    return dict(volume=vg, base_geometry=base, tilt_angles=tilts, projections=projs), ph


def sart_iterates(case, iterations, voxel_type="float", lam=0.3, order="random"):
    """float32 copies of the SART iterates whose numbers are in `iterations` (ascending)."""
    r = A.BlockIterativeReconstructionOperator(algorithm="sart", lam=lam, projection_iteration=order,
                                               voxel_type=voxel_type, **case)
    out, done = [], 0
    for k in iterations:
        r.run_iterations(k - done)
        done = k
        out.append(r.get_reconstructed_volume(dtype=_torch().float32))
    r.close()
    return out


def sart_best_iterate(case, k_max=50, voxel_type="float", lam=0.3, order="random", patience=4):
    """Runs up to k_max SART iterations, stopping `patience` iterations after
    the last new minimum of ||Ax - b||, and returns (float32 copy of the
    minimum-residual iterate, its iteration number, the residual after every
    iteration run)."""
    torch = _torch()
    r = A.BlockIterativeReconstructionOperator(algorithm="sart", lam=lam, projection_iteration=order,
                                               voxel_type=voxel_type, **case)
    norms, best, best_k = [], None, 0
    for k in range(1, k_max + 1):
        r.run_iterations(1)
        norms.append(r.residual_norm())
        if best is None or norms[-1] < norms[best_k - 1]:
            best, best_k = r.get_reconstructed_volume(dtype=torch.float32), k
        elif k - best_k >= patience:
            break
    r.close()
    return best, best_k, norms


def training_sample(case, k0=3, k_max=50, lam=0.3, order="random"):
    """(x, d, target): bf16-storage SART iterate k0, its last update and the
    FP32 SART minimum-residual iterate; plus the target's iteration number."""
    prev, x = sart_iterates(case, [k0 - 1, k0], voxel_type="bfloat16", lam=lam, order=order)
    target, k_best, _ = sart_best_iterate(case, k_max, lam=lam, order=order)
    return (x, x - prev, target), k_best


def train(net, samples, scale, steps, lr=1e-3, log=None):
    """Adam over whole (x, d, target) samples in a fixed cyclic order; the loss per step."""
    losses = []
    for s in range(steps):
        x, d, t = samples[s % len(samples)]
        losses.append(net.train_step(x, d, t, scale, lr))
        if log and (s % 100 == 0 or s == steps - 1):
            log("train step %d: loss %.4e" % (s, losses[-1]))
    return losses


def train_on_crops(net, samples, crop, steps, scale, lr=1e-3, lr_min=5e-5, seed=0, log=None):
    """Adam steps on random crop^3 windows (offsets multiples of 4) of the
    (x, d, target) samples, cycling through them, with the learning rate on a
    cosine from lr to lr_min; net's working size is crop^3."""
    rng = np.random.default_rng(seed)
    losses = []
    for s in range(steps):
        x, d, t = samples[s % len(samples)]
        o = [4 * int(rng.integers(0, (n - crop) // 4 + 1)) for n in x.shape]
        sl = tuple(slice(a, a + crop) for a in o)
        rate = lr_min + 0.5 * (lr - lr_min) * (1.0 + np.cos(np.pi * s / max(steps - 1, 1)))
        losses.append(net.train_step(x[sl].contiguous(), d[sl].contiguous(), t[sl].contiguous(), scale, rate))
        if log and (s % 100 == 0 or s == steps - 1):
            log("train step %d: loss %.4e" % (s, losses[-1]))
    return losses


def time_to_target(case, net, scale, k_max=50, k0=3, lam=0.3, order="random", patience=4, log=None):
    """Wall time until ||Ax - b|| (the reconstruction's FP64 norm, evaluated
    outside the timed regions) reaches the lowest value plain FP32 SART attains
    within k_max iterations: plain FP32 SART against bf16-storage SART that
    jumps with the network after k0 iterations.  Times are sums of synchronised
    iteration (and network) times; both runs stop `patience` iterations after
    their last new minimum.  The jump is `accepted` when its residual is below
    plain SART's after the iterations that fit in the jump's wall-clock time."""
    torch = _torch()

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    plain = A.BlockIterativeReconstructionOperator(algorithm="sart", lam=lam, projection_iteration=order, **case)
    plain.run_iterations(1)          # autotune and graph capture outside the timed region
    plain.set_volume(torch.zeros(plain.shape, dtype=torch.float32, device="cuda"))
    p_times, p_norms, k_best = [], [], 0
    for k in range(1, k_max + 1):
        p_times.append(timed(lambda: plain.run_iterations(1)))
        p_norms.append(plain.residual_norm())
        if k_best == 0 or p_norms[-1] < p_norms[k_best - 1]:
            k_best = k
        elif k - k_best >= patience:
            break
    target = p_norms[k_best - 1]
    t_plain = sum(p_times[:k_best])

    lrn = A.BlockIterativeReconstructionOperator(algorithm="sart", lam=lam, projection_iteration=order,
                                                 voxel_type="bfloat16", **case)
    lrn.run_iterations(1)
    lrn.set_volume(torch.zeros(lrn.shape, dtype=torch.bfloat16, device="cuda"))

    jumped = []

    def jump():
        lrn.run_iterations(k0 - 1)
        prev = lrn.get_reconstructed_volume(dtype=torch.float32)
        lrn.run_iterations(1)
        x = lrn.get_reconstructed_volume(dtype=torch.float32)
        y = net.predict(x, x - prev, scale)
        lrn.set_volume(y.to(torch.bfloat16))
        jumped.append(y)

    t_jump = timed(jump)
    jump_norm = lrn.residual_norm()

    def polish(op, t0):
        """Iterates op from the jumped volume until the target, k_max or patience."""
        t, norm, extra, best_norm, since_best = t0, jump_norm, 0, jump_norm, 0
        while norm > target and k0 + extra < k_max and since_best < patience:
            t += timed(lambda: op.run_iterations(1))
            extra += 1
            norm = op.residual_norm()
            if norm < best_norm:
                best_norm, since_best = norm, 0
            else:
                since_best += 1
        return t, norm, extra

    t_learned, norm, extra = polish(lrn, t_jump)
    # the same jump, then FP32 SART instead of bf16 (the conversion is timed)
    f32 = A.BlockIterativeReconstructionOperator(algorithm="sart", lam=lam, projection_iteration=order, **case)
    f32.run_iterations(1)
    t_conv = timed(lambda: f32.set_volume(jumped[0]))
    t_f32, norm_f32, extra_f32 = polish(f32, t_jump + t_conv)
    del f32
    # plain bf16 SART for reference (untimed): its own lowest residual
    lrn.set_volume(torch.zeros(lrn.shape, dtype=torch.bfloat16, device="cuda"))
    b_norms = []
    for k in range(1, min(k_max, k_best + patience) + 1):
        lrn.run_iterations(1)
        b_norms.append(lrn.residual_norm())
    t_iter = sum(p_times) / len(p_times)
    k_equal = min(max(int(t_jump / t_iter), 1), len(p_norms))
    accepted = jump_norm < p_norms[k_equal - 1]
    reached = norm <= target
    res = {"target_norm": target, "plain_best_iteration": k_best, "plain_s": t_plain,
           "plain_norm_at_k0": p_norms[min(k0, len(p_norms)) - 1],
           "plain_iteration_ms": 1e3 * t_iter, "jump_iterations": k0, "jump_s": t_jump, "jump_norm": jump_norm,
           "plain_norm_at_jump_time": p_norms[k_equal - 1], "plain_iterations_at_jump_time": k_equal,
           "accepted": bool(accepted), "learned_extra_iterations": extra, "learned_s": t_learned,
           "learned_final_norm": norm, "reached": bool(reached),
           "speedup": t_plain / t_learned if reached else None, "plain_norms": p_norms,
           "f32_polish_extra_iterations": extra_f32, "f32_polish_s": t_f32, "f32_polish_final_norm": norm_f32,
           "f32_polish_reached": bool(norm_f32 <= target),
           "f32_polish_speedup": t_plain / t_f32 if norm_f32 <= target else None,
           "bf16_plain_best_norm": min(b_norms), "bf16_plain_best_iteration": 1 + b_norms.index(min(b_norms))}
    if log:
        log("plain FP32 SART: lowest residual %.6e at iteration %d, %.3f s" % (target, k_best, t_plain))
        log("bf16 SART + network: jump after %d iterations in %.3f s -> %.6e (plain at equal time, %d it: %.6e); "
            "%d more iterations, %.3f s total, target %s" % (k0, t_jump, jump_norm, k_equal, p_norms[k_equal - 1],
                                                           extra, t_learned, "reached" if reached else "missed"))
        log("same jump, then FP32 SART: %d more iterations, %.3f s total, %.6e, target %s; plain bf16 SART's "
            "lowest residual %.6e at iteration %d" % (extra_f32, t_f32, norm_f32,
                                                     "reached" if norm_f32 <= target else "missed",
                                                     min(b_norms), 1 + b_norms.index(min(b_norms))))
    return res


def learned_sart(case, net, scale, k0=3, lam=0.3, order="random", voxel_type="bfloat16"):
    """SART with the learned jump from iterate k0, plain SART as the arbiter:
    returns (volume, accepted, jump residual, plain residual), where plain SART
    gets the iterations that fit in the jump's wall-clock time and its volume
    is returned whenever the jump does not beat it."""
    torch = _torch()
    r = A.BlockIterativeReconstructionOperator(algorithm="sart", lam=lam, projection_iteration=order,
                                               voxel_type=voxel_type, **case)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r.run_iterations(k0 - 1)
    prev = r.get_reconstructed_volume(dtype=torch.float32)
    r.run_iterations(1)
    x = r.get_reconstructed_volume(dtype=torch.float32)
    y = net.predict(x, x - prev, scale)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    r.set_volume(y.to(_voxel_dtype(r)))
    jump_norm = r.residual_norm()
    p = A.BlockIterativeReconstructionOperator(algorithm="sart", lam=lam, projection_iteration=order, **case)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    k = 0
    while k == 0 or time.perf_counter() - t1 < elapsed:
        p.run_iterations(1)
        p.synchronize()
        k += 1
    plain_norm = p.residual_norm()
    accepted = jump_norm < plain_norm
    vol = y if accepted else p.get_reconstructed_volume(dtype=torch.float32)
    return vol, accepted, jump_norm, plain_norm


def _voxel_dtype(r):
    torch = _torch()
    from ._lib import ETN_VOXEL_BFLOAT16, ETN_VOXEL_FLOAT16
    return {ETN_VOXEL_FLOAT16: torch.float16, ETN_VOXEL_BFLOAT16: torch.bfloat16}.get(r.voxel_type, torch.float32)


def _main():
    """python -m ettention_amd.learned [edge [repeats]]: HIP-event time of the
    network's forward pass on an edge^3 volume of uniform noise."""
    import sys
    torch = _torch()
    edge = int(sys.argv[1]) if len(sys.argv) > 1 else 512
    repeats = int(sys.argv[2]) if len(sys.argv) > 2 else 3
    net = CorrectionNet((edge, edge, edge))
    x = torch.rand((edge, edge, edge), device="cuda")
    for i in range(repeats):
        ms, flops = net.time_forward(x, 1.0)
        print("forward %d^3 run %d: %.3f ms, %.4e flop, %.1f TFLOP/s" % (edge, i, ms, flops, flops / ms / 1e9),
              flush=True)


if __name__ == "__main__":
    _main()
