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


def train_on_crops(net, pairs, crop, steps, scale, lr=1e-3, seed=0, log=None):
    """Adam steps on random crop^3 windows (offsets multiples of 4) of the
    (early, late) pairs, cycling through the pairs; net's working size is crop^3."""
    rng = np.random.default_rng(seed)
    losses = []
    for s in range(steps):
        x, t = pairs[s % len(pairs)]
        o = [4 * int(rng.integers(0, (n - crop) // 4 + 1)) for n in x.shape]
        sl = tuple(slice(a, a + crop) for a in o)
        losses.append(net.train_step(x[sl].contiguous(), t[sl].contiguous(), scale, lr))
        if log and (s % 100 == 0 or s == steps - 1):
            log("train step %d: loss %.4e" % (s, losses[-1]))
    return losses


def sart_best_iterate(case, k_max=50, voxel_type="float", lam=0.3, order="random", patience=4):
    """SART as the reference implements it semi-converges: ||Ax - b|| falls for
    the first iterations, then grows (the voxel-driven bilinear back projection
    with its half-pixel offset is not the forward projector's adjoint).  Runs up
    to k_max iterations, stopping `patience` iterations after the last new
    minimum, and returns (float32 copy of the minimum-residual iterate, its
    iteration number, the residual after every iteration run)."""
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


def time_to_target(case, net, scale, k_max=50, k0=3, lam=0.3, order="random", patience=4, log=None):
    """Wall time until ||Ax - b|| (the reconstruction's FP64 norm, evaluated
    outside the timed regions) reaches the lowest value plain FP32 SART attains
    within k_max iterations: plain FP32 SART against bf16-storage SART that
    jumps with the network after k0 iterations.  Times are sums of synchronised
    iteration (and network) times; both runs stop `patience` iterations after
    their last new minimum."""
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
    holder = {}

    def jump():
        lrn.run_iterations(k0)
        x = lrn.get_reconstructed_volume(dtype=torch.float32)
        holder["y"] = net.predict(x, scale)
        lrn.set_volume(holder["y"].to(torch.bfloat16))

    t_jump = timed(jump)
    jump_norm = lrn.residual_norm()
    t_learned, norm, extra, best_norm, since_best = t_jump, jump_norm, 0, jump_norm, 0
    while norm > target and k0 + extra < k_max and since_best < patience:
        t_learned += timed(lambda: lrn.run_iterations(1))
        extra += 1
        norm = lrn.residual_norm()
        if norm < best_norm:
            best_norm, since_best = norm, 0
        else:
            since_best += 1
    # the acceptance rule: the jump against plain SART given the same wall-clock time
    t_iter = sum(p_times) / len(p_times)
    k_equal = min(max(int(t_jump / t_iter), 1), len(p_norms))
    accepted = jump_norm < p_norms[k_equal - 1]
    reached = norm <= target
    res = {"target_norm": target, "plain_best_iteration": k_best, "plain_s": t_plain,
           "plain_iteration_ms": 1e3 * t_iter, "jump_iterations": k0, "jump_s": t_jump, "jump_norm": jump_norm,
           "plain_norm_at_jump_time": p_norms[k_equal - 1], "plain_iterations_at_jump_time": k_equal,
           "accepted": bool(accepted), "learned_extra_iterations": extra, "learned_s": t_learned,
           "learned_final_norm": norm, "reached": bool(reached),
           "speedup": t_plain / t_learned if reached else None, "plain_norms": p_norms}
    if log:
        log("plain FP32 SART: lowest residual %.6e at iteration %d, %.3f s" % (target, k_best, t_plain))
        log("bf16 SART + network: jump after %d iterations in %.3f s -> %.6e (plain at equal time, %d it: %.6e); "
            "%d more iterations, %.3f s total, target %s" % (k0, t_jump, jump_norm, k_equal, p_norms[k_equal - 1],
                                                           extra, t_learned, "reached" if reached else "missed"))
    return res


def _voxel_dtype(r):
    torch = _torch()
    from ._lib import ETN_VOXEL_BFLOAT16, ETN_VOXEL_FLOAT16
    return {ETN_VOXEL_FLOAT16: torch.float16, ETN_VOXEL_BFLOAT16: torch.bfloat16}.get(r.voxel_type, torch.float32)
