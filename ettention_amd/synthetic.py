"""Seeded synthetic workloads for benchmarks and config-level tests.

Phantom: additive random ellipsoids (centres uniform in the inner 80 % of the
box, semi-axes U[0.05, 0.35] x edge, uniformly random orientation, densities
U[0.1, 1]) plus optional radius-1.5 "beads" of fixed density.  All random
numbers come from a 64-bit Mersenne twister with explicit transforms
(u = (x >> 11) * 2^-53; Shoemake's uniform quaternion), so the phantom is
identical on every platform.  Voxelisation runs on the GPU with PyTorch when a
device is given (plain tensor arithmetic on the inputs, not the hot path).
"""
from __future__ import annotations

import math

import numpy as np


class MT19937_64:
    """64-bit Mersenne twister (Matsumoto & Nishimura, MT19937-64 parameters)."""

    def __init__(self, seed):
        self.mt = [0] * 312
        self.mt[0] = seed & 0xFFFFFFFFFFFFFFFF
        for i in range(1, 312):
            x = self.mt[i - 1]
            self.mt[i] = (6364136223846793005 * (x ^ (x >> 62)) + i) & 0xFFFFFFFFFFFFFFFF
        self.i = 312

    def next(self):
        if self.i >= 312:
            um, lm = 0xFFFFFFFF80000000, 0x7FFFFFFF
            for k in range(312):
                y = (self.mt[k] & um) | (self.mt[(k + 1) % 312] & lm)
                v = self.mt[(k + 156) % 312] ^ (y >> 1)
                if y & 1:
                    v ^= 0xB5026F5AA96619E9
                self.mt[k] = v
            self.i = 0
        y = self.mt[self.i]
        self.i += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & 0xFFFFFFFFFFFFFFFF

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / 9007199254740992.0))


def _quaternion_matrix(u1, u2, u3):
    a = math.sqrt(1 - u1)
    b = math.sqrt(u1)
    x, y = a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2)
    z, w = b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def phantom_parameters(shape_zyx, seed, ellipsoids=20, beads=0, bead_density=5.0):
    """Ellipsoid and bead parameters in voxel coordinates (x, y, z)."""
    nz, ny, nx = shape_zyx
    ext = np.array([nx, ny, nz], np.float64)
    rng = MT19937_64(seed)
    ells = []
    for _ in range(ellipsoids):
        c = np.array([rng.uniform(0.1, 0.9) for _ in range(3)]) * ext
        ax = np.array([rng.uniform(0.05, 0.35) for _ in range(3)]) * ext
        R = _quaternion_matrix(rng.uniform(), rng.uniform(), rng.uniform())
        d = rng.uniform(0.1, 1.0)
        ells.append((c, ax, R, d))
    bead_list = []
    for _ in range(beads):
        c = np.array([rng.uniform(0.1, 0.9) for _ in range(3)]) * ext
        bead_list.append((c, 1.5, bead_density))
    return ells, bead_list


def ellipsoid_phantom(shape_zyx, seed, ellipsoids=20, beads=0, bead_density=5.0, device=None, dtype=None):
    """Voxelised phantom [z][y][x] (float32).  numpy array when device is None,
    else a torch tensor on that device."""
    ells, bead_list = phantom_parameters(shape_zyx, seed, ellipsoids, beads, bead_density)
    nz, ny, nx = shape_zyx
    if device is None:
        vol = np.zeros(shape_zyx, np.float32)
        zz = np.arange(nz) + 0.5
        yy = np.arange(ny) + 0.5
        xx = np.arange(nx) + 0.5
        for c, ax, R, d in ells:
            r = float(ax.max())
            lo = np.maximum(np.floor(c - r).astype(int), 0)
            hi = np.minimum(np.ceil(c + r).astype(int) + 1, [nx, ny, nz])
            Z, Y, X = np.meshgrid(zz[lo[2]:hi[2]], yy[lo[1]:hi[1]], xx[lo[0]:hi[0]], indexing="ij")
            p = np.stack([X - c[0], Y - c[1], Z - c[2]], -1) @ R
            inside = ((p / ax) ** 2).sum(-1) <= 1.0
            vol[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] += np.float32(d) * inside
        for c, rad, dens in bead_list:
            lo = np.maximum(np.floor(c - rad).astype(int), 0)
            hi = np.minimum(np.ceil(c + rad).astype(int) + 1, [nx, ny, nz])
            Z, Y, X = np.meshgrid(zz[lo[2]:hi[2]], yy[lo[1]:hi[1]], xx[lo[0]:hi[0]], indexing="ij")
            inside = (X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2 <= rad * rad
            sub = vol[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
            sub[inside] = dens
        return vol
    import torch
    vol = torch.zeros(shape_zyx, dtype=torch.float32, device=device)
    for c, ax, R, d in ells:
        r = float(ax.max())
        lo = np.maximum(np.floor(c - r).astype(int), 0)
        hi = np.minimum(np.ceil(c + r).astype(int) + 1, [nx, ny, nz])
        z = torch.arange(lo[2], hi[2], device=device, dtype=torch.float64) + 0.5 - c[2]
        y = torch.arange(lo[1], hi[1], device=device, dtype=torch.float64) + 0.5 - c[1]
        x = torch.arange(lo[0], hi[0], device=device, dtype=torch.float64) + 0.5 - c[0]
        Rt = torch.tensor(R, device=device, dtype=torch.float64)
        acc = torch.zeros((len(z), len(y), len(x)), dtype=torch.float64, device=device)
        for k in range(3):   # (p @ R)_k / a_k = (x R0k + y R1k + z R2k) / a_k
            t = (x[None, None, :] * Rt[0, k] + y[None, :, None] * Rt[1, k] + z[:, None, None] * Rt[2, k]) / ax[k]
            acc += t * t
        vol[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] += (acc <= 1.0).to(torch.float32) * np.float32(d)
    for c, rad, dens in bead_list:
        lo = np.maximum(np.floor(c - rad).astype(int), 0)
        hi = np.minimum(np.ceil(c + rad).astype(int) + 1, [nx, ny, nz])
        z = torch.arange(lo[2], hi[2], device=device, dtype=torch.float64) + 0.5 - c[2]
        y = torch.arange(lo[1], hi[1], device=device, dtype=torch.float64) + 0.5 - c[1]
        x = torch.arange(lo[0], hi[0], device=device, dtype=torch.float64) + 0.5 - c[0]
        inside = (x[None, None, :] ** 2 + y[None, :, None] ** 2 + z[:, None, None] ** 2) <= rad * rad
        sub = vol[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
        sub[inside] = dens
    if dtype is not None:
        vol = vol.to(dtype)
    return vol


def tilt_series(max_angle, step):
    """Angles -max..+max inclusive in `step` degree steps."""
    n = int(round(2 * max_angle / step)) + 1
    return np.array([-max_angle + i * step for i in range(n)], np.float32)


BATCH_EDGES = (64, 96, 128, 192, 256)


def batch_composition(count, seed, edges=BATCH_EDGES, min_tilts=41, max_tilts=121):
    """(edge, tilt count) of each member of a batch: the cube edge uniform over
    `edges`, the tilt count uniform over [min_tilts, max_tilts], both drawn from
    MT19937-64 uniforms u = (x >> 11) * 2^-53."""
    g = MT19937_64(seed)
    out = []
    for _ in range(count):
        e = edges[min(int(g.uniform() * len(edges)), len(edges) - 1)]
        n = min_tilts + min(int(g.uniform() * (max_tilts - min_tilts + 1)), max_tilts - min_tilts)
        out.append((e, n))
    return out


def add_gaussian_noise(projections, sigma_rel, seed):
    """b + N(0, (sigma_rel * max b)^2); numpy or torch input of the same kind out."""
    if isinstance(projections, np.ndarray):
        rng = np.random.default_rng(seed)
        s = sigma_rel * float(projections.max())
        return (projections + rng.normal(0.0, s, projections.shape)).astype(np.float32)
    import torch
    g = torch.Generator(device=projections.device).manual_seed(seed)
    s = sigma_rel * float(projections.max())
    return projections + s * torch.randn(projections.shape, generator=g, device=projections.device)


def add_poisson_noise(projections, counts, seed, peak_attenuation=2.0):
    """Transmission noise: line integrals are scaled so the largest is
    `peak_attenuation`, counts ~ Poisson(counts * exp(-b)), b' = -ln(max(n,1)/counts),
    then scaled back.  numpy in, numpy out; a torch tensor is noised on its own
    device with a torch generator seeded `seed`, and returned as float32."""
    scale = peak_attenuation / max(float(projections.max()), 1e-12)
    if not isinstance(projections, np.ndarray):
        import torch
        g = torch.Generator(device=projections.device).manual_seed(seed)
        rate = counts * torch.exp(-scale * projections.double())
        n = torch.poisson(rate, generator=g)
        return (-torch.log(torch.clamp(n, min=1.0) / counts) / scale).float()
    rng = np.random.default_rng(seed)
    n = rng.poisson(counts * np.exp(-scale * projections.astype(np.float64)))
    return (-np.log(np.maximum(n, 1) / counts) / scale).astype(np.float32)
