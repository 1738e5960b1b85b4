"""Seeded synthetic tomograms for benchmarks and for training the proposal
network: Gaussian-blob phantoms pasted at spaced positions with uniform random
orientations (rotated by the library's own tom_rotate kernel), optionally
missing-wedge filtered, scaled to a signal-to-noise ratio and overlaid with
N(0, 1) noise. Volumes are torch tensors on the GPU, (z, y, x) = x fastest.
"""
from __future__ import annotations

import math

import numpy as np


def phantom(rng, n, k, sigma, R):
    """Sum of k isotropic Gaussian blobs (sigma) at random centres within
    radius R of the box centre floor(n/2), amplitudes U(0.5, 1); (x, y, z)."""
    x = np.arange(n, dtype=np.float64) - n // 2
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    T = np.zeros((n, n, n))
    for _ in range(k):
        while True:
            c = rng.uniform(-R, R, 3)
            if np.linalg.norm(c) <= R:
                break
        T += rng.uniform(0.5, 1.0) * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / (2 * sigma ** 2))
    return T.astype(np.float32)


def place(rng, N, n, P, spacing):
    """P box centres (x, y, z) with the box inside the volume and pairwise
    distance >= spacing (rejection sampling)."""
    pos = []
    lo = n // 2
    tries = 0
    while len(pos) < P:
        tries += 1
        if tries > 400000:
            raise RuntimeError("could not place particles")
        c = np.array([rng.integers(lo, N[d] - n + lo + 1) for d in range(3)])
        if all(np.linalg.norm(c - q) >= spacing for q in pos):
            pos.append(c)
    return np.asarray(pos)


def random_euler(rng, P):
    """Uniform random orientations as [phi psi theta] degrees."""
    return np.stack([rng.uniform(0, 360, P), rng.uniform(0, 360, P),
                     np.degrees(np.arccos(rng.uniform(-1, 1, P)))], 1).astype(np.float32)


def wedge_filter(S, tilt_max_deg, torch):
    """Missing wedge of a single-axis tilt series about Y with tilt range
    +-tilt_max on a (z, y, x) volume: zero the Fourier components with
    |k_z| > tan(tilt_max) |k_x| (a missing wedge of semi-angle
    90 - tilt_max around k_z)."""
    Z, Y, X = S.shape
    kz = torch.fft.fftfreq(Z, device=S.device).view(Z, 1, 1) * Z
    kx = torch.fft.rfftfreq(X, device=S.device).view(1, 1, -1) * X
    keep = (kz.abs() <= math.tan(math.radians(tilt_max_deg)) * kx.abs()) | ((kz == 0) & (kx == 0))
    F = torch.fft.rfftn(S)
    F = F * keep.to(F.dtype)
    return torch.fft.irfftn(F, s=S.shape)


def make_tomogram(N, templates, P, spacing, snr, rng, noise_seed, api, torch, dev, tilt_max=None):
    """Tomogram of P particles drawn from `templates` (list of (n, n, n)
    float32 (x, y, z) arrays of one size n; the class of each particle is
    drawn uniformly when there is more than one). Returns (V, pos (P,3) x y z,
    euler (P,3), cls (P,)). The signal is scaled so var(signal) / var(noise)
    = snr before the unit-variance noise (seeded by noise_seed) is added."""
    n = templates[0].shape[0]
    pos = place(rng, N, n, P, spacing)
    eul = random_euler(rng, P)
    cls = rng.integers(0, len(templates), P) if len(templates) > 1 else np.zeros(P, np.int64)
    S = torch.zeros((N[2], N[1], N[0]), dtype=torch.float32, device=dev)
    h = n // 2
    for c, T in enumerate(templates):
        sel = np.nonzero(cls == c)[0]
        if sel.size == 0:
            continue
        Td = torch.from_numpy(np.ascontiguousarray(T.transpose(2, 1, 0))).to(dev)
        for s0 in range(0, sel.size, 256):
            part = sel[s0:s0 + 256]
            rot = api.tom_rotate_batch(Td, eul[part])  # (k, n, n, n) as (z, y, x)
            for j, p in enumerate(part):
                x, y, z = (int(v) - h for v in pos[p])
                S[z:z + n, y:y + n, x:x + n] += rot[j]
            del rot
    if tilt_max is not None:
        S = wedge_filter(S, tilt_max, torch)
    S *= math.sqrt(snr / float(S.double().var()))
    g = torch.Generator(device=dev)
    g.manual_seed(int(noise_seed))
    V = S + torch.randn(S.shape, generator=g, device=dev, dtype=torch.float32)
    return V, pos, eul, cls
