"""Seeded synthetic workloads of the benchmark configurations (the tomograms,
templates, masks and subtomogram batches that bench.py measures and the
config-size GPU tests check). Each generator is the build's own: Gaussian-blob
phantoms, uniform random orientations, particles placed with a minimum
spacing, noise at a signal-to-noise ratio, and a missing wedge in the data.

Configurations (sizes of the reference workloads they stand for):
  c1  64^3 tomogram, 5 copies of a 24^3 phantom (6 blobs), SNR 0.5, soft mask
      r 10 (Gaussian edge 2), wedge semi-angle 30 and bandpass 2-20 on the
      template, 30-degree grid (1,008 orientations)
  c2  512x512x256, 300 copies of a 48^3 phantom (12 blobs), SNR 0.1, +-60
      degree tilt wedge in the data and the template weight, binary mask r 22,
      15-degree grid (7,488 orientations)
  c5  1024x1024x384, 2,000 particles of three 48^3 phantoms (8, 12, 16
      blobs), SNR 0.05, +-50 degree tilt wedge, 10-degree grid (24,624)
  c3  4,096 subtomograms of edges {32, 48, 64}, one shifted rotated copy of
      the box's phantom each, SNR U(0.05, 0.5), own wedge U(45, 70) degrees of
      tilt, 3x3x3 local grid of 5-degree steps around a prior
"""
from __future__ import annotations

import numpy as np

from . import synth

CONFIGS = {
    # name: volume (X,Y,Z), template edge, blobs, sigma, radius, particles,
    #       min spacing, SNR, wedge semi-angle, mask radius, angle step, seed
    "c1": dict(N=(64, 64, 64), n=24, k=6, sigma=1.5, R=8, P=5, spacing=24, snr=0.5, wedge=30.0,
               mask=("gauss", 10.0, 2.0), bp=(2.0, 20.0), step=(30, 30, 30), seed=1,
               desc="C1 plain template matching 64^3, 24^3 template, 30 deg (1,008 orientations)"),
    "c2": dict(N=(512, 512, 256), n=48, k=12, sigma=2.0, R=18, P=300, spacing=40, snr=0.1, wedge=30.0, tilt=60.0,
               mask=("binary", 22.0, 0.0), bp=None, step=(15, 15, 15), seed=2, tiles="16,32",
               desc="C2 plain exhaustive FLCF 512x512x256, 48^3 template, 15 deg (7,488 orientations)"),
    "c5": dict(N=(1024, 1024, 384), n=48, k=(8, 12, 16), sigma=2.0, R=18, P=2000, spacing=48, snr=0.05, wedge=40.0,
               tilt=50.0, mask=("binary", 22.0, 0.0), bp=None, step=(10, 10, 10), seed=5, tiles="8,16",
               desc="C5 exhaustive FLCF 1024x1024x384, three 48^3 templates (8, 12, 16 blobs), 10 deg "
                    "(24,624 orientations each)"),
}


def angle_grid(step):
    phi = np.arange(0, 360, step[0], dtype=np.float64)
    psi = np.arange(0, 360, step[1], dtype=np.float64)
    the = np.arange(0, 180 + 1e-9, step[2], dtype=np.float64)
    return np.asarray([(a, b, c) for a in phi for b in psi for c in the], np.float32)


def make_tomogram(cfg, api, torch, dev):
    """Seeded synthetic tomogram on the GPU (rotations by the library itself).
    Returns (V (Z, Y, X) float32 on dev, templates [(n, n, n)] one per
    phantom class, positions (P, 3) x y z, Euler angles (P, 3), class (P,))."""
    rng = np.random.default_rng(cfg["seed"])
    ks = cfg["k"] if isinstance(cfg["k"], (tuple, list)) else (cfg["k"],)
    Ts = [synth.phantom(rng, cfg["n"], k, cfg["sigma"], cfg["R"]) for k in ks]
    V, pos, eul, cls = synth.make_tomogram(cfg["N"], Ts, cfg["P"], cfg["spacing"], cfg["snr"], rng, cfg["seed"],
                                           api, torch, dev, tilt_max=cfg.get("tilt"))
    return V, Ts, pos, eul, cls


def masks(cfg, api):
    n = cfg["n"]
    kind, r, s = cfg["mask"]
    M = api.tom_spheremask_gauss((n, n, n), r, s if kind == "gauss" else 0.0)
    lo, hi = cfg["bp"] if cfg["bp"] else (0.0, 0.0)
    W = api.weight((n, n, n), cfg["wedge"], lo, hi)
    return M, W


C3 = dict(P=4096, edges=(32, 48, 64), A=27, step=5.0, seed=3,
          desc="C3 batched subtomogram scoring: 4,096 particles, edges {32,48,64}, 27 local orientations each")


def make_c3(api, torch, dev, P=C3["P"], seed=C3["seed"]):
    """Particles of uneven edges, each a rotated, shifted (+-4) copy of the
    phantom of its box size, missing wedge of its own tilt range
    (U(45, 70) degrees), SNR U(0.05, 0.5); a 3x3x3 local grid of 5-degree
    steps around a prior within 4 degrees of the true orientation."""
    from tom_amd import synth
    rng = np.random.default_rng(seed)
    edges = C3["edges"]
    boxes = []
    for B in edges:
        T = synth.phantom(rng, B, 12, 2.0 * B / 48, 18.0 * B / 48)
        M = api.tom_spheremask_gauss((B, B, B), B / 2 - 4, 2.0)
        Mcc = api.tom_spheremask_gauss((B, B, B), 5.0, 0.0)
        boxes.append({"template": T, "mask": M, "mask_cc": np.fft.ifftshift(Mcc)})
    pbox = rng.integers(0, len(edges), P).astype(np.int32)
    true = synth.random_euler(rng, P)
    shift = rng.integers(-4, 5, (P, 3))
    snr = rng.uniform(0.05, 0.5, P)
    tilt = rng.uniform(45.0, 70.0, P)
    prior = true + rng.uniform(-4, 4, (P, 3)).astype(np.float32)
    off = np.array([-C3["step"], 0, C3["step"]])
    grids = np.stack([[(a[0] + i, a[1] + j, a[2] + k) for i in off for j in off for k in off] for a in prior])
    sizes = np.array([edges[b] ** 3 for b in pbox])
    packed = torch.empty(int(sizes.sum()), dtype=torch.float32, device=dev)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    for bi, B in enumerate(edges):
        sel = np.nonzero(pbox == bi)[0]
        Td = torch.from_numpy(np.ascontiguousarray(boxes[bi]["template"].transpose(2, 1, 0))).to(dev)
        for s0 in range(0, sel.size, 512):
            part = sel[s0:s0 + 512]
            rot = api.tom_rotate_batch(Td, true[part])
            for j, p in enumerate(part):
                v = torch.roll(rot[j], shifts=(int(shift[p, 2]), int(shift[p, 1]), int(shift[p, 0])), dims=(0, 1, 2))
                v = synth.wedge_filter(v, float(tilt[p]), torch)
                v = v * float(np.sqrt(snr[p] / float(v.double().var())))
                v = v + torch.randn(v.shape, generator=g, device=dev)
                packed[offs[p]:offs[p] + sizes[p]] = v.reshape(-1)
            del rot
    semi = (90.0 - tilt).astype(np.float32)
    return boxes, packed, pbox, semi, grids.astype(np.float32), offs, sizes


