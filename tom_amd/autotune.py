"""Autotuner for the exhaustive search. It picks, on the caller's own volume
and template, the orientation batch (orientations per launch group: the
template-spectrum and Q tiles and the register-resident max of the c2r pass)
the column-tile widths of the z- and y-FFT passes and the z-chunking of the
last two passes, by timing a few launch groups of each candidate through
hipGraph replay as in production. The search is staged: batch first (default
tiles), then the z tile, then the y tile, then the chunk.

None of these change the mathematics (each voxel keeps the first strictly
better orientation in list order, whatever the grouping); the FP32 sums may
round differently by a few ulp, which can flip the best orientation only in
near-ties. `tune_search` compares every candidate's maps with the first and
fails loudly when a score differs by more than 1e-6 or when more than 1e-5 of
the voxels change their best orientation.
Results are cached per (device, volume shape, template shape) in
$TOM_AMD_CACHE (default ~/.cache/tom_amd/autotune.json).
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from . import api

BATCHES = (2, 4, 6, 8, 12, 16)
TILES = (8, 16, 32)
ZCHUNKS = (0, 64, 32)
# largest score difference accepted between candidates (FP32 rounding of the
# same sums in another order)
SCORE_ULP_TOL = 1e-6
# largest fraction of voxels whose best orientation may flip (near-ties)
ANGLE_FLIP_FRAC = 1e-5


def _cache_path():
    root = os.environ.get("TOM_AMD_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "tom_amd"))
    return os.path.join(root, "autotune.json")


def _key(volume_dims, template_shape):
    import torch
    name = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "cpu"
    return f"{name}|{tuple(int(d) for d in volume_dims)}|{tuple(int(d) for d in template_shape)}"


def tune_search(volume, tmpl, mask, W, mu_t, sig_t, angles, batches=BATCHES, tiles=TILES, zchunks=ZCHUNKS,
                groups=6, use_cache=True, log=None):
    """Returns {"batch", "tile_z", "tile_y", "zchunk", "ms_per_orientation": {...}}.
    volume: a device tensor (z, y, x) or numpy (x, y, z); tmpl: the prepared
    template (tomhip_prepare_template_f32 output)."""
    import torch
    vb = api._Buf(volume)
    tshape = tuple(reversed(tmpl.shape)) if api._is_torch(tmpl) else np.shape(tmpl)
    key = _key(vb.dims, tshape)
    path = _cache_path()
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            with open(path) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
        if key in cache:
            return dict(cache[key], cached=True)
    ang = np.asarray(angles, np.float32)[:max(batches) * groups]
    timing = {}
    ref = []

    def trial(ka, tz, ty, zc):
        name = f"batch{ka}_z{tz}_y{ty}_c{zc}"
        with api.SearchPlan(volume, tshape, mask, batch=ka) as p:
            try:
                p.set_tiles(tz, ty)
            except api.TomHipError:
                return None  # tile does not fit this volume
            p.set_zchunk(zc)
            p.set_template(tmpl, W, mu_t, sig_t)
            p.run(ang[:2 * ka], 0)  # capture + warm
            p.set_template(tmpl, W, mu_t, sig_t)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            p.run(ang, 0)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) / len(ang) * 1e3
            sc, an = p.finalize()
            sc, an = (sc.cpu().numpy(), an.cpu().numpy()) if api._is_torch(sc) else (sc, an)
        if not ref:
            ref.append((sc, an))
        else:
            # FP32 sums may round differently with another grouping (a few
            # ulp); the best orientation may then differ only in near-ties
            d = np.abs(sc - ref[0][0])
            flips = int(np.count_nonzero(an != ref[0][1]))
            if not (d.max() <= SCORE_ULP_TOL and flips <= ANGLE_FLIP_FRAC * an.size):
                raise RuntimeError(f"autotune: {name} changed the search result "
                                   f"(max |d score| {d.max():.2e}, {flips} orientation changes)")

        timing[name] = ms
        if log:
            log(f"[autotune] {name}: {ms:.4f} ms/orientation")
        torch.cuda.empty_cache()
        return ms

    best = {"batch": 4, "tile_z": 16, "tile_y": 16, "zchunk": 0}
    for stage, values in (("batch", batches), ("tile_z", tiles), ("tile_y", tiles), ("zchunk", zchunks)):
        scores = {}
        for v in values:
            cfg = dict(best, **{stage: v})
            ms = trial(cfg["batch"], cfg["tile_z"], cfg["tile_y"], cfg["zchunk"])
            if ms is not None:
                scores[v] = ms
        if scores:
            best[stage] = min(scores, key=scores.get)
    res = dict(best, ms_per_orientation={k: round(v, 5) for k, v in timing.items()},
               orientations_timed=int(len(ang)))
    if use_cache:
        cache[key] = res
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            with open(path, "w") as f:
                json.dump(cache, f, indent=1)
        except OSError:
            pass
    return dict(res, cached=False)
