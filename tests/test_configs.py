"""The benchmark workloads at their stated sizes on the GPU, checked against
the FP64 oracle (tom_amd.workloads holds the seeded generators).

* C1 (64^3, 24^3 template, 1,008 orientations, soft mask, wedge x bandpass):
  the whole search against the oracle's full FP64 maps.
* C2 (512x512x256, 48^3, 7,488 orientations) and C5 (1024x1024x384, 48^3,
  three phantom classes; an orientation subset): the GPU search over the
  whole volume, checked at sampled voxels (planted centres and random ones)
  against the oracle's FP64 direct sums for every orientation scanned.
* C3 (4,096 subtomograms of edges 32/48/64, 27 orientations each, own
  wedges, one persistent launch): the oracle on a spread sample of particles.

Tolerances: scores within 1e-4 of FP64; best-angle indices (and C3 peaks)
identical wherever the oracle's best/second-best margin exceeds 1e-4."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCORE_TOL = 1e-4
MARGIN = 1e-4
WORKERS = min(16, os.cpu_count() or 1)


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def env():
    import torch

    from tom_amd import api, workloads
    api.context()
    return api, workloads, torch, torch.device("cuda:0")


def _check_sampled(oracle, api, score, ang, Vh, T, M, W, angles, vox):
    """GPU best score / angle at vox against the oracle over all angles."""
    tn, mu_t, sig_t = oracle.prepare_template(T, M.astype(np.float64), W.astype(np.float64))
    s = oracle.flcf_at_voxels(Vh, tn, M, mu_t, sig_t, angles, vox, W=W.astype(np.float64), workers=WORKERS)
    best = s.max(1)
    arg = s.argmax(1)
    srt = np.sort(s, 1)
    margin = srt[:, -1] - srt[:, -2]
    g_s = score[vox[:, 2], vox[:, 1], vox[:, 0]]
    g_a = ang[vox[:, 2], vox[:, 1], vox[:, 0]]
    err = np.abs(g_s - best).max()
    sel = margin > MARGIN
    mism = int(np.sum(g_a[sel] != arg[sel]))
    say(f"  {len(vox)} voxels x {len(angles)} orientations: max |score - fp64| {err:.2e}, "
        f"angle mismatches (margin > 1e-4) {mism} of {int(sel.sum())}")
    assert err < SCORE_TOL
    assert mism == 0
    return best


def _voxels(rng, N, planted, k_random, n_planted):
    pl = np.asarray(planted, np.int64)[:n_planted]
    rnd = np.stack([rng.integers(0, N[d], k_random) for d in range(3)], 1)
    return np.concatenate([pl, rnd])


def test_c1_full_grid(env, oracle):
    api, wl, torch, dev = env
    cfg = wl.CONFIGS["c1"]
    say("C1: 64^3, 1,008 orientations, full maps against FP64")
    V, Ts, pos, eul, cls = wl.make_tomogram(cfg, api, torch, dev)
    T = Ts[0]
    M, W = wl.masks(cfg, api)
    angles = wl.angle_grid(cfg["step"]).astype(np.float64)
    assert angles.shape[0] == 1008
    score, ang = api.tom_os3_findTemplate(V, T, M, angles, W=W, batch=16)
    score = score.cpu().numpy().transpose(2, 1, 0)
    ang = ang.cpu().numpy().transpose(2, 1, 0) - 1  # 1-based, as findTemplate reports it
    # the oracle's FP64 loop, keeping the best and second-best per voxel
    Vh = V.cpu().numpy().transpose(2, 1, 0).astype(np.float64)
    Md = M.astype(np.float64)
    Wd = W.astype(np.float64)
    m = float((Md > 0).sum())
    oracle.set_fft_workers(WORKERS)
    try:
        Tn, mu_T, sig_T = oracle.prepare_template(T, Md, Wd)
        mu_I, sig_I = oracle.local_stats(Vh, Md, m)
        FV = oracle._fftn(Vh)
        b1 = np.full(Vh.shape, -np.inf)
        b2 = np.full(Vh.shape, -np.inf)
        arg = np.zeros(Vh.shape, np.int64)
        for i, a in enumerate(angles):
            s = oracle.flcf_angle(Vh, Tn, Md, m, mu_T, sig_T, mu_I, sig_I, FV, a, Wd)
            up = s > b1
            b2 = np.where(up, b1, np.maximum(b2, s))
            b1 = np.where(up, s, b1)
            arg = np.where(up, i, arg)
    finally:
        oracle.set_fft_workers(1)
    err = np.abs(score - b1).max()
    sel = (b1 - b2) > MARGIN
    mism = int(np.sum(ang[sel] != arg[sel]))
    say(f"  max |score - fp64| {err:.2e}, angle mismatches {mism} of {int(sel.sum())} decided voxels")
    assert err < SCORE_TOL and mism == 0
    # each planted copy is found: the best score within 2 voxels of its centre
    # (a 30-degree grid moves the peak of an off-grid orientation by a voxel
    # or two) is above 0.3
    for p in pos:
        x, y, z = p
        near = score[max(x - 2, 0):x + 3, max(y - 2, 0):y + 3, max(z - 2, 0):z + 3].max()
        say(f"  planted {list(p)}: score {score[x, y, z]:.3f}, best within 2 voxels {near:.3f}")
        assert near > 0.3


def test_c2_full_size_sampled_voxels(env, oracle):
    api, wl, torch, dev = env
    cfg = wl.CONFIGS["c2"]
    say("C2: 512x512x256, 7,488 orientations, 200 sampled voxels against FP64")
    V, Ts, pos, eul, cls = wl.make_tomogram(cfg, api, torch, dev)
    T = Ts[0]
    M, W = wl.masks(cfg, api)
    angles = wl.angle_grid(cfg["step"])
    assert angles.shape[0] == 7488
    tn, mu_t, sig_t = api.prepare_template(T, M, W)
    with api.SearchPlan(V, (48, 48, 48), M, batch=16) as plan:
        plan.set_tiles(16, 32)
        plan.set_template(tn, W, mu_t, sig_t)
        plan.run(angles, 0)
        score, ang = plan.finalize()
    score = score.cpu().numpy()
    ang = ang.cpu().numpy()
    Vh = V.cpu().numpy().transpose(2, 1, 0)
    vox = _voxels(np.random.default_rng(22), cfg["N"], pos, 100, 100)
    best = _check_sampled(oracle, api, score, ang, Vh, T, M, W, angles.astype(np.float64), vox)
    assert np.mean(best[:100] > 0.3) > 0.95  # planted centres score high


def test_c5_full_size_orientation_subset(env, oracle):
    api, wl, torch, dev = env
    cfg = wl.CONFIGS["c5"]
    say("C5: 1024x1024x384, three phantom classes, 48 orientations, sampled voxels against FP64")
    V, Ts, pos, eul, cls = wl.make_tomogram(cfg, api, torch, dev)
    assert len(Ts) == 3 and set(np.unique(cls)) == {0, 1, 2} and len(pos) == 2000
    M, W = wl.masks(cfg, api)
    full = wl.angle_grid(cfg["step"])
    assert full.shape[0] == 24624
    rng = np.random.default_rng(55)
    angles = full[np.sort(rng.choice(full.shape[0], 48, replace=False))]
    c = 1
    T = Ts[c]
    tn, mu_t, sig_t = api.prepare_template(T, M, W)
    with api.SearchPlan(V, (48, 48, 48), M, batch=16) as plan:
        plan.set_tiles(8, 16)
        plan.set_template(tn, W, mu_t, sig_t)
        plan.run(angles, 0)
        score, ang = plan.finalize()
    score = score.cpu().numpy()
    ang = ang.cpu().numpy()
    Vh = V.cpu().numpy().transpose(2, 1, 0)
    del V
    vox = _voxels(rng, cfg["N"], pos[cls == c], 60, 60)
    _check_sampled(oracle, api, score, ang, Vh, T, M, W, angles.astype(np.float64), vox)


def test_c3_full_batch_sampled_particles(env, oracle):
    api, wl, torch, dev = env
    say("C3: 4,096 subtomograms x 27 orientations in one launch, 32 particles against FP64")
    boxes, packed, pbox, semi, grids, offs, sizes = wl.make_c3(api, torch, dev)
    assert len(pbox) == 4096 and grids.shape == (4096, 27, 3)
    best, peak, ccc = api.subtomo_score_batch(boxes, packed, pbox, semi, grids)
    host = packed.cpu().numpy()
    idx = np.linspace(0, 4095, 32).round().astype(int)
    err, mism, decided = 0.0, 0, 0
    for p in idx:
        B = wl.C3["edges"][pbox[p]]
        v = host[offs[p]:offs[p] + sizes[p]].reshape(B, B, B).transpose(2, 1, 0)
        bx = boxes[pbox[p]]
        i, pos, val, vals = oracle.subtomo_score(v, bx["template"], bx["mask"], bx["mask_cc"], float(semi[p]),
                                                 grids[p].astype(np.float64))
        err = max(err, abs(float(ccc[p]) - val))
        srt = np.sort(vals)
        if srt[-1] - srt[-2] > MARGIN:
            decided += 1
            mism += int(best[p] != i or list(peak[p]) != list(pos))
    say(f"  max |ccc - fp64| {err:.2e}, decision mismatches {mism} of {decided}")
    assert err < SCORE_TOL and mism == 0
    assert np.all(np.isfinite(ccc)) and np.all(best >= 0)
