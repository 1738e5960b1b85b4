"""Train the proposal U-Net on tomograms labelled by the exhaustive search.

    python -m tom_amd.train_surrogate --out tom_amd/weights/unet_c2.npz

(the defaults reproduce the shipped weights: 96 tomograms, 24,000 Adam steps,
the 256-orientation coarse score map as the second input channel, targets at
the picks with a heat-map sigma of 1.5 voxels).

Training data: seeded synthetic tomograms (edge --size, --particles copies of
the template phantom at uniform random orientations, SNR U(--snr-lo, --snr-hi),
missing wedge of a +-60 degree tilt series), each searched exhaustively on the
GPU at 15 degrees; the picklist of the score map (tom_os3_returnPicklist
semantics) above --label-threshold gives the particle targets (a Gaussian heat
map of sigma --sigma voxels at each pick, or, with --targets centroid, at the
density centre p + R d of each pick: surrogate.template_centroid) and the best
orientation there the bin targets (bins within 4 voxels); voxels within
--offset-radius of a target learn the vector to it (the candidates are moved
by it). --dense adds targets from the exhaustive maps at every voxel (the best
score mapped linearly from [0.1, 0.5] to [0, 1], and the bin of the best
orientation wherever the score exceeds 0.25), which trains a broader response
that localises worse. The network's second input channel (--aux) is the best
score of a coarse scan over --coarse orientations (PrunedSearch.coarse) or the
FLCF map of the rotationally averaged template (surrogate.aux_map). The
template, mask and wedge weight are those of the 512x512x256 / 48^3 benchmark
workload (seed 2), so the trained network proposes candidates for that search.

Resuming (--state DIR): the picks of each labelled tomogram are kept in
DIR/label-NNNN.npz (a few kB: the exhaustive search is what labelling costs;
the tomogram and the coarse map are regenerated from their seeds), and the
training state (parameters, Adam moments and step, the crop generator, the
loss history) in DIR/train.npz every --ckpt-every steps and when --time-limit
seconds have passed. Running the same command again continues where the last
run stopped and, once training has finished, writes --out again; a state
directory made with other data or training settings is refused. Exit status
3 means the run stopped at --time-limit with its state saved.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload_template():
    """Template, mask, weight and angles of the 48^3 benchmark search."""
    from . import api, synth
    rng = np.random.default_rng(2)
    T = synth.phantom(rng, 48, 12, 2.0, 18)
    M = api.tom_spheremask_gauss((48, 48, 48), 22.0, 0.0)
    W = api.weight((48, 48, 48), 30.0, 0.0, 0.0)
    phi = np.arange(0, 360, 15.0)
    the = np.arange(0, 180 + 1e-9, 15.0)
    angles = np.asarray([(a, b, c) for a in phi for b in phi for c in the], np.float32)
    return T, M, W, angles


def label_tomogram(api, torch, dev, T, M, W, tn, mu, sg, angles, size, P, snr, seed, thr, bins, sigma=2.0,
                   dense=False, offset_radius=10.0, aux_kind="sphere", n_coarse=256, targets="centroid",
                   peaks=None):
    """peaks: the picks of an earlier labelling of the same tomogram (the
    exhaustive search is then skipped; not with dense targets)."""
    from . import surrogate, synth
    if peaks is not None and dense:
        raise ValueError("dense targets need the exhaustive maps")
    rng = np.random.default_rng(seed)
    V, pos, eul, _ = synth.make_tomogram((size, size, size), [T], P, 40, snr, rng, seed, api, torch, dev, tilt_max=60.0)
    with api.SearchPlan(V, T.shape, M, batch=4) as plan:
        plan.set_template(tn, W, mu, sg)
        if peaks is None:
            plan.run(angles, 0)
            score, ang = plan.finalize()
        if aux_kind == "coarse":
            aux = surrogate.PrunedSearch(n_coarse=n_coarse).coarse(plan, tn, W, mu, sg)[0].cpu().numpy()
        else:
            aux = surrogate.aux_map(plan, T, M, W).cpu().numpy()
    if peaks is None:
        peaks = surrogate.exhaustive_peaks(score, ang, T.shape[0], thr)
    peaks = np.asarray(peaks, np.float32).reshape(-1, 5)
    vz = V.cpu().numpy()
    # centroid targets sit at the density centres p + R d of the picks, which
    # a translation-equivariant network can locate without knowing R; peak
    # targets at the picks themselves
    cp = peaks.copy()
    if len(cp) and targets == "centroid":
        cp[:, :3] = np.rint(surrogate.centroid_positions(peaks, angles, surrogate.template_centroid(T)))
        cp[:, :3] = np.clip(cp[:, :3], 0, np.asarray(vz.shape[::-1]) - 1)
    tprob, tbin, toff = surrogate.label_volumes(vz.shape, cp, angles, bins, sigma=sigma,
                                                offset_radius=offset_radius)
    if dense:
        # dense FLCF targets everywhere, the peak heat map on top at the picks
        tp_dense, tb_dense = surrogate.dense_targets(score.cpu().numpy(), ang.cpu().numpy(), angles, bins)
        tprob = np.maximum(tp_dense, tprob)
        tbin = np.where(tbin >= 0, tbin, tb_dense)
    return vz, tprob, tbin, peaks, pos, aux, toff


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(os.path.dirname(__file__), "weights", "unet_c2.npz"))
    ap.add_argument("--tomos", type=int, default=96)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--particles", type=int, default=80)
    ap.add_argument("--snr-lo", type=float, default=0.05)
    ap.add_argument("--snr-hi", type=float, default=0.3)
    ap.add_argument("--label-threshold", type=float, default=0.3)
    ap.add_argument("--steps", type=int, default=24000)
    ap.add_argument("--crop", type=int, default=96)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--wpos", type=float, default=10.0)
    ap.add_argument("--seed", type=int, default=100)
    ap.add_argument("--sigma", type=float, default=1.5, help="heat-map sigma around the exhaustive peaks (voxels)")
    ap.add_argument("--dense", action="store_true", help="add dense targets from the exhaustive score map")
    ap.add_argument("--aux", choices=("sphere", "coarse"), default="coarse",
                    help="second input channel: FLCF of the rotationally averaged template, or the best score of a "
                         "coarse scan over --coarse orientations")
    ap.add_argument("--coarse", type=int, default=256)
    ap.add_argument("--targets", choices=("centroid", "peak"), default="peak")
    ap.add_argument("--offset-radius", type=float, default=10.0,
                    help="voxels within this distance of a peak learn the offset to it")
    ap.add_argument("--state", default=None, help="directory of label caches and training checkpoints (resume)")
    ap.add_argument("--ckpt-every", type=int, default=1000, help="training checkpoint interval (steps)")
    ap.add_argument("--time-limit", type=float, default=None,
                    help="seconds after which training checkpoints and exits with status 3")
    args = ap.parse_args(argv)
    if args.time_limit is not None and not args.state:
        ap.error("--time-limit needs --state")
    t_start = time.perf_counter()
    deadline = None if args.time_limit is None else t_start + args.time_limit

    import torch

    from . import api, orient, surrogate
    if not torch.cuda.is_available():
        raise SystemExit("training needs a GPU")
    dev = torch.device("cuda:0")
    T, M, W, angles = workload_template()
    bins = orient.orientation_bins(64, 0)
    tn, mu, sg = api.prepare_template(T, M, W)
    tr = surrogate.Trainer(crop=args.crop, nbin=64, seed=args.seed, lr=args.lr, wpos=args.wpos)
    # settings a cached label or checkpoint must have been made with
    label_cfg = json.dumps({k: getattr(args, k) for k in
                            ("size", "particles", "snr_lo", "snr_hi", "label_threshold", "seed")}, sort_keys=True)
    train_cfg = json.dumps({k: getattr(args, k) for k in
                            ("tomos", "size", "particles", "snr_lo", "snr_hi", "label_threshold", "steps", "crop",
                             "lr", "wpos", "seed", "sigma", "dense", "aux", "coarse", "targets", "offset_radius")},
                           sort_keys=True)
    ckpt = None
    if args.state:
        os.makedirs(args.state, exist_ok=True)
        ckpt = os.path.join(args.state, "train.npz")
    rng = np.random.default_rng(args.seed)
    t0 = time.perf_counter()
    npk = []
    for i in range(args.tomos):
        snr = float(rng.uniform(args.snr_lo, args.snr_hi))
        cache = os.path.join(args.state, f"label-{i:04d}.npz") if args.state and not args.dense else None
        cached = None
        if cache and os.path.exists(cache):
            z = np.load(cache, allow_pickle=False)
            if str(z["config"]) != label_cfg:
                raise SystemExit(f"{cache} was labelled with other settings: {z['config']}")
            cached = z["peaks"]
        vz, tp, tb, peaks, pos, aux, to = label_tomogram(api, torch, dev, T, M, W, tn, mu, sg, angles, args.size,
                                                         args.particles, snr, args.seed + i, args.label_threshold,
                                                         bins, sigma=args.sigma, dense=args.dense,
                                                         offset_radius=args.offset_radius, aux_kind=args.aux,
                                                         n_coarse=args.coarse, targets=args.targets, peaks=cached)
        if cache and cached is None:
            surrogate.save_npz_atomic(cache, peaks=peaks, snr=np.float64(snr), config=label_cfg)
        tr.add(vz, tp, tb, aux, to)
        npk.append(len(peaks))
        log(f"[label] tomogram {i}{' (cached picks)' if cached is not None else ''}: SNR {snr:.3f}, "
            f"{len(peaks)} peaks > {args.label_threshold}, "
            f"{surrogate.recall(pos, peaks[:, :3]):.3f} of the {len(pos)} planted particles among them "
            f"({time.perf_counter() - t0:.1f} s)")
        torch.cuda.empty_cache()
        # stop only after new work is saved, so every run moves forward
        if deadline is not None and time.perf_counter() > deadline and i + 1 < args.tomos and cache and cached is None:
            log(f"[label] time limit reached after tomogram {i}; run again to continue")
            raise SystemExit(3)
    t_label = time.perf_counter() - t0
    start, train_before = 0, 0.0
    if ckpt and os.path.exists(ckpt):
        start, z = tr.load(ckpt)
        if str(z["config"]) != train_cfg:
            raise SystemExit(f"{ckpt} was written with other settings: {z['config']}")
        train_before = float(z["train_s"])
        log(f"[train] resuming at step {start} of {args.steps}")
    t_seg = [time.perf_counter()]

    def save(step):
        tr.save(ckpt, step, config=train_cfg, train_s=np.float64(train_before + time.perf_counter() - t_seg[0]))
        log(f"[train] checkpoint at step {step}")

    t_train = train_before + tr.train(args.steps, log=log, every=250, start=start,
                                      checkpoint=save if ckpt else None, checkpoint_every=args.ckpt_every,
                                      deadline=deadline)
    if ckpt:
        if not tr.finished:
            log("[train] time limit reached; run again to continue")
            raise SystemExit(3)
        if start < args.steps:
            tr.save(ckpt, args.steps, config=train_cfg, train_s=np.float64(t_train))
    theta = tr.net.get_params()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    meta = dict(vars(args), label_s=t_label, train_s=t_train, peaks_per_tomogram=npk,
                final_loss=[float(np.mean([h[k] for h in tr.history[-100:]])) for k in range(3)])
    extra = {"centroid": surrogate.template_centroid(T)} if args.targets == "centroid" else {}
    np.savez_compressed(args.out, params=theta, bins=bins, meta=json.dumps(meta), **extra)
    log(f"[train] saved {args.out}: label {t_label:.1f} s, train {t_train:.1f} s")
    print(json.dumps(meta))


if __name__ == "__main__":
    main()
