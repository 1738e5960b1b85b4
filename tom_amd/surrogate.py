"""Pruned template matching (no reference counterpart): instead of scanning
every orientation over the whole tomogram, find candidate particles cheaply
and score only a local orientation set at each of them with the exact FLCF
(the same score as the exhaustive search).

* PrunedSearch: the exhaustive machinery scans a few hundred to a few thousand
  orientations spread over SO(3); local maxima of that best-score map are the
  candidates; two or three levels of exact local refinement (Euler grids of
  shrinking step around the best orientation so far, 27 voxels around the
  best position so far) recover the fine orientation and score.
* Surrogate: a 3D U-Net (BF16 MFMA kernels of this library, trained on the
  GPU) proposes particle positions and one of 64 coarse orientation bins from
  the tomogram and a second channel (the FLCF map of the rotationally averaged
  template, or the coarse scan's score map); candidates are refined over
  +-15 degrees in 5-degree steps around the bin centre at 27 voxels, or passed
  to PrunedSearch as its proposer.

Training labels come from the exhaustive search itself: peaks of the FLCF map
(tom_os3_returnPicklist semantics) above a score threshold mark particles, and
the best orientation there gives the bin.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from . import api, orient


# layers of the proposal U-Net: (Cin, Cout, taps, resolution level)
UNET_LAYERS = [(16, 16, 27, 0), (16, 16, 27, 0), (16, 32, 27, 1), (32, 32, 27, 1), (32, 64, 27, 2), (64, 64, 27, 2),
               (96, 32, 27, 1), (32, 32, 27, 1), (48, 16, 27, 0), (16, 16, 27, 0), (16, 80, 1, 0)]


def unet_forward_flops(dims) -> float:
    """Flops of one forward pass over a volume of `dims` voxels: every layer's
    multiply-adds as the network defines them, the first layer with its two
    real input channels (the tomogram and the aux map; its 14 zero padding
    channels are not computed) and the head with all 80 output channels it
    computes. The padding the MFMA tiles add (a tenth K step of zeros for
    16-channel inputs) is not counted."""
    V = float(np.prod(dims))
    return sum(2.0 * V / 8 ** lvl * t * (2 if i == 0 else ci) * co
               for i, (ci, co, t, lvl) in enumerate(UNET_LAYERS))


# ---------------------------------------------------------------- auxiliary input
def spherical_average(T):
    """Rotational average of a template (x, y, z) about its centre floor(n/2):
    each voxel takes the mean of the voxels at the same rounded distance."""
    T = np.asarray(T, np.float64)
    g = np.meshgrid(*[np.arange(s) - s // 2 for s in T.shape], indexing="ij")
    r = np.rint(np.sqrt(sum(x * x for x in g))).astype(np.int64)
    sums = np.bincount(r.ravel(), T.ravel())
    cnt = np.bincount(r.ravel())
    return (sums / np.maximum(cnt, 1))[r].astype(np.float32)


def aux_map(plan, template, mask, W=None):
    """FLCF score map of the rotationally averaged template (one orientation
    of the exhaustive machinery): an orientation-free matched filter that the
    network gets as its second input channel. Leaves the plan's template set to
    the averaged one."""
    Ta = spherical_average(template)
    tn, mu, sg = api.prepare_template(Ta, mask, W)
    plan.set_template(tn, W, mu, sg)
    plan.run(np.zeros((1, 3), np.float32), 0)
    sc, _ = plan.finalize()
    return sc


# ---------------------------------------------------------------- labels
def exhaustive_peaks(score, angle, template_size: int, threshold: float, max_peaks: int = 1000):
    """Picklist (tom_os3_returnPicklist: border erase, tom_peakc radius
    ceil(n/2)) of an exhaustive score map, keeping rows with value > threshold.
    Returns (P, 5) rows x y z (0-based) value angle-index."""
    pk = api.tom_os3_returnPicklist(score, (template_size,) * 3, max_peaks, angles=angle, erase_borders=True)
    pk = np.asarray(pk, np.float64).reshape(-1, 5)
    pk = pk[pk[:, 3] > threshold]
    pk[:, :3] -= 1
    return pk


OFFSET_SCALE = 8.0  # offset channels predict (centre - voxel) / OFFSET_SCALE


def template_centroid(T):
    """Centre of mass of the template's positive density, x y z in voxels from
    the rotation centre ceil(n/2) (0-based) of tom_rotate. A particle whose
    exhaustive peak is at p with orientation R has its density centre at
    p + R d: the proposal network locates that point, which does not depend
    on the orientation the way the peak does."""
    T = np.clip(np.asarray(T, np.float64), 0, None)
    g = np.indices(T.shape, dtype=np.float64)
    c = np.array([(s + 1) // 2 for s in T.shape], np.float64)
    return np.array([(g[i] * T).sum() / T.sum() for i in range(3)]) - c


def centroid_positions(peaks, angles, d):
    """Rows x y z of the density centres p + R(angle) d of picklist rows
    (x y z value angle-index)."""
    if len(peaks) == 0:
        return np.zeros((0, 3))
    R = orient.rotmatrix(np.asarray(angles, np.float64)[peaks[:, 4].astype(np.int64)])
    return peaks[:, :3] + R @ np.asarray(d, np.float64)


def label_volumes(shape_zyx, peaks, angles, bins, sigma: float = 3.0, bin_radius: float = 4.0,
                  offset_radius: float = 0.0):
    """Targets for the proposal network: tprob = exp(-d^2 / (2 sigma^2)) of
    the distance d to the nearest peak (0 beyond 3 sigma), tbin = bin of that
    peak's best orientation within bin_radius of it (-1 elsewhere).
    peaks: (P, 5) rows x y z value angle-index; angles: the search's list.
    With offset_radius > 0 also returns toff (z, y, x, 4): (dx, dy, dz) /
    OFFSET_SCALE from the voxel to the nearest peak and weight 1 within
    offset_radius of it (0 elsewhere)."""
    tprob = np.zeros(shape_zyx, np.float32)
    tbin = np.full(shape_zyx, -1, np.int32)
    toff = np.zeros(tuple(shape_zyx) + (4,), np.float32) if offset_radius > 0 else None
    if len(peaks) == 0:
        return (tprob, tbin) if toff is None else (tprob, tbin, toff)
    b = orient.assign_bins(angles[peaks[:, 4].astype(np.int64)], bins)
    r = int(np.ceil(max(3 * sigma, offset_radius)))
    g = np.arange(-r, r + 1)
    dz, dy, dx = np.meshgrid(g, g, g, indexing="ij")
    d2 = (dx * dx + dy * dy + dz * dz).ravel()
    keep = d2 <= max(3 * sigma, offset_radius) ** 2
    off = np.stack([dz.ravel(), dy.ravel(), dx.ravel()], 1)[keep]
    d2 = d2[keep]
    val = np.where(d2 <= (3 * sigma) ** 2, np.exp(-d2 / (2 * sigma * sigma)), 0.0).astype(np.float32)
    inb = d2 <= bin_radius * bin_radius
    ino = d2 <= offset_radius * offset_radius
    Z, Y, X = shape_zyx
    best = np.zeros(shape_zyx, np.float32)
    near = np.full(shape_zyx, np.inf, np.float32)
    for (x, y, z), bb in zip(peaks[:, :3].astype(np.int64), b):
        zz, yy, xx = z + off[:, 0], y + off[:, 1], x + off[:, 2]
        ok = (zz >= 0) & (zz < Z) & (yy >= 0) & (yy < Y) & (xx >= 0) & (xx < X)
        zz, yy, xx, v, ib, io, dd, o = zz[ok], yy[ok], xx[ok], val[ok], inb[ok], ino[ok], d2[ok], off[ok]
        up = v > best[zz, yy, xx]
        best[zz[up], yy[up], xx[up]] = v[up]
        sel = up & ib
        tbin[zz[sel], yy[sel], xx[sel]] = bb
        if toff is not None:
            nr = io & (dd < near[zz, yy, xx])
            near[zz[nr], yy[nr], xx[nr]] = dd[nr]
            # voxel -> centre is minus the stencil offset (centre -> voxel)
            toff[zz[nr], yy[nr], xx[nr], 0] = -o[nr, 2] / OFFSET_SCALE
            toff[zz[nr], yy[nr], xx[nr], 1] = -o[nr, 1] / OFFSET_SCALE
            toff[zz[nr], yy[nr], xx[nr], 2] = -o[nr, 0] / OFFSET_SCALE
            toff[zz[nr], yy[nr], xx[nr], 3] = 1.0
    tprob[:] = best
    return (tprob, tbin) if toff is None else (tprob, tbin, toff)


def dense_targets(score_zyx, angle_zyx, angles, bins, s0: float = 0.1, s1: float = 0.5, bin_thr: float = 0.25):
    """Dense targets from the exhaustive maps themselves: tprob = clip((S - s0)
    / (s1 - s0), 0, 1) of the best FLCF score S at every voxel, tbin = bin of
    the voxel's best orientation where S > bin_thr (-1 elsewhere)."""
    S = np.asarray(score_zyx, np.float32)
    tprob = np.clip((S - s0) / (s1 - s0), 0.0, 1.0).astype(np.float32)
    tbin = np.full(S.shape, -1, np.int32)
    sel = S > bin_thr
    if sel.any():
        a = np.asarray(angle_zyx)[sel].astype(np.int64)
        ua, inv = np.unique(a, return_inverse=True)
        tbin[sel] = orient.assign_bins(np.asarray(angles)[ua], bins)[inv]
    return tprob, tbin


# ---------------------------------------------------------------- training
class Trainer:
    """Adam on random crops (edge `crop`, a multiple of 4) of labelled
    tomograms. Volumes are normalised by their global mean and std, as
    Unet.infer does."""

    def __init__(self, crop=64, nbin=64, seed=0, lr=1e-3, wpos=20.0, lr_min=5e-5):
        self.crop = int(crop)
        self.net = api.Unet((crop, crop, crop), nbin=nbin, seed=seed, train=True)
        self.rng = np.random.default_rng(seed)
        self.lr, self.lr_min = lr, lr_min
        self.wpos = wpos
        self.data = []
        self.history = []

    def add(self, volume_zyx, tprob, tbin, aux_zyx=None, toff=None):
        """volume_zyx (and the aux map): numpy (z, y, x) float32; targets of
        the same shape (toff: (z, y, x, 4) offset targets, optional)."""
        def norm(a):
            a = np.asarray(a, np.float64)
            return ((a - a.mean()) / a.std()).astype(np.float32)
        vn = norm(volume_zyx)
        an = norm(aux_zyx) if aux_zyx is not None else None
        pos = np.argwhere(tprob > 0.5)
        self.data.append((vn, tprob, tbin, pos, an, toff))

    def _crop(self):
        vn, tp, tb, pos, an, to = self.data[self.rng.integers(len(self.data))]
        c = self.crop
        Z, Y, X = vn.shape
        if len(pos) and self.rng.random() < 0.7:
            z, y, x = pos[self.rng.integers(len(pos))] - self.rng.integers(4, c - 4, 3)
        else:
            z, y, x = self.rng.integers(0, Z - c + 1), self.rng.integers(0, Y - c + 1), self.rng.integers(0, X - c + 1)
        z, y, x = (int(np.clip(z, 0, Z - c)), int(np.clip(y, 0, Y - c)), int(np.clip(x, 0, X - c)))
        sl = (slice(z, z + c), slice(y, y + c), slice(x, x + c))
        f = lambda a: np.asfortranarray(np.ascontiguousarray(a[sl]).transpose(2, 1, 0))  # noqa: E731
        # offsets keep their component axis fastest: (4, x, y, z) F-order
        o = np.ascontiguousarray(to[sl]).transpose(3, 2, 1, 0) if to is not None else None
        return f(vn), f(tp), f(tb), (f(an) if an is not None else None), o

    def step(self, lr=None):
        v, tp, tb, a, o = self._crop()
        self.net.forward(v, a)
        bce, ce, off = self.net.backward(tp, tb, self.wpos, toff=o)
        self.net.adam(self.lr if lr is None else lr)
        self.history.append((bce, ce, off))
        return bce, ce, off

    def train(self, steps, log=None, every=100, start=0, checkpoint=None, checkpoint_every=0, deadline=None):
        """Cosine learning-rate decay from lr to lr_min over the steps, from
        loop index `start` (a resumed run). checkpoint(i) is called with the
        next loop index every checkpoint_every steps; past `deadline`
        (time.perf_counter() value) the loop checkpoints and stops early.
        self.finished tells whether it reached `steps`."""
        t0 = time.perf_counter()
        self.finished = False
        for i in range(start, steps):
            lr = self.lr_min + 0.5 * (self.lr - self.lr_min) * (1 + np.cos(np.pi * i / max(1, steps - 1)))
            bce, ce, off = self.step(lr)
            if log and (i % every == 0 or i == steps - 1):
                log(f"[train] step {i}: bce {bce:.4f} ce {ce:.4f} offset {off:.4f} "
                    f"({time.perf_counter() - t0:.1f} s)")
            if checkpoint is not None and i + 1 < steps:
                late = deadline is not None and time.perf_counter() > deadline
                if late or (checkpoint_every and (i + 1) % checkpoint_every == 0):
                    checkpoint(i + 1)
                if late:
                    return time.perf_counter() - t0
        self.finished = True
        return time.perf_counter() - t0

    # ---- checkpoints: network parameters, Adam moments and step, the crop
    # generator's state and the loss history, so a resumed run takes the same
    # crops and updates as one that was never stopped
    def state(self):
        m, v, t = self.net.get_optimizer()
        return dict(params=self.net.get_params(), adam_m=m, adam_v=v, adam_step=np.int64(t),
                    rng=json.dumps(self.rng.bit_generator.state),
                    history=np.asarray(self.history, np.float64).reshape(-1, 3))

    def restore(self, st):
        self.net.set_params(st["params"])
        self.net.set_optimizer(st["adam_m"], st["adam_v"], int(st["adam_step"]))
        self.rng.bit_generator.state = json.loads(str(st["rng"]))
        self.history = [tuple(float(x) for x in h) for h in np.asarray(st["history"])]

    def save(self, path, step, **extra):
        """Write the checkpoint at loop index `step` (atomically: a run
        killed while writing leaves the previous checkpoint)."""
        save_npz_atomic(path, step=np.int64(step), **self.state(), **extra)

    def load(self, path):
        """Restore a checkpoint; returns (loop index to resume at, the npz)."""
        z = np.load(path, allow_pickle=False)
        self.restore(z)
        return int(z["step"]), z


def save_npz_atomic(path, **arrays):
    """np.savez to a temporary file in the same directory, then rename."""
    tmp = path + ".part"
    with open(tmp, "wb") as f:
        np.savez(f, **arrays)
    os.replace(tmp, path)


# ---------------------------------------------------------------- proposal + refinement
def greedy_nms(points, values, radius, k):
    """Indices of at most k points taken in descending value order, skipping
    any within `radius` of one already taken."""
    from scipy.spatial import cKDTree
    order = np.argsort(-values, kind="stable")
    tree = cKDTree(points)
    taken = []
    dead = np.zeros(len(points), bool)
    for i in order:
        if dead[i]:
            continue
        taken.append(i)
        if len(taken) >= k:
            break
        dead[tree.query_ball_point(points[i], radius - 1e-9)] = True
    return np.asarray(taken, np.int64)


class Surrogate:
    """Proposal network + exact refinement. Volumes of up to `whole_max`
    voxels are inferred in one tile (no seams); larger ones in cubic tiles of
    edge `tile` with `halo` voxels of context."""

    def __init__(self, params, bins, tile=304, halo=24, half=15.0, step=5.0, nbin=64, whole_max=160_000_000,
                 snap_radius=0, shift_iters=2, centroid=None):
        self.params = np.asarray(params, np.float32)
        self.nbin = nbin
        self.bins = np.asarray(bins, np.float64)
        self.tile, self.halo, self.whole_max = tile, halo, whole_max
        self.half, self.step = half, step
        self.snap_radius = snap_radius
        self.shift_iters = shift_iters
        # template_centroid of the template when the network locates density
        # centres (the refinement then scores each orientation around
        # candidate - R d); None when it locates the peaks themselves
        self.centroid = None if centroid is None else np.asarray(centroid, np.float64)
        self.timings = {}
        self._net, self._net_dims = None, None

    def _network(self, dims):
        dims = tuple(int(d) for d in dims)
        if int(np.prod(dims)) <= self.whole_max:
            want, halo = tuple(-(-d // 4) * 4 for d in dims), 0
        else:
            want, halo = (self.tile,) * 3, self.halo
        if self._net_dims != want:
            if self._net is not None:
                self._net.close()
            self._net = api.Unet(want, nbin=self.nbin, seed=0, train=False)
            self._net.set_params(self.params)
            self._net_dims = want
        return self._net, halo

    def propose(self, volume, K=2000, nms_radius=12.0, threshold=0.05, aux=None):
        """Top-K candidates (K, 5): x y z bin logit, taken from local maxima of
        the particle logit above logit(threshold) (the logit does not saturate,
        so a maximum sits at its peak, not on a plateau of probabilities that
        round to 1). aux: the aux_map of the volume."""
        dims = tuple(reversed(volume.shape)) if hasattr(volume, "cpu") else np.shape(volume)
        net, halo = self._network(dims)
        t0 = time.perf_counter()
        shift = self.shift_iters > 0
        out = net.infer(volume, halo, aux=aux, logit=True, offsets=shift)
        prob, b = out[:2]
        _sync(prob)
        t1 = time.perf_counter()
        lm = api.local_maxima(prob, 2, float(np.log(threshold / (1 - threshold))))
        t2 = time.perf_counter()
        if shift and len(lm):
            # move each maximum along the predicted vector to the particle
            # centre (repeated: the vector is read again where it points)
            xyz = lm[:, :3].astype(np.int64)
            for _ in range(self.shift_iters):
                o = _gather(out[2], xyz, vec=True)
                xyz = np.clip(np.rint(xyz + OFFSET_SCALE * o), 0, np.asarray(dims) - 1).astype(np.int64)
            lm = np.concatenate([xyz.astype(np.float32), lm[:, 3:]], 1)
        keep = greedy_nms(lm[:, :3], lm[:, 3], nms_radius, K) if len(lm) else np.zeros(0, np.int64)
        lm = lm[keep]
        if aux is not None and self.snap_radius > 0 and len(lm):
            # recentre on the matched-filter map: its maximum sits at the
            # particle centre, where the network's may drift by a few voxels
            xyz = api.snap_maxima(aux, lm[:, :3].astype(np.int32), self.snap_radius)
            lm = np.concatenate([xyz.astype(np.float32), lm[:, 3:]], 1)
            _, first = np.unique(xyz, axis=0, return_index=True)
            lm = lm[np.sort(first)]
        bins = _gather(b, lm[:, :3].astype(np.int64))
        t3 = time.perf_counter()
        self.timings.update(infer_s=t1 - t0, local_max_s=t2 - t1, nms_s=t3 - t2)
        self.infer_flops = unet_forward_flops(net.dims) * _tiles(dims, net.dims, halo)
        return np.concatenate([lm[:, :3], bins[:, None].astype(np.float32), lm[:, 3:4]], 1)

    def rerank(self, volume, coarse, xyz, noise, alpha=2.0, radius=2):
        """Rank values of coarse-scan local maxima (rows x y z): the robust z
        of the coarse best score, (s - median) / sigma with noise = (median,
        sigma) of the coarse map, plus alpha times the largest particle logit
        of the network within the cube of half-edge `radius` around the
        maximum (the network's peak may sit a voxel or two off the matched
        filter's). The network sees the tomogram and the coarse map, as in
        propose()."""
        dims = tuple(reversed(volume.shape)) if hasattr(volume, "cpu") else np.shape(volume)
        net, halo = self._network(dims)
        t0 = time.perf_counter()
        logit = net.infer(volume, halo, aux=coarse, logit=True)[0]
        _sync(logit)
        t1 = time.perf_counter()
        xyz = np.asarray(xyz, np.int64).reshape(-1, 3)
        peak = api.snap_maxima(logit, xyz, radius).astype(np.int64)
        lv = _gather(logit, peak).astype(np.float64)
        cs = _gather(coarse, xyz).astype(np.float64)
        med, sig = noise
        self.timings.update(infer_s=t1 - t0, rerank_s=time.perf_counter() - t1)
        self.infer_flops = unet_forward_flops(net.dims) * _tiles(dims, net.dims, halo)
        return (cs - med) / max(sig, 1e-30) + alpha * lv

    def refine(self, plan, volume, cand):
        """Exact FLCF over the local grid of each candidate's bin and its 27
        voxels. Returns (score[P], euler[P, 3], pos[P, 3] x y z)."""
        t0 = time.perf_counter()
        ub, grp = np.unique(cand[:, 3].astype(np.int64), return_inverse=True)
        grids = orient.local_grids(self.bins[ub], self.half, self.step)  # (A, G, 3)
        c4 = np.concatenate([cand[:, :3].astype(np.int32), grp[:, None].astype(np.int32)], 1)
        sc, ai, pos = plan.refine(volume, grids, c4, centroid=self.centroid)
        eul = grids[ai, grp]
        self.timings.update(refine_s=time.perf_counter() - t0)
        return sc, eul, pos


def _gather(a, xyz, vec=False):
    """Values of a map at rows x y z: a torch (z, y, x[, 3]) tensor (gathered
    on the device) or a numpy (x, y, z) / (3, x, y, z) F-order array."""
    if hasattr(a, "cpu"):
        import torch
        i = torch.as_tensor(xyz, device=a.device)
        return a[i[:, 2], i[:, 1], i[:, 0]].cpu().numpy()
    if vec:
        return np.asarray(a)[:, xyz[:, 0], xyz[:, 1], xyz[:, 2]].T
    return np.asarray(a)[xyz[:, 0], xyz[:, 1], xyz[:, 2]]


def _tiles(dims, tdims, halo):
    core = [int(t) - 2 * halo for t in tdims]
    return int(np.prod([-(-int(d) // c) for d, c in zip(dims, core)]))


def _sync(t):
    if hasattr(t, "device") and getattr(t.device, "type", "") == "cuda":
        import torch
        torch.cuda.synchronize()


def recall(ref_xyz, got_xyz, dist=3.0):
    """Fraction of reference positions with a found position within dist."""
    if len(ref_xyz) == 0:
        return 1.0
    if len(got_xyz) == 0:
        return 0.0
    from scipy.spatial import cKDTree
    d, _ = cKDTree(np.asarray(got_xyz, np.float64)).query(np.asarray(ref_xyz, np.float64))
    return float(np.mean(d <= dist))


# ---------------------------------------------------------------- coarse-to-fine pruned search
class PrunedSearch:
    """Pruned search without scanning every orientation over the whole
    tomogram: the exhaustive machinery scans `n_coarse` orientations spread
    evenly over SO(3) (orient.super_fibonacci), whose best-score map locates
    the particles; local maxima of that map (greedy suppression within
    `nms_radius`, the `K` best above `threshold`) are refined exactly in
    `levels` rounds, each scoring a +-half grid of `step`-degree offsets on
    every Euler angle around the best orientation so far, at the 27 voxels
    around the best position so far. Scores are the exact FLCF of the search
    (FP32 correlation, FP64 statistics). An optional `rank(volume, xyz)`
    callable (e.g. the proposal network's particle logit at the candidates)
    re-orders the local maxima before the K best are kept.

    With `sigmas` set, the candidate threshold comes from the coarse map
    itself: max(threshold, median + sigmas * 1.4826 MAD) of its best scores
    (a robust noise level; particles are a small fraction of the voxels), and
    K is only a cap on memory. Nothing then depends on how many particles the
    tomogram holds."""

    def __init__(self, n_coarse=512, K=400, threshold=0.05, nms_radius=12.0, levels=((10.0, 10.0), (5.0, 5.0)),
                 bin=0, sigmas=None):
        self.n_coarse, self.K, self.threshold, self.nms_radius = int(n_coarse), int(K), threshold, nms_radius
        self.sigmas = None if sigmas is None else float(sigmas)
        self.noise = None
        self.levels = tuple((float(h), float(s)) for h, s in levels)
        self.orientations = orient.super_fibonacci(self.n_coarse).astype(np.float32)
        self.bin = int(bin)
        self._binned = None
        self.timings = {}

    def set_binned_template(self, T, M, W=None):
        """Template for a coarse scan on tom_bin(volume, bin): tom_bin of the
        template and of the mask (a binary mask stays binary: binned values
        above 0.5), and the central (n >> bin)^3 of the centred weight W,
        i.e. the same frequencies on the binned grid."""
        b = self.bin
        Tb = np.asarray(api.tom_bin(np.asarray(T, np.float32), b), np.float32)
        M = np.asarray(M, np.float32)
        Mb = np.asarray(api.tom_bin(M, b), np.float32)
        if np.all((M == 0) | (M == 1)):
            Mb = (Mb > 0.5).astype(np.float32)
        Wb = None
        if W is not None:
            Wn = np.asarray(W, np.float32)
            lo = [d // 2 - (d >> b) // 2 for d in Wn.shape]
            Wb = np.ascontiguousarray(Wn[lo[0]:lo[0] + (Wn.shape[0] >> b), lo[1]:lo[1] + (Wn.shape[1] >> b),
                                         lo[2]:lo[2] + (Wn.shape[2] >> b)])
        tn, mu, sg = api.prepare_template(Tb, Mb, Wb)
        self._binned = (tn, Mb, Wb, mu, sg)

    def coarse_binned(self, volume):
        """(score, orientation index) maps of the coarse scan of tom_bin(volume,
        bin), in binned voxels; includes the binned plan's precompute."""
        if self._binned is None:
            raise RuntimeError("set_binned_template first")
        tn, Mb, Wb, mu, sg = self._binned
        Vb = api.tom_bin(volume, self.bin)
        with api.SearchPlan(Vb, Mb.shape, Mb, batch=16) as pb:
            try:
                pb.set_tiles(16, 32)
            except api.TomHipError:
                pass
            pb.set_template(tn, Wb, mu, sg)
            pb.run(self.orientations, 0)
            return pb.finalize()

    def coarse(self, plan, tmpl, W, mu_t, sig_t):
        """(score map, orientation index map) of the coarse scan."""
        plan.set_template(tmpl, W, mu_t, sig_t)
        plan.run(self.orientations, 0)
        return plan.finalize()

    def noise_threshold(self, score):
        """median + sigmas * 1.4826 MAD of the coarse best-score map (at most
        2^22 voxels sampled with a fixed stride); sets self.noise."""
        med, sig = self.noise_of(score)
        return med + self.sigmas * sig

    def noise_of(self, score):
        """(median, 1.4826 MAD) of the coarse best-score map (at most 2^22
        voxels sampled with a fixed stride), also kept in self.noise."""
        if hasattr(score, "cpu"):
            import torch
            flat = score.reshape(-1)
            step = max(1, flat.numel() >> 22)
            x = flat[::step].double()
            med = float(torch.median(x))
            mad = float(torch.median((x - med).abs()))
        else:
            x = np.asarray(score, np.float64).ravel()
            x = x[::max(1, x.size >> 22)]
            med = float(np.median(x))
            mad = float(np.median(np.abs(x - med)))
        sig = 1.4826 * mad
        self.noise = (med, sig)
        return self.noise

    def candidates(self, score, index, rank=None, volume=None):
        """(P, 3) x y z and (P, 3) coarse Euler angles of the kept maxima."""
        thr = float(self.threshold)
        if self.sigmas is not None:
            thr = max(thr, self.noise_threshold(score))
        self.candidate_threshold = thr
        lm = api.local_maxima(score, 2, thr)
        if len(lm) == 0:
            return np.zeros((0, 3), np.int64), np.zeros((0, 3), np.float32)
        val = lm[:, 3] if rank is None else np.asarray(rank(volume, lm[:, :3].astype(np.int64)), np.float64)
        keep = greedy_nms(lm[:, :3], val, self.nms_radius / (1 << self.bin), self.K)
        xyz = lm[keep, :3].astype(np.int64)
        idx = _gather(index, xyz).astype(np.int64)
        return xyz, self.orientations[idx]

    def refine(self, plan, volume, xyz, eul):
        """Levels of exact local refinement; returns (score[P], euler[P, 3],
        pos[P, 3] x y z)."""
        P = len(xyz)
        if P == 0:
            return np.zeros(0, np.float32), np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32)
        pos = np.asarray(xyz, np.int32)
        eul = np.asarray(eul, np.float32)
        grp = np.arange(P, dtype=np.int32)[:, None]
        sc = None
        for half, step in self.levels:
            grids = orient.local_grids(eul, half, step)  # (A, P, 3)
            sc, ai, pos = plan.refine(volume, grids, np.concatenate([pos, grp], 1))
            eul = grids[ai, np.arange(P)]
        return sc, eul, pos

    def run(self, plan, volume, tmpl, W, mu_t, sig_t, rank=None, proposer=None, reranker=None):
        """proposer (optional): a Surrogate whose network takes the coarse
        score map as its second input channel; its candidates replace the
        local maxima of the coarse map (orientations still come from the
        coarse scan at each candidate). reranker (optional): a Surrogate that
        orders the coarse map's own local maxima (Surrogate.rerank) before the
        suppression keeps K of them."""
        t0 = time.perf_counter()
        if self.bin:
            score, index = self.coarse_binned(volume)
            plan.set_template(tmpl, W, mu_t, sig_t)  # the refinement scores at full resolution
        else:
            score, index = self.coarse(plan, tmpl, W, mu_t, sig_t)
        _sync(score)
        t1 = time.perf_counter()
        if self.bin:
            # binned peak i <-> template centre at i * 2^bin (the binned voxel
            # i starts at full-resolution voxel i * 2^bin)
            xyz, eul = self.candidates(score, index, rank, volume)
            dims = np.asarray(tuple(reversed(volume.shape)) if hasattr(volume, "cpu") else np.shape(volume))
            xyz = np.minimum(xyz << self.bin, dims - 1)
        elif reranker is not None:
            self.noise_of(score)  # (median, sigma) of the coarse map for the robust z
            xyz, eul = self.candidates(score, index, lambda vol, c: reranker.rerank(vol, score, c, self.noise),
                                       volume)
        elif proposer is None:
            xyz, eul = self.candidates(score, index, rank, volume)
        else:
            cand = proposer.propose(volume, K=self.K, nms_radius=self.nms_radius, aux=score)
            xyz = cand[:, :3].astype(np.int64)
            eul = self.orientations[_gather(index, xyz).astype(np.int64)]
        t2 = time.perf_counter()
        sc, eul, pos = self.refine(plan, volume, xyz, eul)
        t3 = time.perf_counter()
        self.timings = dict(coarse_s=t1 - t0, candidates_s=t2 - t1, refine_s=t3 - t2)
        self.n_candidates = len(xyz)
        return sc, eul, pos
