"""Tiled and chunked searches with resumable state: the job model of the
reference picker (Oscar3/matlabPicker/overlap/tom_os3_split.m,
results/tom_os3_collectResults.m, results/tom_os3_checkMissingResults.m,
search/tom_os3_createJobList.m) laid out for one GPU with 288 GB of HBM.

A job is (z-slab, angle chunk). Slabs cut the tomogram along z only: the x-y
plane stays whole, so every slab keeps the volume's x and y FFT lengths, and
each slab's z extent (its core plus a halo of floor(n/2) + 1 planes on each
side, clipped to the volume, rounded up to a supported FFT length) is what
makes a core voxel's score independent of the cut. Each finished job writes its
maps to `state_dir/job-<slab>-<chunk>.npz`; `missing_jobs` lists the jobs
without a file (tom_os3_checkMissingResults), so an interrupted run resumes
where it stopped; `collect` merges every file into whole-volume maps with the
bestPeak rule (a voxel takes a job's value when it is strictly better; angle
indices are offset by the chunk start, tom_os3_collectResults.m:61-75).
"""
from __future__ import annotations

import glob
import os
from dataclasses import dataclass

import numpy as np

# z lengths the search supports (2^a 3^b, and 2^a 3^b 5 / 2^a 25 with radix-5 stages)
_LENGTHS = (16, 24, 32, 40, 48, 64, 80, 96, 120, 128, 160, 192, 200, 240, 256, 320, 384, 400, 480, 512, 640, 768,
            800, 960, 1024)


@dataclass(frozen=True)
class Job:
    slab: int
    chunk: int
    z0: int       # first plane of the extended slab in the volume
    z1: int       # one past its last plane
    core0: int    # first plane of the slab's core (owned voxels)
    core1: int
    a0: int       # first angle index of the chunk
    a1: int


def _fft_len_at_least(k):
    for L in _LENGTHS:
        if L >= k:
            return L
    raise ValueError(f"no supported FFT length >= {k}")


def make_jobs(Z, n, core, A, chunk):
    """Jobs for a volume of Z planes, template edge n, slab cores of `core`
    planes and angle chunks of `chunk` orientations (A in total)."""
    if core < 1 or chunk < 1:
        raise ValueError("core and chunk must be positive")
    halo = n // 2 + 1
    jobs = []
    for s, c0 in enumerate(range(0, Z, core)):
        c1 = min(c0 + core, Z)
        if c0 == 0 and c1 == Z:
            z0, z1 = 0, Z
        else:
            z0, z1 = max(c0 - halo, 0), min(c1 + halo, Z)
            ext = _fft_len_at_least(z1 - z0)
            if ext >= Z:
                z0, z1 = 0, Z
            else:
                # grow to a supported length, inside the volume
                grow = ext - (z1 - z0)
                z0 = max(z0 - grow // 2, 0)
                z1 = z0 + ext
                if z1 > Z:
                    z1 = Z
                    z0 = Z - ext
        for k, a0 in enumerate(range(0, A, chunk)):
            jobs.append(Job(s, k, z0, z1, c0, c1, a0, min(a0 + chunk, A)))
    return jobs


def _path(state_dir, job):
    return os.path.join(state_dir, f"job-{job.slab:04d}-{job.chunk:04d}.npz")


def missing_jobs(state_dir, jobs):
    """Jobs without a result file."""
    return [j for j in jobs if not os.path.exists(_path(state_dir, j))]


def run_jobs(volume_zyx, tmpl, mask, W, mu_t, sig_t, angles, jobs, state_dir, batch=4, stop_after=None, log=None):
    """Run the jobs that have no result yet. volume_zyx: torch tensor (z, y, x)
    on the GPU; each job scans angles[a0:a1] over its slab and saves the
    core planes of its (score, angle) maps. stop_after: run at most this many
    jobs (an interrupted run)."""
    from . import api
    os.makedirs(state_dir, exist_ok=True)
    todo = missing_jobs(state_dir, jobs)
    if stop_after is not None:
        todo = todo[:stop_after]
    n3 = tuple(reversed(tmpl.shape)) if api._is_torch(tmpl) else np.shape(tmpl)
    for j in todo:
        sub = volume_zyx[j.z0:j.z1].contiguous()
        with api.SearchPlan(sub, n3, mask, batch=batch) as p:
            p.set_template(tmpl, W, mu_t, sig_t)
            p.run(np.asarray(angles[j.a0:j.a1], np.float32), 0)
            sc, an = p.finalize()
        lo, hi = j.core0 - j.z0, j.core1 - j.z0
        sc = sc[lo:hi].cpu().numpy()
        an = an[lo:hi].cpu().numpy()
        tmp = _path(state_dir, j) + ".tmp.npz"
        np.savez(tmp, score=sc, angle=an, job=np.asarray([j.slab, j.chunk, j.core0, j.core1, j.a0, j.a1]))
        os.replace(tmp, _path(state_dir, j))
        if log:
            log(f"[jobs] slab {j.slab} planes {j.z0}:{j.z1} angles {j.a0}:{j.a1} done")


def collect(state_dir, shape_zyx):
    """Merge every job file into (score, angle) maps (z, y, x): a voxel takes a
    job's value where it is strictly better; ties keep the lower angle index,
    as one uninterrupted search over the whole list would."""
    score = np.full(shape_zyx, -np.inf, np.float32)
    angle = np.full(shape_zyx, -1, np.int64)
    files = sorted(glob.glob(os.path.join(state_dir, "job-*.npz")))
    if not files:
        raise FileNotFoundError(f"no job results in {state_dir}")
    for f in files:
        with np.load(f, allow_pickle=False) as d:
            slab, chunk, c0, c1, a0, a1 = (int(v) for v in d["job"])
            s, a = d["score"], d["angle"].astype(np.int64) + a0
        cur_s, cur_a = score[c0:c1], angle[c0:c1]
        take = (s > cur_s) | ((s == cur_s) & (a < cur_a))
        cur_s[take] = s[take]
        cur_a[take] = a[take]
    return score, angle.astype(np.int32)
