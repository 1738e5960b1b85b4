"""Tiled / chunked searches with resumable state (tom_amd.jobs)."""
import numpy as np
import pytest

from tom_amd import jobs as J


@pytest.mark.parametrize("Z,n,core,A,chunk", [(96, 16, 32, 30, 12), (256, 48, 100, 7488, 2000), (64, 16, 64, 5, 5)])
def test_job_list_invariants(Z, n, core, A, chunk):
    jl = J.make_jobs(Z, n, core, A, chunk)
    halo = n // 2 + 1
    slabs = {}
    for j in jl:
        slabs.setdefault(j.slab, j)
        assert 0 <= j.z0 <= j.core0 < j.core1 <= j.z1 <= Z
        # halo on both sides unless clipped by the volume
        assert j.core0 - j.z0 >= min(halo, j.core0) and j.z1 - j.core1 >= min(halo, Z - j.core1)
        ext = j.z1 - j.z0
        assert ext == Z or ext in J._LENGTHS
    cores = sorted((s.core0, s.core1) for s in slabs.values())
    assert cores[0][0] == 0 and cores[-1][1] == Z
    assert all(a[1] == b[0] for a, b in zip(cores, cores[1:]))
    for s in slabs:
        ch = sorted((j.a0, j.a1) for j in jl if j.slab == s)
        assert ch[0][0] == 0 and ch[-1][1] == A and all(a[1] == b[0] for a, b in zip(ch, ch[1:]))


def test_collect_and_missing(tmp_path):
    jl = J.make_jobs(8, 4, 4, 6, 3)   # 2 slabs x 2 chunks
    assert len(J.missing_jobs(tmp_path, jl)) == 4
    rng = np.random.default_rng(0)
    full = {}
    for j in jl[:3]:
        s = rng.random((j.core1 - j.core0, 2, 3)).astype(np.float32)
        a = rng.integers(0, j.a1 - j.a0, s.shape).astype(np.int32)
        if j.slab == 0 and j.chunk == 1:
            s[0, 0, 0] = full[(0, 0)][0][0, 0, 0]  # a tie: the lower angle index wins
        full[(j.slab, j.chunk)] = (s, a)
        np.savez(tmp_path / f"job-{j.slab:04d}-{j.chunk:04d}.npz", score=s, angle=a,
                 job=np.asarray([j.slab, j.chunk, j.core0, j.core1, j.a0, j.a1]))
    assert J.missing_jobs(tmp_path, jl) == [jl[3]]
    sc, an = J.collect(tmp_path, (8, 2, 3))
    s0, a0 = full[(0, 0)]
    s1, a1 = full[(0, 1)]
    better = s1 > s0
    np.testing.assert_array_equal(sc[:4], np.where(better, s1, s0))
    np.testing.assert_array_equal(an[:4], np.where(better, a1 + 3, a0))
    assert an[0, 0, 0] == a0[0, 0, 0]
    s2, a2 = full[(1, 0)]
    np.testing.assert_array_equal(sc[4:], s2)
    np.testing.assert_array_equal(an[4:], a2)


@pytest.mark.gpu
def test_tiled_chunked_resumed_search_matches_whole(api, oracle, tmp_path):
    """Slabs x angle chunks, interrupted after 4 jobs and resumed, give the
    whole-volume maps on every plane at least one halo away from the z
    borders (scores within 1e-5; angles equal where the oracle's margin
    between the best and second-best angle exceeds 1e-4)."""
    import torch
    print("tiled search", flush=True)
    rng = np.random.default_rng(50)
    N = (64, 64, 96)
    V = rng.standard_normal(N).astype(np.float32)
    T = rng.standard_normal((16, 16, 16)).astype(np.float32)
    M = oracle.sphere_mask_binary(16, 7)
    angles = oracle.angle_grid(60, 60, 60)[:30].astype(np.float32)
    tn, mu, sg = api.prepare_template(T, M)
    Vt = torch.from_numpy(np.ascontiguousarray(V.transpose(2, 1, 0))).cuda()
    jl = J.make_jobs(N[2], 16, 32, len(angles), 12)
    J.run_jobs(Vt, tn, M, None, mu, sg, angles, jl, str(tmp_path), stop_after=4)
    assert len(J.missing_jobs(str(tmp_path), jl)) == len(jl) - 4
    J.run_jobs(Vt, tn, M, None, mu, sg, angles, jl, str(tmp_path))
    sc, an = J.collect(str(tmp_path), (N[2], N[1], N[0]))
    ref_s, ref_a = api.tom_os3_findTemplate(V, T, M, angles)
    ref_s, ref_a = ref_s.transpose(2, 1, 0), ref_a.transpose(2, 1, 0) - 1
    _, _, allm = oracle.find_template(V, T, M, angles, return_all=True)
    srt = np.sort(allm, axis=0)
    margin = (srt[-1] - srt[-2]).transpose(2, 1, 0)
    z = slice(9, N[2] - 9)
    err = np.abs(sc[z] - ref_s[z]).max()
    sel = margin[z] > 1e-4
    mism = int((an[z][sel] != ref_a[z][sel]).sum())
    print(f"  max |tiled - whole| = {err:.2e}; angle mismatches {mism} of {int(sel.sum())}", flush=True)
    assert err < 1e-5 and mism == 0
