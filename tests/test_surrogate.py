"""Host logic of the learned proposal (labels, greedy suppression, recall) and
the GPU candidate extraction against brute force."""
import os

import numpy as np
import pytest

from tom_amd import orient, surrogate as S


def test_greedy_nms_brute_force():
    rng = np.random.default_rng(0)
    pts = rng.uniform(0, 50, (400, 3))
    val = rng.random(400)
    got = S.greedy_nms(pts, val, 6.0, 1000)
    taken = []
    for i in np.argsort(-val, kind="stable"):
        if all(np.linalg.norm(pts[i] - pts[j]) >= 6.0 for j in taken):
            taken.append(i)
    np.testing.assert_array_equal(got, taken)
    np.testing.assert_array_equal(S.greedy_nms(pts, val, 6.0, 7), taken[:7])


def test_label_volumes():
    bins = orient.orientation_bins(64, 0)
    angles = np.asarray([[0, 0, 0], [10, 20, 30]], np.float32)
    peaks = np.array([[5, 6, 7, 0.5, 1], [0, 0, 0, 0.4, 0]], float)
    tp, tb = S.label_volumes((12, 14, 16), peaks, angles, bins, sigma=2.0, bin_radius=2.0)
    b = orient.assign_bins(angles, bins)
    assert tp[7, 6, 5] == 1 and tb[7, 6, 5] == b[1]
    assert abs(tp[7, 6, 7] - np.exp(-4 / 8)) < 1e-6 and tb[7, 6, 7] == b[1] and tb[7, 6, 8] == -1
    assert tp[7, 6, 11] == np.float32(np.exp(-36 / 8)) and tp[7, 6, 12] == 0  # cut at 3 sigma
    assert tp[0, 0, 0] == 1 and tb[0, 0, 0] == b[0]
    # the nearer peak wins where the two overlap
    assert tp[2, 2, 2] == np.float32(np.exp(-12 / 8))
    assert (tb >= 0).sum() == 33 + 11


def test_label_offsets_brute_force():
    """Offsets point from each voxel to the nearest peak (brute force over the
    peaks), in units of OFFSET_SCALE, within offset_radius only."""
    bins = orient.orientation_bins(64, 0)
    angles = np.asarray([[0, 0, 0], [10, 20, 30]], np.float32)
    peaks = np.array([[5, 6, 7, 0.5, 1], [12, 3, 2, 0.4, 0], [1, 13, 9, 0.35, 0]], float)
    shape = (12, 14, 16)
    tp0, tb0 = S.label_volumes(shape, peaks, angles, bins, sigma=2.0, bin_radius=2.0)
    tp, tb, to = S.label_volumes(shape, peaks, angles, bins, sigma=2.0, bin_radius=2.0, offset_radius=5.0)
    np.testing.assert_array_equal(tp, tp0)
    np.testing.assert_array_equal(tb, tb0)
    Z, Y, X = shape
    c = peaks[:, :3]
    for z in range(Z):
        for y in range(Y):
            for x in range(X):
                d2 = ((c - [x, y, z]) ** 2).sum(1)
                j = int(np.argmin(d2))
                if d2[j] <= 25 and (d2 == d2[j]).sum() == 1:
                    np.testing.assert_allclose(to[z, y, x, :3], (c[j] - [x, y, z]) / S.OFFSET_SCALE, atol=1e-7)
                    assert to[z, y, x, 3] == 1
                elif d2[j] > 25:
                    assert not to[z, y, x].any()


def test_recall():
    ref = np.array([[0, 0, 0], [10, 10, 10], [20, 20, 20]], float)
    got = np.array([[1, 1, 1], [10, 10, 14]], float)
    assert S.recall(ref, got, 3.0) == pytest.approx(1 / 3)
    assert S.recall(ref, got, 4.0) == pytest.approx(2 / 3)
    assert S.recall(np.zeros((0, 3)), got) == 1.0


@pytest.mark.gpu
def test_local_maxima_brute_force(api):
    print("local maxima", flush=True)
    rng = np.random.default_rng(5)
    a = rng.random((20, 18, 16)).astype(np.float32)
    a[3, 4, 5] = a[3, 4, 6] = 2.0  # a plateau: the lower index wins
    lm = api.local_maxima(a, 2, 0.5)
    X, Y, Z = a.shape
    ref = []
    for x in range(X):
        for y in range(Y):
            for z in range(Z):
                v = a[x, y, z]
                if not v > 0.5:
                    continue
                ok = True
                for dz in range(-2, 3):
                    for dy in range(-2, 3):
                        for dx in range(-2, 3):
                            xx, yy, zz = x + dx, y + dy, z + dz
                            if 0 <= xx < X and 0 <= yy < Y and 0 <= zz < Z:
                                w = a[xx, yy, zz]
                                li, lj = (zz * Y + yy) * X + xx, (z * Y + y) * X + x
                                if w > v or (w == v and li < lj):
                                    ok = False
                if ok:
                    ref.append((x, y, z, v))
    ref = sorted(ref, key=lambda r: (-r[3], (r[2] * Y + r[1]) * X + r[0]))
    np.testing.assert_array_equal(lm, np.asarray(ref, np.float32))
    assert (3, 4, 5, 2.0) in [tuple(r) for r in lm.tolist()]


def test_dense_targets():
    bins = orient.orientation_bins(64, 0)
    angles = np.asarray([[0, 0, 0], [10, 20, 30], [200, 100, 60]], np.float32)
    sc = np.array([[[0.0, 0.1, 0.3], [0.5, 0.7, 0.26]]], np.float32)
    an = np.array([[[0, 1, 2], [1, 2, 0]]], np.int32)
    tp, tb = S.dense_targets(sc, an, angles, bins)
    np.testing.assert_allclose(tp, [[[0, 0, 0.5], [1, 1, 0.4]]], atol=1e-6)
    b = orient.assign_bins(angles, bins)
    np.testing.assert_array_equal(tb, [[[-1, -1, b[2]], [b[1], b[2], b[0]]]])


def test_spherical_average():
    rng = np.random.default_rng(7)
    T = rng.random((16, 16, 16)).astype(np.float32)
    A = S.spherical_average(T)
    g = np.meshgrid(*[np.arange(16) - 8] * 3, indexing="ij")
    r = np.rint(np.sqrt(sum(x * x for x in g))).astype(int)
    for k in (0, 1, 5, 8):
        sel = r == k
        assert np.allclose(A[sel], T[sel].mean(), atol=1e-6)
    # invariant under an exchange of axes (a distance-preserving map about the centre)
    np.testing.assert_allclose(S.spherical_average(T.transpose(1, 2, 0)), A, atol=1e-6)


@pytest.mark.gpu
def test_snap_maxima_brute_force(api):
    print("snap maxima", flush=True)
    rng = np.random.default_rng(9)
    a = rng.random((20, 18, 16)).astype(np.float32)
    a[5, 5, 5] = a[6, 5, 5] = 3.0  # tie: the lower linear index (x fastest) wins
    cand = np.array([[4, 4, 4], [0, 0, 0], [19, 17, 15], [10, 9, 8]], np.int32)
    got = api.snap_maxima(a, cand, 2)
    for c, g in zip(cand, got):
        lo = np.maximum(c - 2, 0)
        hi = np.minimum(c + 3, a.shape)
        sub = a[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
        best = None
        for z in range(sub.shape[2]):
            for y in range(sub.shape[1]):
                for x in range(sub.shape[0]):
                    if best is None or sub[x, y, z] > best[0]:
                        best = (sub[x, y, z], (x + lo[0], y + lo[1], z + lo[2]))
        assert tuple(g) == best[1]
    assert tuple(got[0]) == (5, 5, 5)


@pytest.mark.gpu
def test_pruned_search_finds_exhaustive_peaks_with_exact_scores(api, oracle):
    """Coarse-to-fine search (1,024 coarse orientations, three refinement
    levels; coarse scan at full resolution and on tom_bin-ned volume and
    template) finds every exhaustive peak above 0.3 of a seeded volume with
    three 24^3 particles planted at orientations of the exhaustive 30-degree
    grid, and each reported (score, orientation, position) of the full-resolution
    run equals the FP64 oracle's FLCF of that orientation at that voxel
    within 1e-4."""
    print("pruned search", flush=True)
    from tom_amd import synth
    rng = np.random.default_rng(41)
    N = (128, 96, 64)
    n = 24
    V = rng.standard_normal(N).astype(np.float32)
    T = synth.phantom(np.random.default_rng(42), n, 10, 1.5, 8)
    for corner, e in (((10, 12, 8), (30, 60, 90)), ((60, 50, 28), (210, 120, 120)), ((40, 66, 36), (90, 300, 30))):
        x, y, z = corner
        V[x:x + n, y:y + n, z:z + n] += 3.0 * oracle.rotate(T, e)
    M = oracle.sphere_mask_binary(n, 10)
    tn, mu, sg = api.prepare_template(T, M)
    angles = oracle.angle_grid(30, 30, 30).astype(np.float32)
    with api.SearchPlan(V, T.shape, M, batch=4) as plan:
        plan.set_template(tn, None, mu, sg)
        plan.run(angles, 0)
        score, ang = plan.finalize()
        ref = S.exhaustive_peaks(score, ang, n, 0.3)
        levels = ((30.0, 15.0), (15.0, 7.5), (7.5, 5.0))
        ps = S.PrunedSearch(n_coarse=1024, K=30, threshold=0.05, nms_radius=6.0, levels=levels)
        sc, eul, pos = ps.run(plan, V, tn, None, mu, sg)
        # the same with the coarse scan on tom_bin(volume) and tom_bin(template)
        pb = S.PrunedSearch(n_coarse=1024, K=30, threshold=0.05, nms_radius=6.0, levels=levels, bin=1)
        pb.set_binned_template(T, M)
        scb, _, posb = pb.run(plan, V, tn, None, mu, sg)
    print(f"  {len(ref)} exhaustive peaks > 0.3, {ps.n_candidates} candidates, timings {ps.timings}", flush=True)
    for name, a, b in (("full", sc, pos), ("binned", scb, posb)):
        for r in ref:
            d = np.linalg.norm(b - r[:3], axis=1)
            j = int(np.argmin(d))
            print(f"  {name}: exhaustive {r[:4]} nearest refined {b[j]} score {a[j]:.4f}", flush=True)
    assert len(ref) >= 3
    assert S.recall(ref[:, :3], pos[sc > 0.3]) == 1.0
    assert S.recall(ref[:, :3], posb[scb > 0.3]) == 1.0
    Vd = V.astype(np.float64)
    Md = M.astype(np.float64)
    m = float((Md > 0).sum())
    Tn, mu_T, sig_T = oracle.prepare_template(T, Md)
    mu_I, sig_I = oracle.local_stats(Vd, Md, m)
    FV = oracle._fftn(Vd)
    for i in np.argsort(-sc)[:3]:
        s64 = oracle.flcf_angle(Vd, Tn, Md, m, mu_T, sig_T, mu_I, sig_I, FV, eul[i])
        x, y, z = pos[i]
        print(f"  candidate {i}: gpu {sc[i]:.6f} fp64 {s64[x, y, z]:.6f}", flush=True)
        assert abs(sc[i] - s64[x, y, z]) < 1e-4


@pytest.mark.gpu
def test_train_script_resumes_across_runs(api, tmp_path):
    """train_surrogate with a state directory: a run stopped by its time
    limit during labelling, then one stopped after its first training step,
    then one without a limit, ends at the weights of an uninterrupted run
    (to the float-atomic noise of the weight gradients: 1 % of the distance
    the weights moved), and the picks came from the label cache."""
    import json
    from tom_amd import train_surrogate
    print("train_surrogate resume", flush=True)
    base = ["--tomos", "2", "--size", "128", "--particles", "6", "--steps", "12", "--crop", "32", "--coarse", "64",
            "--ckpt-every", "4", "--seed", "7"]
    st, out_r, out_s = str(tmp_path / "st"), str(tmp_path / "r.npz"), str(tmp_path / "s.npz")
    for expect_files in (["label-0000.npz"], ["label-0000.npz", "label-0001.npz", "train.npz"]):
        with pytest.raises(SystemExit) as e:
            train_surrogate.main(base + ["--state", st, "--time-limit", "0", "--out", out_r])
        assert e.value.code == 3
        assert sorted(os.listdir(st)) == expect_files
    z = np.load(os.path.join(st, "train.npz"), allow_pickle=False)
    assert int(z["step"]) == 1 and len(z["history"]) == 1
    train_surrogate.main(base + ["--state", st, "--out", out_r])
    z = np.load(os.path.join(st, "train.npz"), allow_pickle=False)
    assert int(z["step"]) == 12 and len(z["history"]) == 12
    train_surrogate.main(base + ["--out", out_s])
    r, s = np.load(out_r, allow_pickle=False), np.load(out_s, allow_pickle=False)
    assert json.loads(str(r["meta"]))["peaks_per_tomogram"] == json.loads(str(s["meta"]))["peaks_per_tomogram"]
    with api.Unet((32, 32, 32), nbin=64, seed=7, train=True) as net:
        theta0 = net.get_params()
    rel = np.linalg.norm(r["params"] - s["params"]) / np.linalg.norm(s["params"] - theta0)
    print(f"  |resumed - uninterrupted| / |moved| = {rel:.2e}", flush=True)
    assert rel <= 1e-2
    # a state directory of other settings is refused
    with pytest.raises(SystemExit) as e:
        train_surrogate.main(base[:-1] + ["8", "--state", st, "--out", out_r])
    assert "other settings" in str(e.value.code)


@pytest.mark.gpu
def test_rerank_is_robust_z_plus_logit_near_the_maximum(api):
    """Surrogate.rerank = (coarse - median) / sigma + 2 x the largest network
    logit in the 5^3 cube around each row (checked against the logit map of
    Unet.infer and a brute-force cube maximum, borders clipped)."""
    import torch
    print("rerank", flush=True)
    w = np.load(os.path.join(os.path.dirname(S.__file__), "weights", "unet_c2.npz"), allow_pickle=False)
    net = S.Surrogate(w["params"], w["bins"])
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    vol = torch.randn((40, 48, 64), generator=g, device="cuda")
    coarse = torch.rand((40, 48, 64), generator=g, device="cuda") * 0.3
    rng = np.random.default_rng(4)
    xyz = np.stack([rng.integers(0, 64, 50), rng.integers(0, 48, 50), rng.integers(0, 40, 50)], 1)
    xyz[:3] = [[0, 0, 0], [63, 47, 39], [1, 46, 0]]  # cubes clipped at the borders
    noise = (0.15, 0.08)
    got = net.rerank(vol, coarse, xyz, noise)
    unet, halo = net._network((64, 48, 40))
    lg = unet.infer(vol, halo, aux=coarse, logit=True)[0].cpu().numpy()  # (z, y, x)
    cs = coarse.cpu().numpy()
    want = []
    for x, y, z in xyz:
        cube = lg[max(z - 2, 0):z + 3, max(y - 2, 0):y + 3, max(x - 2, 0):x + 3]
        want.append((cs[z, y, x] - noise[0]) / noise[1] + 2.0 * cube.max())
    np.testing.assert_allclose(got, np.asarray(want), rtol=1e-6, atol=1e-6)
