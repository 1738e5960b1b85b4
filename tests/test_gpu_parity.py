"""GPU parity: libtomhip.so through its C ABI (tom_amd.api, ctypes) against the
oracle and the golden vectors. Tolerances are written per test:
rotation, peaks, masks: bit-exact; scores: 1e-4 absolute against the FP64
oracle; best-angle indices: identical wherever the oracle's margin between the
best and second-best angle exceeds 1e-4."""
import hashlib
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "rot3d_golden.npz")

SCORE_TOL = 1e-4
MARGIN = 1e-4


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def api():
    from tom_amd import api
    api.context()
    return api


def test_library_is_native(api):
    import tom_amd
    assert os.path.exists(tom_amd.LIB_PATH)
    assert api.context().h


# ---------------------------------------------------------------- rotation
def test_rotate_golden_bit_exact(api):
    say("rotate golden")
    g = np.load(GOLD)
    vols = {24: g["vol24"], 25: g["vol25"], 48: g["vol48"], 2030: g["vol2030"]}
    bad = []
    for i in range(len(g["case_sha256"])):
        v = vols[int(g["case_vol"][i])]
        ne = int(g["case_neuler"][i])
        e = g["case_euler"][i][:ne].astype(np.float32)
        if ne == 9:
            e = e.reshape(3, 3, order="F")
        o = api.tom_rotate(v, e, "linear", g["case_center"][i])
        if hashlib.sha256(np.asfortranarray(o).tobytes(order="F")).hexdigest() != g["case_sha256"][i]:
            bad.append(i)
    assert not bad, f"{len(bad)} golden rotations differ, first {bad[:5]}"


def test_rotate2d_matches_oracle(api, oracle):
    """tom_rotate of a 2D image (rot2d) through the C ABI, bit-exact against
    the oracle's restatement: integer and fractional pivots (the pixel
    coordinates accumulate by float increments), the matrix form, a 1-pixel
    wide image."""
    say("rotate 2d")
    rng = np.random.default_rng(12)
    img = rng.standard_normal((45, 38)).astype(np.float32)
    for phi, c in ((0.0, None), (30.0, None), (-112.7, (22.5, 19.0)), (355.0, (7.3, 30.9)), (90.0, (0, 0))):
        got = api.tom_rotate(img, phi, "linear", c)
        np.testing.assert_array_equal(got, oracle.rotate2d(img, phi, c))
    m = np.array([[0.8, -0.6], [0.6, 0.8]], np.float32)  # rm00 rm10 ; rm01 rm11 column-major
    np.testing.assert_array_equal(api.tom_rotate(img, m, "linear", (20, 17)),
                                  oracle.rotate2d(img, m.ravel(order="F"), (20, 17)))
    line = rng.standard_normal((1, 25)).astype(np.float32)
    np.testing.assert_array_equal(api.tom_rotate(line, 10.0), oracle.rotate2d(line, 10.0))


def test_rotate_batch_matches_single_and_oracle(api, oracle):
    say("rotate batch")
    rng = np.random.default_rng(0)
    v = rng.standard_normal((48, 48, 48)).astype(np.float32)
    M = oracle.sphere_mask_binary(48, 22)
    ang = np.concatenate([oracle.angle_grid(30, 30, 30)[::37], rng.uniform(0, 360, (8, 3))]).astype(np.float32)
    out = api.tom_rotate_batch(v, ang, mask=M)
    for a in range(ang.shape[0]):
        ref = oracle.rotate(v, ang[a]) * M
        np.testing.assert_array_equal(out[..., a], ref)


# ---------------------------------------------------------------- peaks, masks
def test_peak_first_max(api, oracle):
    say("peak")
    rng = np.random.default_rng(1)
    for shape in [(7, 6, 5), (64, 64, 64), (130, 3, 17)]:
        a = rng.integers(0, 50, size=shape).astype(np.float32)
        v, p = api.tom_peakinc(a)
        vo, po = oracle.find_max(a)
        assert v == vo and list(p) == list(po)
    v, p = api.tom_peakinc(np.full((4, 4, 4), -2e5, np.float32))
    assert v == -99999.0
    A = np.array([[1, 2, 3, 4, 17], [5, 6, 7, 8, 18], [10, 9, 40, 6, 20], [12, 16, 17, 20, 30],
                  [5, 23, 35, 12, 6]], np.float32)
    c, val = api.tom_peak(A)
    assert list(c) == [3, 3] and val == 40  # Analysis/tom_peak.m:27-35


def test_masks_and_weights(api, oracle):
    say("masks")
    m, cnt = api.tom_os3_sphereMask((24, 24, 24))
    mo, co = oracle.sphere_mask_os3(24)
    np.testing.assert_array_equal(m, mo)
    assert cnt == co
    g = api.tom_spheremask_gauss((24, 24, 24), 10, 2)
    np.testing.assert_allclose(g, oracle.sphere_mask_gauss(24, 10, 2), rtol=2e-7, atol=0)
    b = api.tom_spheremask_gauss((48, 48, 48), 22, 0)
    np.testing.assert_array_equal(b, oracle.sphere_mask_binary(48, 22))
    w = api.weight((24, 24, 24), 30, 2, 20)
    np.testing.assert_array_equal(w, oracle.wedge((24, 24, 24), 30) * oracle.bandpass_weight((24, 24, 24), 2, 20))


def test_prepare_template(api, oracle):
    say("prepare template")
    rng = np.random.default_rng(2)
    T = rng.standard_normal((24, 24, 24)).astype(np.float32)
    M = oracle.sphere_mask_gauss(24, 10, 2)
    W = oracle.wedge((24, 24, 24), 30) * oracle.bandpass_weight((24, 24, 24), 2, 20)
    for w in (None, W):
        tn, mu, sg = api.prepare_template(T, M, None if w is None else w.astype(np.float32))
        Tn, mu_o, sg_o = oracle.prepare_template(T, M, w)
        np.testing.assert_allclose(tn, Tn, atol=2e-6)
        assert abs(mu - mu_o) < 1e-9 and abs(sg - sg_o) < 1e-9


# ---------------------------------------------------------------- FLCF
def _phantom(rng, n, k, sigma, R):
    x = np.arange(n) - n // 2
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    T = np.zeros((n, n, n))
    for _ in range(k):
        while True:
            c = rng.uniform(-R, R, 3)
            if np.linalg.norm(c) <= R:
                break
        T += rng.uniform(0.5, 1) * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / (2 * sigma ** 2))
    return T.astype(np.float32)


@pytest.mark.parametrize("N,n,soft", [((64, 64, 64), 24, True), ((32, 48, 24), 12, False)])
def test_os3_corr_single(api, oracle, N, n, soft):
    """tom_os3_corr FLCF map (no rotation) against the FP64 oracle."""
    say(f"os3_corr N={N} n={n}")
    rng = np.random.default_rng(sum(N))
    V = rng.standard_normal(N).astype(np.float32)
    T = _phantom(rng, n, 5, 1.5, n // 3)
    M = oracle.sphere_mask_gauss(n, n / 2 - 2, 2) if soft else oracle.sphere_mask_binary(n, n / 2 - 1)
    out = api.tom_os3_corr(V, T, M)
    m = float((M > 0).sum())
    mu_T, sig_T = oracle.norm_under_mask(T, M)
    mu_I, sig_I = oracle.local_stats(V, M, m)
    ref = oracle.flcf(V, T.astype(np.float64) * M, m, mu_T, sig_T, mu_I, sig_I)
    err = np.abs(out - ref).max()
    say(f"  max |score - fp64| = {err:.3e}")
    assert err < SCORE_TOL


def _check_search(ccc, ang, allm, tag):
    """ccc/ang (GPU, ang 1-based) against per-angle FP64 oracle maps."""
    best = allm.max(axis=0)
    srt = np.sort(allm, axis=0)
    margin = srt[-1] - srt[-2] if allm.shape[0] > 1 else np.full(best.shape, np.inf)
    ref_ang = np.argmax(allm == best[None], axis=0) + 1
    err = np.abs(ccc - best).max()
    sel = margin > MARGIN
    mism = int(((ang != ref_ang) & sel).sum())
    say(f"  {tag}: max |score - fp64| = {err:.3e}; angle mismatches where margin>1e-4: {mism} of {int(sel.sum())}")
    assert err < SCORE_TOL
    assert mism == 0


def test_find_template_c1_subset(api, oracle):
    """Config-1 shaped search (64^3, 24^3 phantom, soft mask, wedge x bandpass)
    over 40 angles of the 30-degree grid."""
    say("findTemplate C1 subset")
    rng = np.random.default_rng(100)
    n, N = 24, (64, 64, 64)
    T = _phantom(rng, n, 6, 1.5, 8)
    V = rng.standard_normal(N).astype(np.float32)
    for c in [(20, 20, 20), (44, 40, 22), (30, 46, 44)]:
        e = rng.uniform(0, 360, 3)
        r = oracle.rotate(T, e)
        V[c[0] - 12:c[0] + 12, c[1] - 12:c[1] + 12, c[2] - 12:c[2] + 12] += 3 * r
    M = oracle.sphere_mask_gauss(n, 10, 2)
    W = (oracle.wedge((n, n, n), 30) * oracle.bandpass_weight((n, n, n), 2, 20)).astype(np.float32)
    angles = oracle.angle_grid(30, 30, 30)[::25]
    ccc, ang = api.tom_os3_findTemplate(V, T, M, angles, W=W, batch=3)
    _, _, allm = oracle.find_template(V, T, M, angles, W=W, return_all=True)
    _check_search(ccc, ang, allm, "C1 subset")


@pytest.mark.parametrize("N", [(160, 80, 40), (96, 120, 240), (800, 24, 200)])
def test_find_template_radix5_sizes(api, oracle, N):
    """Volumes with a factor 5 in their FFT lengths (radix-5 stages in every
    pass: X/2 = 80 / 400, Y = 80 / 120, Z = 40 / 240 / 200) against the FP64 oracle,
    with a wedge weight and planted copies."""
    say(f"findTemplate radix-5 sizes {N}")
    rng = np.random.default_rng(31)
    n = 16
    T = _phantom(rng, n, 4, 1.5, 5)
    V = rng.standard_normal(N).astype(np.float32)
    for c in [(20, min(20, N[1] // 2), 20), (N[0] - 24, N[1] // 2, N[2] // 2)]:
        r = oracle.rotate(T, rng.uniform(0, 360, 3))
        V[c[0] - 8:c[0] + 8, c[1] - 8:c[1] + 8, c[2] - 8:c[2] + 8] += 3 * r
    M = oracle.sphere_mask_binary(n, 7)
    W = oracle.wedge((n, n, n), 30).astype(np.float32)
    angles = oracle.angle_grid(60, 60, 60)[::7]
    ccc, ang = api.tom_os3_findTemplate(V, T, M, angles, W=W, batch=3)
    _, _, allm = oracle.find_template(V, T, M, angles, W=W, return_all=True)
    _check_search(ccc, ang, allm, f"radix-5 {N}")


def test_find_template_no_weight_binary_mask(api, oracle):
    say("findTemplate no weight")
    rng = np.random.default_rng(7)
    V = rng.standard_normal((32, 64, 48)).astype(np.float32)
    T = _phantom(rng, 16, 4, 1.5, 5)
    M = oracle.sphere_mask_binary(16, 7)
    angles = np.asarray([[0, 0, 0], [15, 30, 45], [90, 0, 90], [200, 10, 170], [0, 0, 0]], float)
    ccc, ang = api.tom_os3_findTemplate(V, T, M, angles, batch=2)
    _, _, allm = oracle.find_template(V, T, M, angles, return_all=True)
    _check_search(ccc, ang, allm, "no weight")
    assert not np.any(ang == 5)  # duplicate of angle 1 never wins (strict update)


def test_search_plan_batches_equal(api, oracle):
    """Batch size, split runs, pass tile widths, z-chunking and hipGraph
    replay versus eager launches do not change the result (ties included) on
    this volume."""
    say("plan batches")
    rng = np.random.default_rng(8)
    V = rng.standard_normal((64, 32, 32)).astype(np.float32)
    T = _phantom(rng, 16, 4, 1.5, 5)
    M = oracle.sphere_mask_binary(16, 7)
    ang = oracle.angle_grid(60, 60, 60)
    tn, mu, sg = api.prepare_template(T, M)
    outs = []
    for batch, split, graph, tiles, zc in [(1, 1, False, (16, 16), 0), (4, 1, True, (16, 16), 0),
                                           (4, 1, False, (8, 32), 8), (7, 3, True, (32, 8), 12),
                                           (3, 1, True, (32, 32), 5)]:
        with api.SearchPlan(V, T.shape, M, batch=batch) as p:
            p.set_graph(graph)
            p.set_tiles(*tiles)
            p.set_zchunk(zc)
            p.set_template(tn, None, mu, sg)
            step = int(np.ceil(len(ang) / split))
            for s0 in range(0, len(ang), step):
                p.run(ang[s0:s0 + step], s0)
            outs.append(p.finalize())
    # a fresh context whose first launches are a graph capture (twiddle
    # tables must exist before the capture starts)
    fresh = api.Context(0)
    with api.SearchPlan(V, T.shape, M, batch=4, ctx=fresh) as p:
        p.set_graph(True)
        p.set_template(tn, None, mu, sg)
        p.run(ang, 0)
        outs.append(p.finalize())
    for sc, an in outs[1:]:
        np.testing.assert_array_equal(sc, outs[0][0])
        np.testing.assert_array_equal(an, outs[0][1])


def test_oscar(api, oracle):
    say("oscar")
    rng = np.random.default_rng(9)
    V = rng.standard_normal((32, 32, 32)).astype(np.float32)
    T = _phantom(rng, 8, 3, 1.2, 2)
    ccf, phi, psi, the = api.tom_oscar(V, T, 0, 90, 90, 0, 0, 30, 0, 60, 60)
    ang = np.asarray([(p, s, t) for p in (0, 90) for s in (0,) for t in (0, 60)], float)
    ref, idx = oracle.oscar(V, T, ang)
    scale = np.abs(ref).max()
    assert np.abs(ccf - ref).max() < 1e-5 * scale
    allm = []
    for a in ang:
        c, _ = oracle.oscar(V, T, a[None])
        allm.append(c)
    allm = np.stack(allm)
    srt = np.sort(allm, axis=0)
    sel = (srt[-1] - srt[-2]) > 1e-4 * scale
    np.testing.assert_array_equal(phi[sel], ang[idx, 0][sel].astype(np.int16))
    np.testing.assert_array_equal(the[sel], ang[idx, 2][sel].astype(np.int16))


def test_refine_matches_exhaustive_maps(api, oracle):
    """tomhip_search_refine: for each candidate, the best FP64 score over its
    3x3x3 neighbourhood (wrapping at the borders) and its group's
    orientations; the orientation and voxel agree wherever the oracle's margin
    between the best and second-best (voxel, orientation) pair exceeds 1e-4."""
    say("refine")
    rng = np.random.default_rng(21)
    N = (64, 48, 32)
    n = 16
    V = rng.standard_normal(N).astype(np.float32)
    T = _phantom(rng, n, 4, 1.5, 5)
    V[20:36, 10:26, 8:24] += 2 * oracle.rotate(T, [30, 60, 90])
    M = oracle.sphere_mask_gauss(n, 6, 1.5)
    W = (oracle.wedge((n, n, n), 30) * oracle.bandpass_weight((n, n, n), 1, 7)).astype(np.float32)
    angles = np.concatenate([oracle.angle_grid(60, 60, 60)[::7], [[30, 60, 90]]])[:20]
    G, A = 2, 10
    ga = angles.reshape(G, A, 3).transpose(1, 0, 2)  # (A, G, 3)
    _, _, allm = oracle.find_template(V, T, M, angles, W=W, return_all=True)
    cand = np.array([[28, 18, 16, 0], [28, 18, 16, 1], [0, 0, 0, 0], [63, 47, 31, 1], [5, 40, 30, 0],
                     [40, 3, 1, 1]], np.int32)
    tn, mu, sg = api.prepare_template(T, M, W)
    with api.SearchPlan(V, T.shape, M, batch=4) as p:
        p.set_template(tn, W, mu, sg)
        sc, an, pos = p.refine(V, ga, cand)
    for i, (x, y, z, g) in enumerate(cand):
        vals = []
        for a in range(A):
            for q in range(27):
                dx, dy, dz = q % 3 - 1, (q // 3) % 3 - 1, q // 9 - 1
                px, py, pz = (x + dx) % N[0], (y + dy) % N[1], (z + dz) % N[2]
                vals.append((allm[g * A + a][px, py, pz], a, (px, py, pz)))
        best = max(v[0] for v in vals)
        srt = sorted((v[0] for v in vals), reverse=True)
        err = abs(sc[i] - best)
        say(f"  cand {i}: gpu {sc[i]:.6f} fp64 {best:.6f} |d| {err:.2e}")
        assert err < SCORE_TOL
        if srt[0] - srt[1] > MARGIN:
            ref = next(v for v in vals if v[0] == best)
            assert an[i] == ref[1] and tuple(pos[i]) == ref[2]


def test_refine_centroid_windows(api, oracle):
    """tomhip_search_refine with a centroid d: candidates locate the rotated
    template point R d + p; orientations a = 4k .. 4k + 3 of group g are
    scored around cand - round(mean R d) over the 27 neighbours (FP64 oracle
    maps as in the test above, R of tom_rotate's convention)."""
    say("refine with centroid")
    from tom_amd import orient
    rng = np.random.default_rng(22)
    N = (64, 48, 32)
    n = 16
    V = rng.standard_normal(N).astype(np.float32)
    T = _phantom(rng, n, 4, 1.5, 5)
    V[20:36, 10:26, 8:24] += 2 * oracle.rotate(T, [30, 60, 90])
    M = oracle.sphere_mask_gauss(n, 6, 1.5)
    angles = np.concatenate([oracle.angle_grid(60, 60, 60)[::7], [[30, 60, 90]]])[:20]
    G, A = 2, 10
    ga = angles.reshape(G, A, 3).transpose(1, 0, 2)  # (A, G, 3)
    _, _, allm = oracle.find_template(V, T, M, angles, return_all=True)
    d = np.array([2.6, -1.7, 3.2])
    cand = np.array([[31, 15, 19, 0], [31, 15, 19, 1], [0, 0, 0, 0], [63, 47, 31, 1], [5, 40, 30, 0]], np.int32)
    tn, mu, sg = api.prepare_template(T, M)
    with api.SearchPlan(V, T.shape, M, batch=4) as p:
        p.set_template(tn, None, mu, sg)
        sc, an, pos = p.refine(V, ga, cand, centroid=d)
    R = orient.rotmatrix(ga.astype(np.float64))  # (A, G, 3, 3)
    Rd = R @ d
    for i, (x, y, z, g) in enumerate(cand):
        vals = []
        for a in range(A):
            k = a // 4
            sh = np.rint(Rd[4 * k:min(A, 4 * k + 4), g].mean(0)).astype(int)
            cx, cy, cz = x - sh[0], y - sh[1], z - sh[2]
            for q in range(27):
                dx, dy, dz = q % 3 - 1, (q // 3) % 3 - 1, q // 9 - 1
                px, py, pz = (cx + dx) % N[0], (cy + dy) % N[1], (cz + dz) % N[2]
                vals.append((allm[g * A + a][px, py, pz], a, (px, py, pz)))
        best = max(v[0] for v in vals)
        srt = sorted((v[0] for v in vals), reverse=True)
        err = abs(sc[i] - best)
        say(f"  cand {i}: gpu {sc[i]:.6f} fp64 {best:.6f} |d| {err:.2e}")
        assert err < SCORE_TOL
        if srt[0] - srt[1] > MARGIN:
            ref = next(v for v in vals if v[0] == best)
            assert an[i] == ref[1] and tuple(pos[i]) == ref[2]


def test_tom_bin_bit_exact(api, oracle):
    """tomhip_bin_f32 equals the tom_bin restatement bit for bit (odd edges,
    one and two rounds, numpy and device tensors)."""
    say("tom_bin")
    import torch
    rng = np.random.default_rng(23)
    V = rng.standard_normal((37, 22, 19)).astype(np.float32) * 100
    for nb in (1, 2):
        np.testing.assert_array_equal(api.tom_bin(V, nb), oracle.tom_bin(V, nb))
    Vt = torch.from_numpy(np.ascontiguousarray(V.transpose(2, 1, 0))).cuda()
    got = api.tom_bin(Vt, 1).cpu().numpy().transpose(2, 1, 0)
    np.testing.assert_array_equal(got, oracle.tom_bin(V, 1))


def test_autotune_picks_a_batch_with_identical_results(api, oracle):
    """The autotuner times candidate batches and checks that every batch gives
    the same maps before it reports a choice."""
    say("autotune")
    from tom_amd import autotune
    rng = np.random.default_rng(31)
    V = rng.standard_normal((64, 32, 32)).astype(np.float32)
    T = _phantom(rng, 16, 4, 1.5, 5)
    M = oracle.sphere_mask_binary(16, 7)
    tn, mu, sg = api.prepare_template(T, M)
    res = autotune.tune_search(V, tn, M, None, mu, sg, oracle.angle_grid(60, 60, 60), batches=(1, 3, 4, 7),
                               tiles=(8, 16, 32), zchunks=(0, 8), groups=3, use_cache=False)
    say(f"  {res}")
    assert res["batch"] in (1, 3, 4, 7) and res["tile_z"] in (8, 16, 32) and res["tile_y"] in (8, 16, 32)
    assert res["zchunk"] in (0, 8)
    assert not res["cached"]


def test_psr_matches_oracle(api, oracle):
    say("psr")
    rng = np.random.default_rng(41)
    x = rng.standard_normal((64, 48, 32)).astype(np.float32)
    x[30, 20, 10] += 8
    for n in (16, 24):
        got = api.tom_os3_psr(x, n)
        ref = oracle.psr(x, n)
        err = np.abs(got - ref).max()
        say(f"  n={n}: max |psr - fp64| = {err:.2e}")
        assert err < 1e-5


def test_batch_position_independent_with_weight(api, oracle):
    """With a box weight, an orientation scores the same at every position of
    a launch group (a duplicate never beats its first occurrence), and batch
    sizes give bit-identical maps (the weighting pass handles each template
    in whole blocks)."""
    say("batch position with weight")
    rng = np.random.default_rng(33)
    V = rng.standard_normal((64, 48, 32)).astype(np.float32)
    T = _phantom(rng, 24, 5, 1.5, 7)
    M = oracle.sphere_mask_gauss(24, 10, 2)
    W = (oracle.wedge((24, 24, 24), 30) * oracle.bandpass_weight((24, 24, 24), 1, 10)).astype(np.float32)
    tn, mu, sg = api.prepare_template(T, M, W)
    A, B = [30, 60, 90], [100, 20, 45]
    dup = np.asarray([A, A, A, A, B, A, A], np.float32)
    with api.SearchPlan(V, T.shape, M, batch=4) as p:
        p.set_template(tn, W, mu, sg)
        p.run(dup, 0)
        _, an = p.finalize()
    assert set(np.unique(an).tolist()) <= {0, 4}
    ang = oracle.angle_grid(60, 60, 60)[::3]
    outs = []
    for batch in (1, 3, 4):
        with api.SearchPlan(V, T.shape, M, batch=batch) as p:
            p.set_template(tn, W, mu, sg)
            p.run(ang, 0)
            outs.append(p.finalize())
    for sc, an in outs[1:]:
        np.testing.assert_array_equal(sc, outs[0][0])
        np.testing.assert_array_equal(an, outs[0][1])
