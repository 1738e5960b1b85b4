"""Peak analysis on the GPU against the oracle: tom_peak's forms (every tie,
subpixel parabola, radius erase), the box-size FLCF (template
autocorrelation), the tom_os3_bestPeak / os3_results combined update, the
secondary-score picker loop, and tom_os3_corr's low-std branch.
Tolerances: positions and erased maps exact; subpixel positions to the
0.001 grid; FLCF / PSR / SOC scores 1e-4 absolute against FP64 (PSR, a ratio
of order 10, 1e-4 relative on top); the update kernel bit-exact on given maps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def test_tom_peak_forms(api, oracle):
    say("tom_peak forms")
    A = np.array([[1, 2, 3, 4, 17], [5, 6, 7, 8, 18], [10, 9, 40, 6, 20], [12, 16, 17, 20, 30],
                  [5, 23, 35, 12, 6]], np.float32)
    c, v = api.tom_peak(A)
    assert list(c) == [3, 3] and v == 40  # Analysis/tom_peak.m:27-35
    c, v = api.tom_peak(A, "spline")
    np.testing.assert_allclose(c, [3.089, 2.977], atol=1e-9)  # tom_peak.m:37-50
    assert round(v, 4) == 40.2232
    rng = np.random.default_rng(3)
    a = rng.integers(0, 6, (9, 8, 7)).astype(np.float32)
    c, v = api.tom_peak(a)
    exp = np.argwhere(np.asfortranarray(a) == a.max())
    exp = exp[np.lexsort((exp[:, 0], exp[:, 1], exp[:, 2]))] + 1  # find order: x fastest
    assert v == a.max() and np.array_equal(np.atleast_2d(c), exp)
    # a smooth 3D peak: parabola refinement per axis as the oracle's tom_intpeak
    x = np.arange(20)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    g = np.exp(-((X - 9.3) ** 2 + (Y - 10.6) ** 2 + (Z - 8.8) ** 2) / 9.0).astype(np.float32)
    c, v = api.tom_peak(g, "spline")
    co, vo = oracle.peak(g, spline=True)
    np.testing.assert_allclose(c, co, atol=1e-9)
    assert abs(v - vo) < 1e-6
    # radius: the sphere |d| <= R around the first maximum is zeroed, clipped
    for R, where in ((3, (9, 10, 9)), (5, (1, 18, 0))):
        h = rng.standard_normal((20, 20, 20)).astype(np.float32)
        h[where] = 10.0
        c, v, M = api.tom_peak(h, R)
        assert list(c) == [w + 1 for w in where] and v == 10.0
        d = np.sqrt((X - where[0]) ** 2 + (Y - where[1]) ** 2 + (Z - where[2]) ** 2)
        exp = np.where(d <= R, 0.0, h).astype(np.float32)
        np.testing.assert_array_equal(M, exp)
        c2, v2, M2 = api.tom_peak(h, R, "spline")
        np.testing.assert_array_equal(M2, exp)
        co, vo = oracle.peak(h, spline=True)
        np.testing.assert_allclose(c2, co, atol=1e-9)


def test_flcf_box_matches_oracle(api, oracle):
    say("box FLCF")
    rng = np.random.default_rng(5)
    for n in (16, 24):
        T = rng.standard_normal((n, n, n)).astype(np.float32)
        S = (rng.standard_normal((n, n, n)) + 0.3).astype(np.float32)
        M = oracle.sphere_mask_gauss(n, n / 2 - 3, 1.5)
        got = api.flcf_box(S, T, M)
        ref = oracle.flcf_box(S, T, M)
        assert np.abs(got - ref).max() < 1e-4
        got = api.flcf_box(T, T)  # default mask: tom_os3_sphereMask
        ref = oracle.flcf_box(T, T)
        assert np.abs(got - ref).max() < 1e-4
        assert abs(got[n // 2, n // 2, n // 2] - 1.0) < 1e-5


def test_bestpeak_update_exact(api, oracle):
    """The update kernel on given maps equals the oracle's rule bit for bit,
    for each rule, over a sequence of orientations (numpy in, numpy out)."""
    say("bestPeak update")
    rng = np.random.default_rng(6)
    N = (33, 20, 17)
    maps = [tuple(rng.standard_normal(N).astype(np.float32) for _ in range(3)) for _ in range(5)]
    for rule in ("ccc", "psr", "soc", "ccc+psr", "ccc+soc", "psr+soc", "ccc+psr+soc"):
        g = {k: np.full(N, -1000.0, np.float32, order="F") for k in ("ccc", "psr", "soc")}
        g["angle"] = np.zeros(N, np.int32, order="F")
        o = {k: np.full(N, -1000.0) for k in ("ccc", "psr", "soc")}
        o["angle"] = np.zeros(N, np.int64)
        for i, (c, p, q) in enumerate(maps):
            api.tom_os3_bestPeak(g, c, p, q, angle=i + 1, rule=rule)
            oracle.bestpeak(o, c.astype(np.float64), p.astype(np.float64), q.astype(np.float64), i + 1, rule)
        for k in ("ccc", "psr", "soc"):
            np.testing.assert_array_equal(g[k], o[k].astype(np.float32))
        np.testing.assert_array_equal(g["angle"], o["angle"])


def test_os3_picker_scores_match_oracle(api, oracle):
    """One orientation's ccc, psr and soc maps of the picker loop against the
    oracle, then two orientations through the combined update."""
    import torch
    say("os3 picker scores")
    rng = np.random.default_rng(8)
    n, N = 16, (64, 64, 64)
    x = np.arange(n) - n // 2
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    T = (np.exp(-((X - 2) ** 2 + Y ** 2 + (Z + 1) ** 2) / 5.0) + 0.7 * np.exp(-((X + 3) ** 2 + (Y - 2) ** 2) / 3.0))
    T = T.astype(np.float32)
    V = rng.standard_normal(N).astype(np.float32)
    V[20:36, 18:34, 30:46] += 3 * oracle.rotate(T, [30, 60, 90])
    M = oracle.sphere_mask_binary(n, 7)
    Vd = torch.from_numpy(np.ascontiguousarray(V.transpose(2, 1, 0))).cuda()
    angles = np.array([[30, 60, 90], [0, 0, 0]], np.float32)
    res = api.os3_picker(Vd, T, M, angles[:1], rule="ccc+psr+soc")
    ccc, psr, soc = oracle.os3_scores(V, T, M, angles[0])
    for k, ref in (("ccc", ccc), ("psr", psr), ("soc", soc)):
        got = res[k].cpu().numpy().transpose(2, 1, 0)
        tol = 1e-4 * (1 + np.abs(ref)) if k == "psr" else 1e-4
        err = np.abs(got - ref)
        say(f"  {k}: max |gpu - fp64| {err.max():.2e}")
        assert np.all(err <= tol)
    # the planted copy (template centre ceil(n/2) at V[28, 26, 38]) is the
    # best-scoring voxel; its height is set by the noise under the mask
    c = res["ccc"].cpu().numpy().transpose(2, 1, 0)
    assert np.unravel_index(np.argmax(c), c.shape) == (28, 26, 38) and c[28, 26, 38] > 0.3
    res2 = api.os3_picker(Vd, T, M, angles, rule="ccc+psr+soc")
    ang = res2["angle"].cpu().numpy()
    assert set(np.unique(ang)) <= {0, 1, 2}
    assert np.all(res2["ccc"].cpu().numpy()[ang == 1] == res["ccc"].cpu().numpy()[ang == 1])


def test_os3_corr_low_std_branch(api, oracle):
    """tom_os3_corr.m:110-125. (1) An image of mean local std below 0.01 is
    scored with the renormalised template (T*1000+1000).*M, 'mean0+1std',
    which changes the score under a soft mask. (2) When even the renormalised
    image is featureless (a lone spike in a near-zero volume, mean local std
    below 0.01 of the global std), voxels of local std below 0.05 of the
    global std score 0 while the spike's neighbourhood keeps its FLCF."""
    say("os3_corr low-std branch")
    rng = np.random.default_rng(11)

    def expected(V, T, M):
        Md = M.astype(np.float64)
        m = float((Md > 0).sum())
        Vd = V.astype(np.float64)
        mu_I, sig_I = oracle.local_stats(Vd, Md, m)
        Tr = (T.astype(np.float64) * 1000 + 1000) * Md
        Tr = (Tr - Tr.mean()) / Tr.std()
        mu_t, sig_t = oracle.norm_under_mask(Tr, Md)
        return oracle.flcf(Vd, Tr * Md, m, mu_t, sig_t, mu_I, sig_I), sig_I, Vd.std()

    n = 12
    T = rng.standard_normal((n, n, n)).astype(np.float32)
    Ms = oracle.sphere_mask_gauss(n, 4, 1.5)
    V = (1e-3 * rng.standard_normal((64, 32, 32))).astype(np.float32)
    got = api.tom_os3_corr(V, T, Ms)
    ref, sig_I, s_img = expected(V, T, Ms)
    assert sig_I.mean() < 0.01 and sig_I.mean() / s_img >= 0.01  # first test only
    assert not np.any(sig_I < 0.05 * s_img)
    err = np.abs(got - ref).max()
    plain = oracle.flcf_angle(V.astype(np.float64), T.astype(np.float64), Ms.astype(np.float64),
                              float((Ms > 0).sum()), *oracle.norm_under_mask(T.astype(np.float64), Ms),
                              *oracle.local_stats(V.astype(np.float64), Ms.astype(np.float64), float((Ms > 0).sum())),
                              None, [0, 0, 0])
    say(f"  renormalised template: max |gpu - fp64| {err:.2e}; differs from the plain score by "
        f"{np.abs(plain - ref).max():.2e}")
    assert err < 1e-4 and np.abs(plain - ref).max() > 1e-3
    Mb = oracle.sphere_mask_binary(n, 5)
    N = (256, 256, 128)
    V = (1e-7 * rng.standard_normal(N)).astype(np.float32)
    V[100, 120, 60] = 1.0
    got = api.tom_os3_corr(V, T, Mb)
    ref, sig_I, s_img = expected(V, T, Mb)
    # both tests of tom_os3_corr.m:110-125 fire: mean local std below 0.01,
    # and below 0.01 again after the image is brought to unit std
    assert sig_I.mean() / s_img < 0.01
    hot = sig_I < 0.05 * s_img
    assert hot.mean() > 0.99 and (~hot).sum() > 100
    say(f"  hot voxels {int(hot.sum())}, kept {int((~hot).sum())}")
    assert np.all(got[hot] == 0)
    assert np.abs(got - ref)[~hot].max() < 1e-4
