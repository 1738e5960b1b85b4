"""Pins the CPU oracle (oracle/) against the reference's own material:
golden vectors from the reference C rotation core, the KATs in the
reference's docstrings, closed forms, brute force and invariants.
No GPU needed."""
import hashlib
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "rot3d_golden.npz")


def _golden():
    g = np.load(GOLD)
    vols = {24: g["vol24"], 25: g["vol25"], 48: g["vol48"], 2030: g["vol2030"]}
    return g, vols


# ---------------------------------------------------------------- rotation
def test_rotation_golden_bit_exact(oracle):
    """rot3d restatement == reference rot3dlin on every golden case
    (Sptrans/tom_rotatec.c:146-401)."""
    g, vols = _golden()
    for i in range(len(g["case_sha256"])):
        v = vols[int(g["case_vol"][i])]
        ne = int(g["case_neuler"][i])
        o = oracle.rotate(v, g["case_euler"][i][:ne], g["case_center"][i])
        assert hashlib.sha256(o.tobytes(order="F")).hexdigest() == g["case_sha256"][i], i
    for k, i in enumerate(g["full_idx"]):
        v = vols[int(g["case_vol"][i])]
        o = oracle.rotate(v, g["case_euler"][i][:3], g["case_center"][i])
        np.testing.assert_array_equal(o, g["full_out"][k])


def test_rotation_against_built_reference(oracle):
    """Extra random angles against oracle/_ref when it is built here."""
    so = os.path.join(ROOT, "oracle", "_ref", "libtomref.so")
    if not os.path.exists(so):
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    import ctypes
    L = ctypes.CDLL(so)
    fp = ctypes.POINTER(ctypes.c_float)
    L.rot3dlin.argtypes = [fp, fp, ctypes.c_long, ctypes.c_long, ctypes.c_long, fp, ctypes.c_float,
                           ctypes.c_float, ctypes.c_float, ctypes.c_int]
    rng = np.random.default_rng(7)
    v = np.asfortranarray(rng.standard_normal((24, 20, 28)).astype(np.float32))
    for _ in range(40):
        e = rng.uniform(-400, 400, 3).astype(np.float32)
        out = np.zeros_like(v, order="F")
        L.rot3dlin(v.ctypes.data_as(fp), out.ctypes.data_as(fp), 24, 20, 28, e.ctypes.data_as(fp), 12, 10, 14, 1)
        np.testing.assert_array_equal(oracle.rotate(v, e, [12, 10, 14]), out)


def test_rotation_axis_invariants(oracle):
    """Sptrans/tom_rotate.m:59-61: X-axis [0 0 a], Y-axis [270 90 a], Z-axis [a 0 0];
    90-degree turns permute voxels exactly (exact cos/sin table)."""
    rng = np.random.default_rng(3)
    v = rng.standard_normal((16, 16, 16)).astype(np.float32)
    # rotation about z by 90 degrees maps voxel (x,y) -> (c - (y-c), c + (x-c))
    r = oracle.rotate(v, [90, 0, 0])
    c = 8
    for x in range(1, 16):
        for y in range(1, 16):
            # inverse mapping: out(i,j) = in(c + rm00*(i-c) + rm10*(j-c), ...)
            rm = oracle.rotmat([90, 0, 0])
            sx = c + rm[0] * (x - c) + rm[1] * (y - c)
            sy = c + rm[3] * (x - c) + rm[4] * (y - c)
            if 0 <= sx <= 15 and 0 <= sy <= 15:
                assert r[x, y, 5] == v[int(sx), int(sy), 5]
    # the three axis conventions are rotations about the named axis: the axis
    # coordinate is preserved by the matrix
    a = 37.0
    rmx = oracle.rotmat([0, 0, a]).reshape(3, 3, order="F")  # rows i, cols j = rm_ij
    rmy = oracle.rotmat([270, 90, a]).reshape(3, 3, order="F")
    rmz = oracle.rotmat([a, 0, 0]).reshape(3, 3, order="F")
    np.testing.assert_allclose(rmx @ [1, 0, 0], [1, 0, 0], atol=1e-6)
    np.testing.assert_allclose(rmy @ [0, 1, 0], [0, 1, 0], atol=1e-6)
    np.testing.assert_allclose(rmz @ [0, 0, 1], [0, 0, 1], atol=1e-6)


def test_rotmat_formula(oracle):
    """tom_rotatec.c:17-25 explicit matrix == Rz(psi) Rx(theta) Rz(phi)."""
    rng = np.random.default_rng(11)
    for _ in range(20):
        phi, psi, the = rng.uniform(0, 360, 3)
        rm = oracle.rotmat([phi, psi, the]).astype(np.float64)
        f, p, t = np.deg2rad([np.float32(phi), np.float32(psi), np.float32(the)])

        def rz(a):
            return np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])

        def rx(a):
            return np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])

        R = rz(p) @ rx(t) @ rz(f)
        # rm_ij (i = line) stored rm00 rm10 rm20 rm01 ...: column order of R;
        # the sampling is src = p + R^T (dst - p) (tom_rotatec.c:315-327)
        M = rm.reshape(3, 3, order="F")  # M[i, j] = rm_ij
        np.testing.assert_allclose(M, R, atol=2e-6)


# ---------------------------------------------------------------- peaks
def test_tom_peak_kat(oracle):
    """Analysis/tom_peak.m:27-50 worked example."""
    A = np.array([[1, 2, 3, 4, 17], [5, 6, 7, 8, 18], [10, 9, 40, 6, 20], [12, 16, 17, 20, 30],
                  [5, 23, 35, 12, 6]], float)
    c, v = oracle.peak(A)
    assert list(c) == [3, 3] and v == 40
    c, v = oracle.peak(A, spline=True)
    np.testing.assert_allclose(c, [3.089, 2.977], atol=1e-9)
    assert round(v, 4) == 40.2232
    # the parabola vertex in closed form: 2 + 10/112 and 40 + 100/448
    assert abs((2 + 10 / 112) - 2.0893) < 1e-4 and abs((40 + 100 / 448) - 40.2232) < 1e-4


def test_find_max_first_in_x_fastest(oracle):
    rng = np.random.default_rng(5)
    a = rng.integers(0, 5, size=(7, 6, 5)).astype(np.float32)  # many ties
    v, pos = oracle.find_max(a)
    flat = np.asfortranarray(a).ravel(order="F")
    i = int(np.argmax(flat))
    assert v == flat[i]
    assert list(pos) == [i % 7, (i // 7) % 6, i // 42]
    v, pos = oracle.find_max(np.full((3, 3, 3), -1e6, np.float32))
    assert v == -99999.0


def test_fourier_kat(oracle):
    """Filtrans/tom_fourier.m:18-22 (pins the forward sign and no scaling).

    The printed matrix has two typos that contradict its own Hermitian
    symmetry (a real input needs c(1,2) = conj(c(1,4)) and
    c(2,2) = conj(c(4,4))): the docstring prints 3+19i and 50+9i there. Those
    two entries are checked against the values their printed conjugate
    partners imply (3-19i, 59+9i); the other 14 are checked as printed."""
    a = np.array([[11, 3, -6, -14], [1, -13, 9, 6], [0, 8, 13, -3], [-8, 17, -15, 7]], float)
    c = oracle.fourier(a)
    printed = np.array([[16, 3 + 19j, -6, 3 + 19j], [-24 - 2j, 50 + 9j, 8 - 64j, 1 + 21j],
                        [8, 5 - 37j, 54, 5 + 37j], [-24 + 2j, 1 - 21j, 8 + 64j, 59 - 9j]])
    typo = np.zeros((4, 4), bool)
    typo[0, 1] = typo[1, 1] = True
    np.testing.assert_allclose(c[~typo], printed[~typo], atol=1e-9)
    np.testing.assert_allclose(c[0, 1], np.conj(printed[0, 3]), atol=1e-9)
    np.testing.assert_allclose(c[1, 1], np.conj(printed[3, 3]), atol=1e-9)


# ---------------------------------------------------------------- masks/weights
def test_tom_norm_mean0_1std_closed_form(oracle):
    """tom_norm(1e5*(T+1), 'mean0+1std') (Sptrans/tom_norm.m:83-119 with
    tom_dev's population std, Analysis/tom_dev.m:55-60) is affine-invariant:
    it equals (T - mean T) / std_pop(T). The sample std (n - 1) fails it;
    a constant template gives zeros (the st == 0 branch)."""
    rng = np.random.default_rng(12)
    T = (3.0 * rng.standard_normal((10, 9, 8)) + 0.4).astype(np.float32)
    Td = T.astype(np.float64)
    out = oracle.tom_norm_mean0_1std(T)
    np.testing.assert_allclose(out, (Td - Td.mean()) / Td.std(), atol=1e-9)
    assert abs(out.mean()) < 1e-12 and abs(out.std() - 1) < 1e-12
    assert np.abs(out - (Td - Td.mean()) / Td.std(ddof=1)).max() > 1e-4
    np.testing.assert_array_equal(oracle.tom_norm_mean0_1std(np.full((4, 4, 4), 2.5)), 0.0)


def test_mask_cc_sphere_lattice_counts(oracle):
    """The peak-search mask (init_spheremask radius r, sigma 0, then the
    inverse fftshift of CorrelationHandlerPeaks, correlation_handler.cpp:
    65-80): shifts s with |s| <= r around the origin of the unshifted
    correlation volume. Its size is the number of integer points in the ball
    of radius r (1, 7, 33, 123, 257, 515 for r = 0..5; 19 and 27 for
    r^2 = 2 and 3), it holds the zero shift and wraps to negative shifts."""
    n = 16
    for r, cnt in ((0, 1), (1, 7), (2, 33), (3, 123), (4, 257), (5, 515), (np.sqrt(2), 19), (np.sqrt(3), 27)):
        m = oracle.mask_cc_sphere((n, n, n), r + 1e-9)
        assert int(m.sum()) == cnt, r
    m = oracle.mask_cc_sphere((n, n, n), 3)
    assert m[0, 0, 0] == 1 and m[n - 3, 0, 0] == 1 and m[3, 0, 0] == 1 and m[4, 0, 0] == 0
    assert m[n // 2, n // 2, n // 2] == 0
    np.testing.assert_array_equal(m, np.roll(m[::-1, ::-1, ::-1], 1, axis=(0, 1, 2)))


def test_masks_closed_form(oracle):
    m, cnt = oracle.sphere_mask_os3(24)
    x = np.arange(24) - 12
    r = np.sqrt(x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2)
    np.testing.assert_array_equal(m, (r < 11).astype(np.float32))
    assert cnt == int((r < 11).sum())
    g = oracle.sphere_mask_gauss(24, 10, 2)
    assert g[12, 12, 12] == 1 and g[12, 12, 22] == 1  # r = 10
    np.testing.assert_allclose(g[12, 12, 23], np.exp(-0.25), rtol=1e-6)  # r = 11
    assert g[12, 12, 0] == 0 or r[12, 12, 0] <= 12  # r = 12 is the last ring kept
    np.testing.assert_allclose(g[12, 12, 0], np.exp(-1.0), rtol=1e-6)
    assert g[0, 0, 0] == 0


def test_wedge_bandpass_predicates(oracle):
    w = oracle.wedge((16, 16, 16), 30)
    x = np.arange(16) - 8
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    assert np.all(w[:, :, 8] == 1)  # kz = 0 plane kept
    miss = (Z != 0) & (np.tan(np.pi / 6) > np.abs(X) / np.maximum(np.abs(Z), 1e-300))
    np.testing.assert_array_equal(w == 0, miss)
    b = oracle.bandpass_weight((16, 16, 16), 2, 5)
    L = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    np.testing.assert_array_equal(b == 1, (L >= 2) & (L <= 5))


def test_apply_weight_identity_and_linearity(oracle):
    rng = np.random.default_rng(2)
    t = rng.standard_normal((12, 12, 12))
    np.testing.assert_allclose(oracle.apply_weight(t, np.ones((12, 12, 12))), t, atol=1e-12)
    W = oracle.bandpass_weight((12, 12, 12), 1, 4)
    u = rng.standard_normal((12, 12, 12))
    np.testing.assert_allclose(oracle.apply_weight(2 * t + u, W),
                               2 * oracle.apply_weight(t, W) + oracle.apply_weight(u, W), atol=1e-10)
    # projection: applying a binary weight twice changes nothing
    a = oracle.apply_weight(t, W)
    np.testing.assert_allclose(oracle.apply_weight(a, W), a, atol=1e-10)


# ---------------------------------------------------------------- statistics / FLCF
def _direct(oracle, V, tm, M, m, mu_T, sig_T):
    import ctypes
    L = oracle.lib()
    V = np.asfortranarray(V, np.float32)
    N = np.asarray(V.shape, np.int64)
    n = np.asarray(M.shape, np.int64)
    tmd = np.asfortranarray(tm, np.float64)
    Md = np.asfortranarray(M, np.float64)
    sc = np.zeros(V.shape, np.float64, order="F")
    mu = np.zeros(V.shape, np.float64, order="F")
    sg = np.zeros(V.shape, np.float64, order="F")
    dp = ctypes.POINTER(ctypes.c_double)
    lp = ctypes.POINTER(ctypes.c_long)
    L.orc_flcf_direct(V.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), N.ctypes.data_as(lp),
                      tmd.ctypes.data_as(dp), Md.ctypes.data_as(dp), n.ctypes.data_as(lp), m, mu_T, sig_T,
                      sc.ctypes.data_as(dp), mu.ctypes.data_as(dp), sg.ctypes.data_as(dp))
    return sc, mu, sg


@pytest.mark.parametrize("N,n", [((12, 10, 8), (5, 5, 5)), ((11, 9, 13), (6, 4, 5)), ((8, 8, 8), (6, 6, 6))])
def test_flcf_fft_matches_direct(oracle, N, n):
    """FFT restatement (pasteCenter/ifftshift/convolution conventions of
    tom_os3_corr/mean/std) == brute-force direct sums, even and odd sizes."""
    rng = np.random.default_rng(sum(N) + sum(n))
    V = rng.standard_normal(N).astype(np.float32) + 0.3
    M = rng.uniform(0, 1, n) * (rng.uniform(0, 1, n) > 0.3)
    tm = rng.standard_normal(n) * M
    m = float((M > 0).sum())
    mu_I, sig_I = oracle.local_stats(V, M, m)
    s_fft = oracle.flcf(V, tm, m, 0.2, 1.3, mu_I, sig_I)
    s_dir, mu_d, sg_d = _direct(oracle, V, tm, M, m, 0.2, 1.3)
    np.testing.assert_allclose(mu_I, mu_d, atol=1e-12)
    np.testing.assert_allclose(sig_I, sg_d, atol=1e-10)
    np.testing.assert_allclose(s_fft, s_dir, atol=1e-10)


def test_flcf_planted_copy_scores_one(oracle):
    """A noise-free copy of the template inside a binary mask scores 1 at
    its centre (Roseman's FLCF)."""
    rng = np.random.default_rng(1)
    n = 8
    M = oracle.sphere_mask_binary(n, 3.5).astype(np.float64)
    T = rng.standard_normal((n, n, n))
    V = rng.standard_normal((24, 24, 24)) * 0.0 + 0.0
    V = rng.standard_normal((24, 24, 24))
    c = (10, 13, 9)
    V[c[0] - 4:c[0] + 4, c[1] - 4:c[1] + 4, c[2] - 4:c[2] + 4] = T
    m = float((M > 0).sum())
    mu_T, sig_T = oracle.norm_under_mask(T, M)
    mu_I, sig_I = oracle.local_stats(V, M, m)
    s = oracle.flcf(V, T * M, m, mu_T, sig_T, mu_I, sig_I)
    assert abs(s[c] - 1.0) < 1e-9
    assert s.max() <= 1 + 1e-9


def test_flcf_affine_invariance_and_zero_variance(oracle):
    rng = np.random.default_rng(4)
    V = rng.standard_normal((16, 16, 16))
    M = oracle.sphere_mask_binary(6, 2.5).astype(np.float64)
    T = rng.standard_normal((6, 6, 6))
    m = float((M > 0).sum())
    mu_T, sig_T = oracle.norm_under_mask(T, M)
    tm = (T - mu_T) * M  # zero-mean under the mask: score is affine invariant in V
    s1 = oracle.flcf(V, tm, m, 0.0, sig_T, *oracle.local_stats(V, M, m))
    V2 = 3.5 * V - 7
    s2 = oracle.flcf(V2, tm, m, 0.0, sig_T, *oracle.local_stats(V2, M, m))
    np.testing.assert_allclose(s1, s2, atol=1e-9)
    # constant volume: zero local variance -> score 0
    Vc = np.full((16, 16, 16), 2.5)
    s3 = oracle.flcf(Vc, tm, m, 0.0, sig_T, *oracle.local_stats(Vc, M, m))
    assert np.all(s3 == 0)


def test_find_template_bruteforce(oracle):
    """The reduction: strict '<' update, first angle wins, 1-based index."""
    rng = np.random.default_rng(9)
    V = rng.standard_normal((16, 16, 16)).astype(np.float32)
    T = rng.standard_normal((6, 6, 6)).astype(np.float32)
    M = oracle.sphere_mask_binary(6, 3).astype(np.float64)
    angles = np.array([[0, 0, 0], [90, 0, 0], [0, 0, 0], [30, 45, 60]], float)  # duplicate angle -> tie
    ccc, ang, allm = oracle.find_template(V, T, M, angles, return_all=True)
    best = allm.max(axis=0)
    np.testing.assert_array_equal(ccc, best)
    first = np.argmax(allm == best[None], axis=0) + 1
    np.testing.assert_array_equal(ang, first)
    assert not np.any(ang == 3)  # the duplicate of angle 1 never wins


def test_angle_grid_counts(oracle):
    assert oracle.angle_grid(30, 30, 30).shape == (1008, 3)
    assert oracle.angle_grid(15, 15, 15).shape == (7488, 3)
    assert oracle.angle_grid(10, 10, 10).shape == (24624, 3)
    g = oracle.angle_grid(30, 30, 30)
    assert list(g[0]) == [0, 0, 0] and list(g[1]) == [0, 0, 30] and list(g[7]) == [0, 30, 0]


def test_oscar_orientation(oracle):
    """tom_oscar correlates the other way round: its map at k is the FLCF-style
    raw correlation at -k."""
    rng = np.random.default_rng(12)
    V = rng.standard_normal((12, 12, 12))
    T = rng.standard_normal((4, 4, 4)).astype(np.float32)
    ccf, idx = oracle.oscar(V, T, np.zeros((1, 3)))
    raw = oracle.correlation_raw(V, T.astype(np.float64))
    neg = raw[np.ix_(*[(-np.arange(12)) % 12] * 3)]
    np.testing.assert_allclose(ccf, neg, atol=1e-10)


def test_psr_direct_sums(oracle):
    """PSR = (x - local mean) / local std under the os3_psr ring, against
    direct sums mu[k] = sum_u x[k + h - u] M[u] / m (h = floor(n/2), indices
    modulo N), population variance."""
    rng = np.random.default_rng(40)
    N, n = (16, 14, 12), 12
    x = rng.standard_normal(N)
    M = oracle.psr_ring(n)
    assert M.min() == 0 and M.max() == 1
    r = np.sqrt(sum(g * g for g in np.meshgrid(*[np.arange(n) - n // 2] * 3, indexing="ij")))
    np.testing.assert_array_equal(M > 0, (r < 6) & ~(r < 5))
    m = M.sum()
    h = n // 2
    mu = np.zeros(N)
    e2 = np.zeros(N)
    for u in zip(*np.nonzero(M)):
        sh = tuple(h - ui for ui in u)  # x[k + h - u] = roll(x, u - h)[k]
        xs = np.roll(x, tuple(-s for s in sh), axis=(0, 1, 2))
        mu += xs
        e2 += xs * xs
    mu /= m
    sd = np.sqrt(np.maximum(e2 / m - mu * mu, 0))
    np.testing.assert_allclose(oracle.psr(x, n), (x - mu) / sd, rtol=1e-9, atol=1e-9)


def test_picklist_planted_peaks(oracle):
    """Separated sharp peaks of known heights come back in descending order at
    their positions (1-based), each pick erasing a sphere of radius
    ceil(n/2) around it; border voxels (tom_os3_eraseBorders) never appear."""
    N, n = (40, 36, 32), 6
    a = np.zeros(N)
    pos = [(10, 10, 10), (28, 12, 20), (12, 26, 22), (30, 28, 9)]
    hts = [0.5, 0.9, 0.7, 0.6]
    rng = np.random.default_rng(3)
    a += 0.01 * rng.random(N)
    for p, h in zip(pos, hts):
        a[p] = h
    a[0, 0, 0] = 5.0  # in the erased border
    ang = np.arange(np.prod(N)).reshape(N, order="F").astype(np.int32)
    rows = oracle.return_picklist(a.astype(np.float32), ang, (n, n, n), 4)
    order = np.argsort(hts)[::-1]
    for r, k in zip(rows, order):
        p = pos[k]
        assert tuple(int(v) for v in r[:3]) == tuple(q + 1 for q in p)
        assert abs(r[3] - hts[k]) < 1e-7 and r[4] == ang[p]
    # the first pick erased its sphere (radius 3): a neighbour at distance 2 is 0
    _, _, after = oracle.peakc(oracle.erase_borders(a, (n, n, n)).astype(np.float32), 3)
    assert after[28, 12, 22] == 0 and after[28, 12, 20] == 0 and after[28, 12, 24] != 0


def test_peakc_border_paste_by_hand(oracle):
    """tom_peakc near the borders, values worked out by hand from
    Sptrans/tom_cut_out.m (start clamped to 1, size 2r kept) and
    Sptrans/tom_paste.m (pasted at the unclamped start c - r).
    Peak at MATLAB (1,1,1), r = 3: the cut-out is in(1:6,1:6,1:6) with sphere
    centre 4; it is pasted at -2, so a(1:3,1:3,1:3) = box(4:6,4:6,4:6). Of
    those 27 voxels only the four at box distance >= 3 from the centre keep a
    value, a(3,3,3) = in(6,6,6), a(3,3,2) = in(6,6,5), a(3,2,3) = in(6,5,6),
    a(2,3,3) = in(5,6,6); the other 23 become 0; nothing else changes.
    Peak at the last x voxel: the clipped box is 4 wide, its centre 3 sits one
    voxel inside, so a voxel 3 away along -x is zeroed while one 3 away along
    +y (distance sqrt(10) to the centre) is kept."""
    N = (12, 11, 10)
    a = (np.arange(np.prod(N), dtype=np.float32).reshape(N, order="F") + 1) / 1000
    a[0, 0, 0] = 50.0
    c, v, b = oracle.peakc(a, 3)
    assert list(c) == [1, 1, 1] and v == 50.0
    exp = a.copy()
    exp[0:3, 0:3, 0:3] = 0
    exp[2, 2, 2], exp[2, 2, 1], exp[2, 1, 2], exp[1, 2, 2] = a[5, 5, 5], a[5, 5, 4], a[5, 4, 5], a[4, 5, 5]
    np.testing.assert_array_equal(b, exp)
    a2 = a.copy()
    a2[0, 0, 0] = 0.5
    a2[11, 5, 5] = 60.0
    c, v, b = oracle.peakc(a2, 3)
    assert list(c) == [12, 6, 6] and v == 60.0
    assert b[11, 5, 5] == 0 and b[8, 5, 5] == 0 and b[11, 8, 5] == a2[11, 8, 5] and b[7, 5, 5] == a2[7, 5, 5]
    # the erased voxels are the box voxels closer than 3 to the centre (10.5 .. ): count them
    changed = b != a2
    assert changed.sum() == np.count_nonzero(
        [(i - 2) ** 2 + (j - 3) ** 2 + (k - 3) ** 2 < 9 for i in range(4) for j in range(6) for k in range(6)])


def test_picklist_stops_when_flat_or_not_finite(oracle):
    """The loop runs while std(peaks(:)) > 0: a constant map and a map with
    a NaN give no rows; a map of four separated spikes on zeros gives exactly
    four rows (after them every voxel is 0)."""
    N, n = (30, 30, 30), 4
    # (a template of the map's size: no border erase)
    assert oracle.return_picklist(np.full(N, 2.0, np.float32), None, N, 10).shape == (0, 5)
    a = np.zeros(N, np.float32)
    for i, p in enumerate([(8, 8, 8), (20, 9, 12), (10, 21, 19), (19, 20, 8)]):
        a[p] = 1 + i
    rows = oracle.return_picklist(a, None, (n, n, n), 10)
    assert rows.shape == (4, 5) and list(rows[:, 3]) == [4, 3, 2, 1]
    a[15, 15, 15] = np.nan
    assert oracle.return_picklist(a, None, (n, n, n), 10).shape == (0, 5)


def test_tom_bin_closed_forms(oracle):
    """tom_bin restatement: a constant stays constant, a linear ramp
    a + bx + cy + dz bins to its value at the block centres, odd trailing
    samples are dropped, two rounds halve twice, and one block equals its
    eight samples summed by hand in the .m file's order."""
    c = np.full((6, 4, 8), 3.25, np.float32)
    np.testing.assert_array_equal(oracle.tom_bin(c), np.full((3, 2, 4), 3.25, np.float32))
    x, y, z = np.meshgrid(np.arange(8.0), np.arange(6.0), np.arange(4.0), indexing="ij")
    ramp = (1.0 + 0.5 * x + 0.25 * y + 2.0 * z).astype(np.float32)
    xb, yb, zb = np.meshgrid(np.arange(4) * 2 + 0.5, np.arange(3) * 2 + 0.5, np.arange(2) * 2 + 0.5, indexing="ij")
    np.testing.assert_array_equal(oracle.tom_bin(ramp), (1.0 + 0.5 * xb + 0.25 * yb + 2.0 * zb).astype(np.float32))
    odd = np.random.default_rng(3).standard_normal((7, 5, 9)).astype(np.float32)
    assert oracle.tom_bin(odd).shape == (3, 2, 4)
    np.testing.assert_array_equal(oracle.tom_bin(odd), oracle.tom_bin(odd[:6, :4, :8]))
    assert oracle.tom_bin(np.ones((16, 8, 12), np.float32), 2).shape == (4, 2, 3)
    v = odd
    acc = np.float32(v[0, 0, 0])
    for (i, j, k) in ((0, 1, 0), (1, 0, 0), (1, 1, 0), (0, 0, 1), (0, 1, 1), (1, 0, 1), (1, 1, 1)):
        acc = np.float32(acc + v[i, j, k])
    assert oracle.tom_bin(v)[0, 0, 0] == np.float32(acc / np.float32(8))


def test_cpu_fp32_baseline_matches_oracle(oracle):
    """The FP32 CPU baseline (the reference's loop in single precision) agrees
    with the FP64 oracle's find_template to FP32 rounding on a small case,
    best angles identical where the margin exceeds 1e-3."""
    from oracle.cpu_fp32 import CpuFlcfFP32
    rng = np.random.default_rng(4)
    n, N = 12, (32, 24, 28)
    x = np.arange(n) - n // 2
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    T = (np.exp(-((X - 2) ** 2 + Y ** 2 + (Z + 1) ** 2) / 4.0) + 0.6 * np.exp(-((X + 2) ** 2 + (Y - 2) ** 2) / 3.0))
    T = T.astype(np.float32)
    V = rng.standard_normal(N).astype(np.float32)
    V[6:18, 5:17, 8:20] += 3 * oracle.rotate(T, [30, 60, 90])
    M = oracle.sphere_mask_binary(n, 5)
    W = oracle.wedge((n, n, n), 30) * oracle.bandpass_weight((n, n, n), 1, 5)
    angles = np.array([[0, 0, 0], [30, 60, 90], [90, 30, 45]], float)
    ref, ref_ang, allm = oracle.find_template(V, T, M, angles, W=W, return_all=True)
    cpu = CpuFlcfFP32(V, T, M, W)
    for a in angles:
        cpu.step(a)
    got = cpu.best.numpy()
    assert np.abs(got - ref).max() < 2e-3
    srt = np.sort(allm, axis=0)
    sel = (srt[-1] - srt[-2]) > 1e-3
    assert np.all(cpu.arg.numpy()[sel] == ref_ang[sel])
    assert cpu.arg.numpy()[12, 11, 14] == 2


def test_flcf_at_voxels_matches_fft_maps(oracle):
    """The sampled-voxel direct-sum FLCF equals the FFT restatement's maps
    (find_template's per-angle scores) at every sampled voxel, including
    voxels whose template window wraps around the volume edge."""
    rng = np.random.default_rng(6)
    n, N = 10, (24, 20, 22)
    T = rng.standard_normal((n, n, n)).astype(np.float32)
    V = rng.standard_normal(N).astype(np.float32) + 0.5
    M = oracle.sphere_mask_gauss(n, 3.5, 1.0)
    W = oracle.wedge((n, n, n), 30) * oracle.bandpass_weight((n, n, n), 1, 4)
    angles = np.array([[0, 0, 0], [30, 60, 90], [200, 10, 45]], float)
    _, _, allm = oracle.find_template(V, T, M, angles, W=W, return_all=True)
    Tn, mu_T, sig_T = oracle.prepare_template(T, M.astype(np.float64), W)
    vox = np.array([[0, 0, 0], [23, 19, 21], [12, 10, 11], [3, 17, 5], [20, 2, 9]])
    for workers in (1, 2):
        s = oracle.flcf_at_voxels(V, Tn, M, mu_T, sig_T, angles, vox, W=W, chunk=2, workers=workers)
        ref = allm[:, vox[:, 0], vox[:, 1], vox[:, 2]].T
        np.testing.assert_allclose(s, ref, atol=1e-10)


def test_rotate2d_closed_forms(oracle):
    """rot2d (Sptrans/tom_rotatec.c:49-142): 0 degrees is the identity; a
    linear ramp f = 2x + 3y + 1 rotates exactly under bilinear interpolation,
    with the content turned by +phi about the pivot (the output at d samples
    the input at p + R(-phi)(d - p)), where the source lies inside the image;
    pixels whose source falls outside are 0; the matrix form with the same
    entries gives the same image."""
    n = 16
    rng = np.random.default_rng(2)
    img = rng.standard_normal((n, n)).astype(np.float32)
    np.testing.assert_array_equal(oracle.rotate2d(img, 0.0), img)
    x = np.arange(n, dtype=np.float64)
    X, Y = np.meshgrid(x, x, indexing="ij")
    ramp = (2 * X + 3 * Y + 1).astype(np.float32)
    for phi, p in ((30.0, (8, 8)), (-70.0, (6, 9)), (135.0, (8, 7))):
        out = oracle.rotate2d(ramp, phi, p)
        a = np.radians(phi)
        sx = p[0] + np.cos(a) * (X - p[0]) + np.sin(a) * (Y - p[1])
        sy = p[1] - np.sin(a) * (X - p[0]) + np.cos(a) * (Y - p[1])
        inside = (sx > 1e-3) & (sx < n - 1 - 1e-3) & (sy > 1e-3) & (sy < n - 1 - 1e-3)
        outside = (sx < -1e-3) | (sx > n - 1 + 1e-3) | (sy < -1e-3) | (sy > n - 1 + 1e-3)
        np.testing.assert_allclose(out[inside], (2 * sx + 3 * sy + 1)[inside], atol=1e-3)
        assert np.all(out[outside] == 0)
        # the opposite sense would put the ramp elsewhere
        assert np.abs(out[inside] - (2 * (2 * p[0] - sx) + 3 * (2 * p[1] - sy) + 1)[inside]).max() > 1.0
    a = np.float32(np.radians(np.float32(30.0)))
    c, s = np.float32(np.cos(np.float64(a))), np.float32(np.sin(np.float64(a)))
    np.testing.assert_array_equal(oracle.rotate2d(img, np.array([c, s, -s, c], np.float32), (8, 8)),
                                  oracle.rotate2d(img, 30.0, (8, 8)))


def test_flcf_box_direct_sums_and_autocorrelation(oracle):
    """Box-size tom_os3_corr equals the brute-force direct-sum FLCF at N = n,
    and a template's autocorrelation scores 1 at zero shift (the box centre
    floor(n/2), pasteCenter + ifftshift) under a binary mask."""
    rng = np.random.default_rng(9)
    n = 10
    T = rng.standard_normal((n, n, n))
    S = rng.standard_normal((n, n, n)) + 0.2
    M = oracle.sphere_mask_binary(n, 4.2).astype(np.float64)
    got = oracle.flcf_box(S, T, M)
    m = float((M > 0).sum())
    mu_T, sig_T = oracle.norm_under_mask(T, M)
    sc, _, _ = _direct(oracle, S.astype(np.float32), T * M, M, m, mu_T, sig_T)
    np.testing.assert_allclose(got, sc, atol=1e-5)  # S rounded to float32 for the C sums
    auto = oracle.flcf_box(T * M, T * M, M)
    assert abs(auto[n // 2, n // 2, n // 2] - 1.0) < 1e-12
    assert np.nanmax(auto) <= 1.0 + 1e-12


def test_bestpeak_rules(oracle):
    """The ccc rule over an angle list reproduces find_template's maps; with
    all three scores a voxel's recorded maps are those of its recorded angle,
    and each recorded update improved every score strictly."""
    rng = np.random.default_rng(10)
    n, N = 8, (20, 18, 16)
    T = rng.standard_normal((n, n, n)).astype(np.float32)
    V = rng.standard_normal(N).astype(np.float32)
    M = oracle.sphere_mask_binary(n, 3.5)
    angles = np.array([[0, 0, 0], [30, 60, 90], [90, 30, 45], [10, 20, 30]], float)
    ref_ccc, ref_ang, allm = oracle.find_template(V, T, M, angles, return_all=True)
    res = {k: np.full(N, -1000.0) for k in ("ccc", "psr", "soc")}
    res["angle"] = np.zeros(N, np.int64)
    for i, s in enumerate(allm):
        oracle.bestpeak(res, s, None, None, i + 1, "ccc")
    np.testing.assert_array_equal(res["ccc"], ref_ccc)
    np.testing.assert_array_equal(res["angle"], ref_ang)
    maps = [tuple(rng.standard_normal(N) for _ in range(3)) for _ in angles]
    res = {k: np.full(N, -1000.0) for k in ("ccc", "psr", "soc")}
    res["angle"] = np.zeros(N, np.int64)
    for i, (c, p, q) in enumerate(maps):
        before = {k: res[k].copy() for k in ("ccc", "psr", "soc")}
        oracle.bestpeak(res, c, p, q, i + 1, "ccc+psr+soc")
        up = res["angle"] == i + 1
        assert np.all((before["ccc"] < c)[up] & (before["psr"] < p)[up] & (before["soc"] < q)[up])
        keep = ~up
        for k in ("ccc", "psr", "soc"):
            np.testing.assert_array_equal(res[k][keep], before[k][keep])
    for i, (c, p, q) in enumerate(maps):
        sel = res["angle"] == i + 1
        np.testing.assert_array_equal(res["ccc"][sel], c[sel])
        np.testing.assert_array_equal(res["soc"][sel], q[sel])
