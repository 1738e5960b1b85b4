"""Batched subtomogram scoring (corr3d_batch semantics): oracle pins (CPU)
and GPU parity through the C ABI. Tolerances: CCC within 1e-4 of the FP64
oracle; best angle and peak identical where the oracle's best/second-best
margin exceeds 1e-4."""
import numpy as np
import pytest


def _phantom(n, seed):
    rng = np.random.default_rng(seed)
    x = np.arange(n) - n // 2
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    T = np.zeros((n, n, n))
    for _ in range(8):
        c = rng.uniform(-n / 5, n / 5, 3)
        T += rng.uniform(0.5, 1) * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / (2 * (n / 16) ** 2))
    return T.astype(np.float32)


# ---------------------------------------------------------------- oracle pins
def test_norm_mask_closed_forms(oracle):
    """tom::norm_mask against closed forms (volume_fcn.cpp:168-247).
    Binary mask over the lower half of the box, v = x: the masked values are
    the rows 0..n-1, population variance (n^2 - 1)/12; norm_mask's variance
    is that sum over the masked voxels divided by numel (not by the count).
    Soft mask constant along x and v = a + b (-1)^x: the in-mask mean is a,
    the output is b (-1)^x M and the variance b^2 sum(M^2) / numel."""
    n = 8
    x = np.arange(n, dtype=np.float64)
    v = np.broadcast_to(x[:, None, None], (n, n, n)).copy()
    M = np.zeros((n, n, n))
    M[:, :, : n // 2] = 1
    out, var = oracle.norm_mask(v, M, True)
    cnt = n * n * (n // 2)
    assert abs(var - cnt * (n * n - 1) / 12.0 / n ** 3) < 1e-12
    np.testing.assert_allclose(out[:, :, : n // 2], v[:, :, : n // 2] - (n - 1) / 2.0, atol=1e-12)
    assert np.all(out[:, :, n // 2:] == 0)
    rng = np.random.default_rng(4)
    Myz = rng.uniform(0, 1, (n, n)) * (rng.uniform(0, 1, (n, n)) > 0.3)
    Ms = np.broadcast_to(Myz[None], (n, n, n)).copy()
    a, b = 3.5, -1.25
    u = np.broadcast_to(((-1.0) ** x)[:, None, None], (n, n, n))
    out, var = oracle.norm_mask(a + b * u, Ms, False)
    np.testing.assert_allclose(out, b * u * Ms, atol=1e-12)
    assert abs(var - b * b * (Ms ** 2).sum() / n ** 3) < 1e-12
    out, var = oracle.norm_mask(np.full((n, n, n), 7.0), Ms, False)
    assert var == 0 and np.all(out == 0)


def test_subtomo_oracle_planted_copy_scores_one(oracle):
    """No wedge, binary mask, particle = the rotated template itself: the
    true angle and zero shift score exactly 1."""
    T = _phantom(32, 1)
    M = oracle.sphere_mask_binary(32, 14)
    a = np.array([[30.0, 60.0, 90.0]])
    part = oracle.rotate(T - T.mean(), a[0])
    i, pos, val, _ = oracle.subtomo_score(part, T, M, np.ones((32,) * 3), None,
                                          np.array([[0, 0, 0], [30, 60, 90], [40, 60, 90]], float))
    assert i == 1 and list(pos) == [0, 0, 0]
    assert abs(val - 1.0) < 1e-9


def _wedge64(n, semi):
    """Missing-wedge predicate in float64 from the definition (cutoff sphere
    of radius (n-1)//2, kz = 0 kept, keep tan(semi) <= |kx|/|kz|)."""
    f = np.fft.fftfreq(n, 1.0 / n)
    X, Y, Z = np.meshgrid(f, f, f, indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z)
    with np.errstate(divide="ignore", invalid="ignore"):
        keep = (r <= (n - 1) // 2) & ((Z == 0) | (np.tan(np.radians(semi)) <= np.abs(X) / np.abs(Z)))
    return keep.astype(np.float64)


def test_subtomo_oracle_particle_wedge_planted_scores_one(oracle):
    """Particle-wedge branch (corr.cpp:1146-1171) by its closed form: with a
    binary mask, a particle equal to the wedged, boolean-re-masked, mean-free
    rotated template scores exactly 1 at the true angle and zero shift.
    The particle is built here from numpy FFTs and the float64 predicate; at
    30 degrees no grid point lies on the wedge boundary, so the oracle's
    float predicate selects the same frequencies."""
    n = 32
    T = _phantom(n, 2).astype(np.float64)
    M = oracle.sphere_mask_binary(n, 13).astype(np.float64)
    B = M != 0
    true = np.array([40.0, 70.0, 110.0])
    t = oracle.rotate((T - T.mean()).astype(np.float32), true).astype(np.float64)
    t = np.where(B, t - t[B].mean(), 0.0)
    tw = np.real(np.fft.ifftn(np.fft.fftn(t) * _wedge64(n, 30.0)))
    tw = np.where(B, tw - tw[B].mean(), 0.0)
    angles = np.array([[0, 0, 0], true, true + [5, 0, 0], true + [0, 0, 5]], float)
    i, pos, val, vals = oracle.subtomo_score(tw.astype(np.float32), T, M, np.ones((n,) * 3), 30.0, angles)
    assert i == 1 and list(pos) == [0, 0, 0]
    assert abs(val - 1.0) < 1e-9
    assert np.all(np.delete(vals, 1) < 1.0 - 1e-3)
    # without re-estimating the template's variance after the wedge the same
    # particle scores below 1: the pin sees a dropped variance update
    Tf = np.fft.fftn(t)
    raw = np.real(np.fft.ifftn(Tf * _wedge64(n, 30.0)))
    assert np.sum(raw * tw) / np.sqrt(np.sum(t * t) * np.sum(tw * tw)) < 0.99


def test_subtomo_oracle_correlation_direct(oracle):
    """The FFT correlation of the oracle equals the direct circular sum
    sum_x tw[x] p[x + s] on a tiny box (no wedge)."""
    rng = np.random.default_rng(3)
    n = 6
    T = rng.standard_normal((n, n, n)).astype(np.float32)
    p = rng.standard_normal((n, n, n)).astype(np.float32)
    M = np.ones((n, n, n), np.float32)
    _, _, _, vals = oracle.subtomo_score(p, T, M, np.ones((n, n, n)), None, np.zeros((1, 3)))
    t, vt = oracle.norm_mask(oracle.rotate(T - T.mean(), [0, 0, 0]), M, False)
    q, vp = oracle.norm_mask(p, M, False)
    best = -1e30
    for s in np.ndindex(n, n, n):
        c = sum(t[x] * q[tuple((np.array(x) + s) % n)] for x in np.ndindex(n, n, n))
        best = max(best, c / (np.sqrt(vp * vt) * t.size))
    assert abs(vals[0] - best) < 1e-9


def test_simple_wedge_known_points(oracle):
    """SimpleWedge about Y with semi-angle 30 deg at hand-checked frequencies:
    a swapped axis, the full instead of the semi-angle, or a dropped cutoff
    or kz = 0 rule fails one of them."""
    n = 16
    W = oracle.simple_wedge((n, n, n), 30.0)

    def w(kx, ky, kz):
        return W[kx % n, ky % n, kz % n]
    assert w(1, 0, 1) == 1       # |kx|/|kz| = 1 >= tan 30
    assert w(1, 0, 2) == 0       # 0.5 < tan 30 = 0.577
    assert w(2, 0, 3) == 1       # 0.667
    assert w(0, 5, 1) == 0       # independent of ky
    assert w(0, 1, 1) == 0       # y is not the tilt axis' partner: swapped x/y keeps it
    assert w(1, 2, 0) == 1       # kz = 0 plane kept (swapped y/z removes it)
    assert w(-3, 2, 5) == 1      # 0.6 >= 0.577 (r = 6.2); 60 deg would remove it
    assert w(4, 4, 4) == 1 and w(7, 0, 0) == 1 and w(4, 4, 5) == 0  # r = 6.9, 7, 7.5 against cutoff 7
    assert w(0, 0, 0) == 1
    W60 = oracle.simple_wedge((n, n, n), 60.0)
    assert W60[3 % n, 0, 5] == 0 and W60[2, 0, 1] == 1
    # the predicate is symmetric under k -> -k on the whole grid
    np.testing.assert_array_equal(W, np.roll(W[::-1, ::-1, ::-1], 1, axis=(0, 1, 2)))
    # equals the float64 definition wherever no rounding is involved
    np.testing.assert_array_equal(W, _wedge64(n, 30.0))


def test_simple_wedge_rotation(oracle):
    """The rotated wedge of corr3d_single (wedge.cpp:149-224 with
    init_rotmat z,x,z) is W(R^T k) for the rotation R = Rz(psi) Rx(theta)
    Rz(phi) that tom_rotate applies to the object: the wedge turns with the
    template. Checked against that formula in float64 at every frequency
    whose predicate is not within rounding of a boundary."""
    n = 24

    def Rz(a):
        c, s = np.cos(a), np.sin(a)
        return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])

    def Rx(a):
        c, s = np.cos(a), np.sin(a)
        return np.array([[1.0, 0, 0], [0, c, -s], [0, s, c]])
    f = np.fft.fftfreq(n, 1.0 / n)
    K = np.stack(np.meshgrid(f, f, f, indexing="ij"), -1).reshape(-1, 3)
    np.testing.assert_array_equal(oracle.simple_wedge((n,) * 3, 35.0, [0, 0, 0]), oracle.simple_wedge((n,) * 3, 35.0))
    for eul in ([90.0, 0.0, 0.0], [0.0, 0.0, 45.0], [30.0, 70.0, 110.0], [200.0, 10.0, 65.0]):
        phi, psi, the = np.radians(eul)
        R = Rz(psi) @ Rx(the) @ Rz(phi)
        Kp = K @ R  # rows: R^T k
        ax, az = np.abs(Kp[:, 0]), np.abs(Kp[:, 2])
        r = np.linalg.norm(Kp, axis=1)
        tan_a = np.tan(np.radians(35.0))
        with np.errstate(divide="ignore", invalid="ignore"):
            ratio = np.where(az > 0, ax / az, np.inf)
        keep = (r <= (n - 1) // 2) & ((az < 1e-4) | (tan_a <= ratio))
        clear = (np.abs(r - (n - 1) // 2) > 1e-4) & ((az < 1e-5) | ((az > 2e-4) & (np.abs(ratio - tan_a) > 1e-4)))
        W = oracle.simple_wedge((n,) * 3, 35.0, eul).reshape(-1)
        assert clear.mean() > 0.99
        np.testing.assert_array_equal(W[clear], keep[clear].astype(float), err_msg=str(eul))
    # sense: a 45 deg theta turn keeps (1, 2, 2) (it lands on kz' = 0) and
    # removes (1, -2, 2) (kz' = 2.83, |kx'|/|kz'| = 0.35 < tan 35)
    W = oracle.simple_wedge((n,) * 3, 35.0, [0, 0, 45])
    assert W[1, 2, 2] == 1 and W[1, n - 2, 2] == 0


def test_subtomo_oracle_template_wedge_swap_symmetry(oracle):
    """Template-wedge branch (corr.cpp:1086-1111) against the pinned
    particle-wedge branch: at angle 0 the two roles swap. Scoring particle X
    against template T with the particle's wedge gives the map c(s); scoring
    particle T against template X with the template's wedge gives c(-s).
    Same best value, peak at the negated shift."""
    n = 32
    T = _phantom(n, 3)
    rng = np.random.default_rng(8)
    X = (2.0 * oracle.rotate(T, [10, 20, 30]) + rng.standard_normal((n, n, n))).astype(np.float32)
    X = np.roll(X, (2, -1, 3), axis=(0, 1, 2))
    M = oracle.sphere_mask_binary(n, 12)
    ones = np.ones((n,) * 3)
    a0 = np.zeros((1, 3))
    ia, pa, va, _ = oracle.subtomo_score(X, T, M, ones, 45.0, a0)
    ib, pb, vb, _ = oracle.subtomo_score(T, X, M, ones, None, a0, wedge_ref_deg=45.0)
    assert abs(va - vb) < 1e-9
    assert list((-pa) % n) == list(pb)
    # a template wedge that is ignored would not reproduce the swapped map
    _, _, vn, _ = oracle.subtomo_score(T, X, M, ones, None, a0)
    assert abs(vn - va) > 1e-3


def test_subtomo_oracle_template_wedge_turns_with_the_template(oracle, monkeypatch):
    """Scoring at the 90-degree turn a = [90 0 0] (an exact voxel permutation
    about the centre voxel) equals scoring the particle turned back by a^-1
    at angle 0: the template's wedge is rotated with the template. The same
    check with the wedge left unrotated (planted in the oracle) fails."""
    n = 32
    T = _phantom(n, 5)
    rng = np.random.default_rng(9)
    part = (oracle.rotate(T, [90, 0, 0]) + 0.7 * rng.standard_normal((n, n, n))).astype(np.float32)
    M = oracle.sphere_mask_binary(n, 12)
    ones = np.ones((n,) * 3)
    back = oracle.rotate(part, [270, 0, 0])
    _, p1, v1, _ = oracle.subtomo_score(part, T, M, ones, None, np.array([[90.0, 0, 0]]), wedge_ref_deg=30.0)
    _, p2, v2, _ = oracle.subtomo_score(back, T, M, ones, None, np.array([[0.0, 0, 0]]), wedge_ref_deg=30.0)
    assert abs(v1 - v2) < 1e-9
    # the peak shift turns with it: s1 = R s2 for R the 90-degree turn about z
    s1 = ((p1 + n // 2) % n) - n // 2
    s2 = ((p2 + n // 2) % n) - n // 2
    assert list(s1) == [-s2[1], s2[0], s2[2]]
    plain = oracle.simple_wedge
    monkeypatch.setattr(oracle, "simple_wedge", lambda shape, semi, euler=None: plain(shape, semi))
    _, _, vbad, _ = oracle.subtomo_score(part, T, M, ones, None, np.array([[90.0, 0, 0]]), wedge_ref_deg=30.0)
    assert abs(vbad - v2) > 1e-3


# ---------------------------------------------------------------- GPU parity
def _case(oracle, rng, n, true, shift, snr, semi):
    T = _phantom(n, n)
    sig = oracle.rotate(T, true)
    sig = np.roll(sig, shift, axis=(0, 1, 2))
    noise = rng.standard_normal((n, n, n)).astype(np.float32)
    s = np.sqrt(snr / sig.var())
    return T, (sig * s + noise).astype(np.float32)


@pytest.mark.gpu
def test_subtomo_gpu_matches_oracle(oracle):
    from tom_amd import api
    rng = np.random.default_rng(11)
    boxes, parts, pbox, semi, angs, meta = [], [], [], [], [], []
    for bi, n in enumerate((32, 48, 64)):
        T = _phantom(n, n)
        M = oracle.sphere_mask_gauss(n, n / 2 - 4, 2)  # soft mask
        Mcc = oracle.mask_cc_sphere((n, n, n), 5)
        boxes.append({"template": T, "mask": M, "mask_cc": Mcc})
        for k in range(3):
            true = rng.uniform(0, 360, 3)
            shift = tuple(int(x) for x in rng.integers(-3, 4, 3))
            _, part = _case(oracle, rng, n, true, shift, [0.5, 0.1, 5.0][k], 0)
            prior = true + rng.uniform(-4, 4, 3)
            grid = np.array([[prior[0] + i * 5, prior[1] + j * 5, prior[2] + l * 5]
                             for i in (-1, 0, 1) for j in (-1, 0, 1) for l in (-1, 0, 1)], np.float32)
            parts.append(part)
            pbox.append(bi)
            semi.append([-1.0, 30.0, 45.0][k])
            angs.append(grid)
            meta.append((T, M, Mcc))
    best, peak, ccc = api.subtomo_score_batch(boxes, parts, pbox, semi, np.stack(angs))
    for p in range(len(parts)):
        T, M, Mcc = meta[p]
        i, pos, val, vals = oracle.subtomo_score(parts[p], T, M, Mcc, semi[p] if semi[p] >= 0 else None,
                                                 angs[p].astype(np.float64))
        srt = np.sort(vals)
        margin = srt[-1] - srt[-2]
        print(f"p{p} box {parts[p].shape[0]} wedge {semi[p]}: gpu ({best[p]}, {list(peak[p])}, {ccc[p]:.6f}) "
              f"oracle ({i}, {list(pos)}, {val:.6f}) margin {margin:.2e}", flush=True)
        assert abs(ccc[p] - val) < 1e-4
        if margin > 1e-4:
            assert best[p] == i and list(peak[p]) == list(pos)


@pytest.mark.gpu
def test_subtomo_gpu_template_wedge_matches_oracle(oracle):
    """WEDGE_REF (the template's wedge turned with each orientation and
    applied to the particle, corr.cpp:1086-1111) alone and together with the
    particle's wedge, including a scalar 0 (the cutoff sphere only)."""
    from tom_amd import api
    rng = np.random.default_rng(21)
    boxes, parts, pbox, semi, angs, meta = [], [], [], [], [], []
    for bi, (n, wref) in enumerate(((32, 30.0), (48, 40.0), (64, 25.0), (32, 0.0))):
        T = _phantom(n, n + bi)
        M = oracle.sphere_mask_gauss(n, n / 2 - 4, 2) if bi != 3 else oracle.sphere_mask_binary(n, n / 2 - 3)
        Mcc = oracle.mask_cc_sphere((n, n, n), 4)
        boxes.append({"template": T, "mask": M, "mask_cc": Mcc, "wedge_ref": wref})
        for k in range(2):
            true = rng.uniform(0, 360, 3)
            shift = tuple(int(x) for x in rng.integers(-3, 4, 3))
            _, part = _case(oracle, rng, n, true, shift, [0.5, 2.0][k], 0)
            prior = true + rng.uniform(-4, 4, 3)
            grid = np.array([[prior[0] + i * 5, prior[1] + j * 5, prior[2] + l * 5]
                             for i in (-1, 0, 1) for j in (-1, 0, 1) for l in (-1, 0, 1)], np.float32)
            parts.append(part)
            pbox.append(bi)
            semi.append([-1.0, 35.0][k])
            angs.append(grid)
            meta.append((T, M, Mcc, wref))
    best, peak, ccc = api.subtomo_score_batch(boxes, parts, pbox, semi, np.stack(angs))
    for p in range(len(parts)):
        T, M, Mcc, wref = meta[p]
        i, pos, val, vals = oracle.subtomo_score(parts[p], T, M, Mcc, semi[p] if semi[p] >= 0 else None,
                                                 angs[p].astype(np.float64), wedge_ref_deg=wref)
        srt = np.sort(vals)
        margin = srt[-1] - srt[-2]
        print(f"p{p} box {parts[p].shape[0]} wedge_ref {wref} wedge {semi[p]}: gpu ({best[p]}, {list(peak[p])}, "
              f"{ccc[p]:.6f}) oracle ({i}, {list(pos)}, {val:.6f}) margin {margin:.2e}", flush=True)
        assert abs(ccc[p] - val) < 1e-4
        if margin > 1e-4:
            assert best[p] == i and list(peak[p]) == list(pos)
