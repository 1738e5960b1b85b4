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
def test_norm_mask_direct_sums(oracle):
    rng = np.random.default_rng(0)
    v = rng.standard_normal((8, 8, 8)) + 2
    M = rng.uniform(0, 1, (8, 8, 8)) * (rng.uniform(0, 1, (8, 8, 8)) > 0.4)
    out, var = oracle.norm_mask(v, M, False)
    B = M != 0
    mean = v[B].mean()
    v2 = np.where(B, (v - mean) * M, 0)
    m2 = v2[B].mean()
    assert abs(var - (((v2[B]) ** 2).sum() - B.sum() * m2 * m2) / v.size) < 1e-12
    np.testing.assert_allclose(out[B].sum(), 0, atol=1e-9)
    assert np.all(out[~B] == 0)


def test_subtomo_oracle_planted_copy_scores_one(oracle):
    """No wedge, binary mask, particle = the rotated template itself: the
    true angle and zero shift score exactly 1."""
    T = _phantom(32, 1)
    M = oracle.sphere_mask_binary(32, 14)
    a = np.array([[30.0, 60.0, 90.0]])
    part = oracle.rotate(T - T.mean(), a[0])
    i, pos, val, _ = oracle.subtomo_score(part, T, M, None if False else np.ones((32,) * 3), 0.0,
                                          np.array([[0, 0, 0], [30, 60, 90], [40, 60, 90]], float))
    assert i == 1 and list(pos) == [0, 0, 0]
    assert abs(val - 1.0) < 1e-9


def test_subtomo_oracle_correlation_direct(oracle):
    """The FFT correlation of the oracle equals the direct circular sum
    sum_x tw[x] p[x + s] on a tiny box (no wedge)."""
    rng = np.random.default_rng(3)
    n = 6
    T = rng.standard_normal((n, n, n)).astype(np.float32)
    p = rng.standard_normal((n, n, n)).astype(np.float32)
    M = np.ones((n, n, n), np.float32)
    _, _, _, vals = oracle.subtomo_score(p, T, M, np.ones((n, n, n)), 0.0, np.zeros((1, 3)))
    t, vt = oracle.norm_mask(oracle.rotate(T - T.mean(), [0, 0, 0]), M, False)
    q, vp = oracle.norm_mask(p, M, False)
    best = -1e30
    for s in np.ndindex(n, n, n):
        c = sum(t[x] * q[tuple((np.array(x) + s) % n)] for x in np.ndindex(n, n, n))
        best = max(best, c / (np.sqrt(vp * vt) * t.size))
    assert abs(vals[0] - best) < 1e-9


def test_simple_wedge_predicate(oracle):
    W = oracle.simple_wedge_half((16, 16, 16), 30)
    f = np.fft.fftfreq(16, 1 / 16)
    X, Y, Z = np.meshgrid(f, f, f, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    assert np.all(W[r > 7] == 0)                 # cutoff (16-1)//2 = 7
    assert np.all(W[(Z == 0) & (r <= 7)] == 1)   # kz = 0 kept
    sel = (Z != 0) & (r <= 7)
    exp = np.tan(np.pi / 6) <= np.abs(X[sel]) / np.abs(Z[sel]) + 1e-7
    assert np.mean(W[sel] == exp) > 0.99


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
            semi.append([0.0, 30.0, 45.0][k])
            angs.append(grid)
            meta.append((T, M, Mcc))
    best, peak, ccc = api.subtomo_score_batch(boxes, parts, pbox, semi, np.stack(angs))
    for p in range(len(parts)):
        T, M, Mcc = meta[p]
        i, pos, val, vals = oracle.subtomo_score(parts[p], T, M, Mcc, semi[p], angs[p].astype(np.float64))
        srt = np.sort(vals)
        margin = srt[-1] - srt[-2]
        print(f"p{p} box {parts[p].shape[0]} wedge {semi[p]}: gpu ({best[p]}, {list(peak[p])}, {ccc[p]:.6f}) "
              f"oracle ({i}, {list(pos)}, {val:.6f}) margin {margin:.2e}", flush=True)
        assert abs(ccc[p] - val) < 1e-4
        if margin > 1e-4:
            assert best[p] == i and list(peak[p]) == list(pos)
