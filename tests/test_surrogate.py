"""Host logic of the learned proposal (labels, greedy suppression, recall) and
the GPU candidate extraction against brute force."""
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
    tp, tb = S.label_volumes((10, 12, 14), peaks, angles, bins, radius=2.0)
    b = orient.assign_bins(angles, bins)
    assert tp[7, 6, 5] == 1 and tb[7, 6, 5] == b[1]
    assert tp[7, 6, 7] == 1 and tp[7, 6, 8] == 0  # radius 2 along x
    assert tp[0, 0, 0] == 1 and tb[0, 0, 0] == b[0]
    assert tp.sum() == (tb >= 0).sum()
    # a sphere of radius 2 has 33 voxels; at the corner one octant (11) remains
    assert tp.sum() == 33 + 11


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
