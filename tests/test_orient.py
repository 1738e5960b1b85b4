"""Orientation helpers and the 64 coarse bins of the learned proposal.

Pins: rotmatrix equals the reference-convention matrix (oracle.rotmat, the
restatement of tom_rotatec's rotation matrix, stored in column order) and
tom_amd.api's tom_sum_rotation restatement; Euler and quaternion round trips;
bin invariants (unit quaternions, deterministic, near-uniform nearest-neighbour
spacing, covering radius well below that of random rotations)."""
import numpy as np

from tom_amd import orient as O
from tom_amd.api import _rotmatrix


def _rand_euler(rng, k):
    return np.stack([rng.uniform(0, 360, k), rng.uniform(0, 360, k), rng.uniform(0, 180, k)], 1)


def test_rotmatrix_matches_reference_conventions(oracle):
    rng = np.random.default_rng(0)
    for e in _rand_euler(rng, 50):
        R = O.rotmatrix(e)
        np.testing.assert_allclose(R, _rotmatrix(*e), atol=1e-14)
        np.testing.assert_allclose(R.ravel(order="F"), oracle.rotmat(e.astype(np.float32)), atol=2e-6)


def test_round_trips():
    rng = np.random.default_rng(1)
    e = _rand_euler(rng, 500)
    e[:5, 2] = [0, 180, 0, 180, 90]  # degenerate theta
    R = O.rotmatrix(e)
    np.testing.assert_allclose(O.rotmatrix(O.euler_from_matrix(R)), R, atol=1e-12)
    q = O.quat_from_matrix(R)
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1, atol=1e-14)
    np.testing.assert_allclose(O.matrix_from_quat(q), R, atol=1e-12)
    # distance: a rotation about z by 30 degrees is 30 degrees away
    a = O.quat_from_matrix(O.rotmatrix([10, 0, 0]))
    b = O.quat_from_matrix(O.rotmatrix([40, 0, 0]))
    assert abs(O.rotation_distance_deg(a, b) - 30) < 1e-9


def test_bins_invariants():
    b = O.orientation_bins(64, 0)
    assert b.shape == (64, 3)
    np.testing.assert_array_equal(b, O.orientation_bins(64, 0))
    q = O.quat_from_matrix(O.rotmatrix(b))
    D = O.rotation_distance_deg(q[:, None], q[None])
    np.fill_diagonal(D, np.inf)
    nn = D.min(1)
    assert nn.min() / nn.max() > 0.75
    rng = np.random.default_rng(2)
    r = rng.standard_normal((50000, 4))
    r /= np.linalg.norm(r, axis=1, keepdims=True)
    cover = np.rad2deg(2 * np.arccos(np.clip(np.abs(r @ q.T).max(1), 0, 1)))
    qr = rng.standard_normal((64, 4))
    qr /= np.linalg.norm(qr, axis=1, keepdims=True)
    cover_rand = np.rad2deg(2 * np.arccos(np.clip(np.abs(r @ qr.T).max(1), 0, 1)))
    assert cover.max() < 52 and cover.max() < 0.8 * cover_rand.max()
    # every bin centre is its own nearest bin
    np.testing.assert_array_equal(O.assign_bins(b, b), np.arange(64))


def test_super_fibonacci_covers_so3():
    """The coarse-scan orientations: deterministic, distinct, and covering
    SO(3) more tightly than the same number of random rotations (largest
    distance from 50,000 random rotations to the nearest member)."""
    e = O.super_fibonacci(512)
    assert e.shape == (512, 3)
    np.testing.assert_array_equal(e, O.super_fibonacci(512))
    q = O.quat_from_matrix(O.rotmatrix(e))
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-12)
    D = O.rotation_distance_deg(q[:, None], q[None])
    np.fill_diagonal(D, np.inf)
    assert D.min() > 3.0
    rng = np.random.default_rng(5)
    r = rng.standard_normal((50000, 4))
    r /= np.linalg.norm(r, axis=1, keepdims=True)
    cover = np.rad2deg(2 * np.arccos(np.clip(np.abs(r @ q.T).max(1), 0, 1)))
    qr = rng.standard_normal((512, 4))
    qr /= np.linalg.norm(qr, axis=1, keepdims=True)
    cover_rand = np.rad2deg(2 * np.arccos(np.clip(np.abs(r @ qr.T).max(1), 0, 1)))
    print(f"super-Fibonacci 512 covering {cover.max():.1f} deg, random {cover_rand.max():.1f} deg")
    assert cover.max() < 0.8 * cover_rand.max()


def test_local_grids_order_and_wrap():
    """(A, P, 3) grids: phi outermost, theta innermost, phi and psi wrapped
    into [0, 360), theta not wrapped (loops over the offsets)."""
    rng = np.random.default_rng(7)
    c = _rand_euler(rng, 9)
    c[0] = [355.0, 2.0, 90.0]
    g = O.local_grids(c, 15, 7.5)
    off = [-15.0, -7.5, 0.0, 7.5, 15.0]
    assert g.shape == (125, 9, 3)
    for p in range(9):
        ref = [((c[p][0] + a) % 360, (c[p][1] + b) % 360, c[p][2] + t) for a in off for b in off for t in off]
        np.testing.assert_allclose(g[:, p], np.asarray(ref, np.float32), atol=1e-4)
        np.testing.assert_array_equal(g[:, p], O.local_grid(c[p], 15, 7.5))


def test_local_grid():
    g = O.local_grid([350, 10, 90], 15, 5)
    assert g.shape == (343, 3)
    assert np.any(np.all(np.isclose(g, [350, 10, 90]), 1))
    assert g[:, :2].min() >= 0 and g[:, :2].max() < 360
    # phi outermost, theta innermost
    assert g[1, 2] - g[0, 2] == 5 and g[0, 0] == g[48, 0]


def test_template_centroid_rotates_with_the_template(oracle):
    """The centre of mass of tom_rotate(T, e) sits at R(e) d from the rotation
    centre, d = template_centroid(T): the convention the proposal's centroid
    targets and the refinement's windows rely on."""
    from tom_amd import surrogate, synth
    rng = np.random.default_rng(2)
    T = synth.phantom(rng, 32, 6, 2.0, 9)
    d = surrogate.template_centroid(T)
    assert np.linalg.norm(d) > 1.0
    for e in ([30, 50, 70], [100, 200, 40], [0, 0, 90], [250, 10, 135]):
        got = surrogate.template_centroid(oracle.rotate(T, e))
        np.testing.assert_allclose(got, O.rotmatrix(np.array(e, float)) @ d, atol=0.05)
    pk = np.array([[10, 20, 30, 0.5, 1]], float)
    ang = np.array([[0, 0, 0], [30, 50, 70]], float)
    np.testing.assert_allclose(surrogate.centroid_positions(pk, ang, d)[0],
                               [10, 20, 30] + O.rotmatrix(ang[1]) @ d)
