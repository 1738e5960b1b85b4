"""Coarse orientation bins for the learned proposal, and the rotation helpers
they need. No counterpart in the reference path (its nearest relative is the
Thomson-repulsion sphere sampler Oscar3/c++Picker/src/matlab/tom_equi_rotation.m,
which samples directions, not rotations).

Rotations follow TOM's ZXZ convention: [phi psi theta] in degrees with
R = Rz(psi) Rx(theta) Rz(phi) (Sptrans/tom_sum_rotation.m:115-124). The bins
are `n` unit quaternions (q and -q the same rotation) spread over SO(3) by
repulsion: each step moves every quaternion away from its neighbours (and
their antipodes) with a force 1/d^2 in the chordal distance, then renormalises.
The result is deterministic for a given (n, seed, steps).
"""
import numpy as np


def rotmatrix(euler) -> np.ndarray:
    """(..., 3) [phi psi theta] degrees -> (..., 3, 3) Rz(psi) Rx(theta) Rz(phi)."""
    e = np.deg2rad(np.asarray(euler, np.float64))
    f, p, t = e[..., 0], e[..., 1], e[..., 2]
    cf, sf, cp, sp, ct, st = np.cos(f), np.sin(f), np.cos(p), np.sin(p), np.cos(t), np.sin(t)
    R = np.empty(e.shape[:-1] + (3, 3))
    R[..., 0, 0] = cp * cf - sp * ct * sf
    R[..., 0, 1] = -cp * sf - sp * ct * cf
    R[..., 0, 2] = sp * st
    R[..., 1, 0] = sp * cf + cp * ct * sf
    R[..., 1, 1] = -sp * sf + cp * ct * cf
    R[..., 1, 2] = -cp * st
    R[..., 2, 0] = st * sf
    R[..., 2, 1] = st * cf
    R[..., 2, 2] = ct
    return R


def euler_from_matrix(R) -> np.ndarray:
    """Inverse of rotmatrix: theta in [0, 180], phi and psi in [0, 360)."""
    R = np.asarray(R, np.float64)
    t = np.arccos(np.clip(R[..., 2, 2], -1.0, 1.0))
    st = np.sin(t)
    gen = st > 1e-9
    f = np.where(gen, np.arctan2(R[..., 2, 0], R[..., 2, 1]), 0.0)
    # degenerate theta (0 or 180): only psi + phi (or psi - phi) is defined;
    # with phi = 0 both cases give psi = atan2(R10, R00)
    p_deg = np.arctan2(R[..., 1, 0], R[..., 0, 0])
    p = np.where(gen, np.arctan2(R[..., 0, 2], -R[..., 1, 2]), p_deg)
    out = np.stack([np.rad2deg(f), np.rad2deg(p), np.rad2deg(t)], -1)
    out[..., :2] = np.mod(out[..., :2], 360.0)
    return out


def quat_from_matrix(R) -> np.ndarray:
    """(..., 3, 3) -> (..., 4) unit quaternions (w, x, y, z), w >= 0."""
    R = np.asarray(R, np.float64)
    tr = R[..., 0, 0] + R[..., 1, 1] + R[..., 2, 2]
    # Shepperd's method: pick the largest of 4 candidate squared components
    c = np.stack([1 + tr, 1 + R[..., 0, 0] - R[..., 1, 1] - R[..., 2, 2],
                  1 - R[..., 0, 0] + R[..., 1, 1] - R[..., 2, 2], 1 - R[..., 0, 0] - R[..., 1, 1] + R[..., 2, 2]], -1)
    k = np.argmax(c, -1)
    q = np.empty(R.shape[:-2] + (4,))
    s = np.sqrt(np.maximum(np.take_along_axis(c, k[..., None], -1)[..., 0], 1e-300)) * 2
    r = R
    cand = [
        np.stack([s / 4, (r[..., 2, 1] - r[..., 1, 2]) / s, (r[..., 0, 2] - r[..., 2, 0]) / s,
                  (r[..., 1, 0] - r[..., 0, 1]) / s], -1),
        np.stack([(r[..., 2, 1] - r[..., 1, 2]) / s, s / 4, (r[..., 0, 1] + r[..., 1, 0]) / s,
                  (r[..., 0, 2] + r[..., 2, 0]) / s], -1),
        np.stack([(r[..., 0, 2] - r[..., 2, 0]) / s, (r[..., 0, 1] + r[..., 1, 0]) / s, s / 4,
                  (r[..., 1, 2] + r[..., 2, 1]) / s], -1),
        np.stack([(r[..., 1, 0] - r[..., 0, 1]) / s, (r[..., 0, 2] + r[..., 2, 0]) / s,
                  (r[..., 1, 2] + r[..., 2, 1]) / s, s / 4], -1),
    ]
    for i in range(4):
        q = np.where((k == i)[..., None], cand[i], q)
    q /= np.linalg.norm(q, axis=-1, keepdims=True)
    return np.where(q[..., :1] < 0, -q, q)


def matrix_from_quat(q) -> np.ndarray:
    q = np.asarray(q, np.float64)
    q = q / np.linalg.norm(q, axis=-1, keepdims=True)
    w, x, y, z = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    R = np.empty(q.shape[:-1] + (3, 3))
    R[..., 0, 0] = 1 - 2 * (y * y + z * z)
    R[..., 0, 1] = 2 * (x * y - w * z)
    R[..., 0, 2] = 2 * (x * z + w * y)
    R[..., 1, 0] = 2 * (x * y + w * z)
    R[..., 1, 1] = 1 - 2 * (x * x + z * z)
    R[..., 1, 2] = 2 * (y * z - w * x)
    R[..., 2, 0] = 2 * (x * z - w * y)
    R[..., 2, 1] = 2 * (y * z + w * x)
    R[..., 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def rotation_distance_deg(qa, qb) -> np.ndarray:
    """Angle of the rotation taking one orientation to the other, degrees."""
    d = np.abs(np.sum(np.asarray(qa) * np.asarray(qb), -1))
    return np.rad2deg(2 * np.arccos(np.clip(d, 0.0, 1.0)))


def orientation_bins(n: int = 64, seed: int = 0, steps: int = 400) -> np.ndarray:
    """(n, 3) [phi psi theta] bin centres spread over SO(3)."""
    rng = np.random.default_rng(seed)
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    eye = np.eye(n, dtype=bool)
    for it in range(steps):
        force = np.zeros_like(q)
        for sgn in (1.0, -1.0):
            d = q[:, None, :] - sgn * q[None, :, :]
            r2 = np.sum(d * d, -1)
            if sgn > 0:
                r2[eye] = np.inf
            force += np.sum(d / (r2[..., None] ** 1.5 + 1e-12), 1)
        # tangential part only, step shrinking over time
        force -= np.sum(force * q, 1, keepdims=True) * q
        fn = np.linalg.norm(force, axis=1, keepdims=True)
        step = 0.05 * (1 - it / steps) + 0.002
        q = q + step * force / (fn.max() + 1e-300)
        q /= np.linalg.norm(q, axis=1, keepdims=True)
    q = np.where(q[:, :1] < 0, -q, q)
    return euler_from_matrix(matrix_from_quat(q))


def super_fibonacci(n: int) -> np.ndarray:
    """(n, 3) [phi psi theta] of n orientations from the super-Fibonacci
    spiral (M. Alexa, "Super-Fibonacci Spirals: Fast, Low-Discrepancy
    Sampling of SO(3)", CVPR 2022): quaternion i = (r sin a, r cos a,
    R sin b, R cos b) with s = i + 1/2, r = sqrt(s/n), R = sqrt(1 - s/n),
    a = 2 pi s / sqrt(2), b = 2 pi s / psi, psi^4 = psi + 4. O(n), so
    thousands of near-uniform orientations come without an iterative spread."""
    psi = 1.533751168755204288118041
    s = np.arange(int(n), dtype=np.float64) + 0.5
    r, R = np.sqrt(s / n), np.sqrt(1.0 - s / n)
    a, b = 2.0 * np.pi * s / np.sqrt(2.0), 2.0 * np.pi * s / psi
    q = np.stack([r * np.sin(a), r * np.cos(a), R * np.sin(b), R * np.cos(b)], 1)
    return euler_from_matrix(matrix_from_quat(q)).astype(np.float64)


def assign_bins(euler, bins) -> np.ndarray:
    """Index of the nearest bin (largest |q . q_bin|) for each [phi psi theta]."""
    qa = quat_from_matrix(rotmatrix(np.asarray(euler, np.float64).reshape(-1, 3)))
    qb = quat_from_matrix(rotmatrix(np.asarray(bins, np.float64).reshape(-1, 3)))
    return np.argmax(np.abs(qa @ qb.T), axis=1).astype(np.int32)


def local_grid(center, half: float = 15.0, step: float = 5.0) -> np.ndarray:
    """[phi psi theta] grid center +- half in steps of `step` on each Euler
    angle (phi outermost, theta innermost, as tom_os3_angleList orders)."""
    return local_grids(np.asarray(center, np.float64).reshape(1, 3), half, step)[:, 0]


def local_grids(centers, half: float = 15.0, step: float = 5.0) -> np.ndarray:
    """local_grid of every row of centers (P, 3), as one (A, P, 3) array."""
    off = np.arange(-half, half + 1e-9, step)
    a, b, t = np.meshgrid(off, off, off, indexing="ij")  # phi outermost, theta innermost
    d = np.stack([a.ravel(), b.ravel(), t.ravel()], 1)  # (A, 3)
    g = np.asarray(centers, np.float64).reshape(1, -1, 3) + d[:, None, :]
    g[..., :2] = np.mod(g[..., :2], 360.0)
    return g.astype(np.float32)
