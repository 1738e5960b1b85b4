"""Write golden rotation vectors from the reference's own C rotation core.

Runs only in a container that has the reference tree: it builds
``oracle/_ref/libtomref.so`` from the reference sources with
``oracle/Makefile.ref`` and calls its ``rot3dlin`` (same arithmetic as
``Sptrans/tom_rotatec.c`` ``rot3d``; centre = ceil(n/2), 0-based, as passed by
``Sptrans/tom_rotate.m:105-108``).

Output: ``tests/golden/rot3d_golden.npz`` with, per case, the input volume
index, the Euler triple (or 9 matrix entries), the centre, the SHA-256 of the
output bytes (x-fastest) and, for a few cases, the full output volume.

    python tests/golden/make_rot3d_golden.py
"""
import ctypes
import hashlib
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libtomref.so")


def load_ref():
    if not os.path.exists(REF_SO):
        subprocess.check_call(["make", "-f", "oracle/Makefile.ref"], cwd=ROOT)
    L = ctypes.CDLL(REF_SO)
    fp = ctypes.POINTER(ctypes.c_float)
    L.rot3dlin.argtypes = [fp, fp, ctypes.c_long, ctypes.c_long, ctypes.c_long, fp,
                           ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int]
    L.rot3dlin.restype = None
    return L


def ref_rotate(L, vol, euler, center):
    v = np.asfortranarray(vol, dtype=np.float32)
    out = np.zeros_like(v, order="F")
    e = np.ascontiguousarray(np.asarray(euler, np.float32).ravel())
    euler_dim = 1 if e.size == 3 else 3
    fp = ctypes.POINTER(ctypes.c_float)
    L.rot3dlin(v.ctypes.data_as(fp), out.ctypes.data_as(fp), v.shape[0], v.shape[1], v.shape[2],
               e.ctypes.data_as(fp), float(center[0]), float(center[1]), float(center[2]), euler_dim)
    return out


def cases(rng):
    """Euler triples: the C1 30-degree grid (exact-table hits), 15-degree and
    10-degree grid samples, random non-table angles, negatives and >= 360."""
    out = []
    g30 = [(a, b, c) for a in range(0, 360, 30) for b in range(0, 360, 30) for c in range(0, 181, 30)]
    idx = rng.choice(len(g30), 60, replace=False)
    out += [g30[i] for i in sorted(idx)]
    out += [(15 * rng.integers(0, 24), 15 * rng.integers(0, 24), 15 * rng.integers(0, 13)) for _ in range(30)]
    out += [(10 * rng.integers(0, 36), 10 * rng.integers(0, 36), 10 * rng.integers(0, 19)) for _ in range(20)]
    out += [tuple(np.round(rng.uniform(-30, 390, 3), 3)) for _ in range(30)]
    out += [(0, 0, 0), (90, 0, 0), (0, 0, 90), (270, 90, 37), (360, 0, 0), (-90, 45, 180), (0, 0, 45.5)]
    return [tuple(float(x) for x in c) for c in out]


def main():
    L = load_ref()
    rng = np.random.default_rng(20240601)
    vols = {
        24: rng.standard_normal((24, 24, 24)).astype(np.float32),
        25: rng.standard_normal((25, 25, 25)).astype(np.float32),
        48: rng.standard_normal((48, 48, 48)).astype(np.float32),
        # a non-cubic box exercises the per-axis bounds and border taps
        2030: rng.standard_normal((20, 30, 16)).astype(np.float32),
    }
    rec_vol, rec_euler, rec_neuler, rec_center, rec_hash = [], [], [], [], []
    full_out, full_idx = [], []
    keys = sorted(vols)
    for key in keys:
        v = vols[key]
        cs = cases(rng)
        if key == 48:
            cs = cs[::4]
        for ci, e in enumerate(cs):
            center = [int(np.ceil(s / 2)) for s in v.shape]
            o = ref_rotate(L, v, e, center)
            rec_vol.append(key)
            rec_euler.append(list(e) + [0.0] * 6)
            rec_neuler.append(3)
            rec_center.append(center)
            rec_hash.append(hashlib.sha256(o.tobytes(order="F")).hexdigest())
            if key == 24 and ci % 25 == 3:
                full_out.append(o)
                full_idx.append(len(rec_hash) - 1)
        # two 9-entry matrix cases and one off-centre pivot per volume
        for e9 in (rng.standard_normal(9), np.eye(3).ravel(order="F")):
            q, _ = np.linalg.qr(np.asarray(e9).reshape(3, 3))
            m9 = np.asarray(q, np.float32).ravel(order="F")
            center = [int(np.ceil(s / 2)) for s in v.shape]
            o = ref_rotate(L, v, m9, center)
            rec_vol.append(key)
            rec_euler.append(list(m9.astype(np.float64)))
            rec_neuler.append(9)
            rec_center.append(center)
            rec_hash.append(hashlib.sha256(o.tobytes(order="F")).hexdigest())
        center = [3.5, 7.0, 5.25]
        o = ref_rotate(L, v, (33.0, 71.0, 12.0), center)
        rec_vol.append(key)
        rec_euler.append([33.0, 71.0, 12.0] + [0.0] * 6)
        rec_neuler.append(3)
        rec_center.append(center)
        rec_hash.append(hashlib.sha256(o.tobytes(order="F")).hexdigest())

    np.savez_compressed(
        os.path.join(ROOT, "tests", "golden", "rot3d_golden.npz"),
        vol24=vols[24], vol25=vols[25], vol48=vols[48], vol2030=vols[2030],
        case_vol=np.asarray(rec_vol, np.int32), case_euler=np.asarray(rec_euler, np.float64),
        case_neuler=np.asarray(rec_neuler, np.int32), case_center=np.asarray(rec_center, np.float64),
        case_sha256=np.asarray(rec_hash), full_idx=np.asarray(full_idx, np.int32),
        full_out=np.stack(full_out).astype(np.float32),
    )
    print(f"{len(rec_hash)} cases, {len(full_idx)} full outputs")


if __name__ == "__main__":
    sys.exit(main())
