"""ORACLE — CPU restatement of the TOM template-matching path (test infrastructure).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import this module, and only as the checker: the product package
``tom_amd`` never imports it.

Everything here is FP64 numpy except the rotation, which calls the float-exact C
restatement in ``oracle/tomoracle.c`` (``liboracle.so``). Volumes use MATLAB
indexing: a numpy array ``a`` of shape ``(nx, ny, nz)`` with ``a[x, y, z]`` equal
to MATLAB's ``a(x+1, y+1, z+1)``; the C side sees them Fortran-ordered
(x fastest).

Function -> reference it restates -> how it is pinned (tests/test_oracle.py):
  angle_list            Oscar3/matlabPicker/search/tom_os3_findTemplate.m:255 loop
                        order over tools/tom_os3_angleList.m:62-75 (no dedup)
                        -> explicit enumeration count/order
  rotate                Sptrans/tom_rotate.m:89-128 + Sptrans/tom_rotatec.c:146-401
                        -> bit-exact vs the reference's rot3dlin (oracle/_ref) and
                           golden vectors; axis invariants tom_rotate.m:59-61
  rotate2d              Sptrans/tom_rotatec.c:49-142 (rot2d; not buildable here,
                        its header includes mex.h) -> identity, exact rotation of
                        a linear ramp (sense and pivot), zero outside, the last
                        row (no y tap, vy2 = 0) by hand (test_rotate2d_closed_forms)
  sphere_mask_os3       search/tom_os3_sphereMask.m:4-61 -> closed form
  sphere_mask_gauss     src/tom/core/volume_fcn.cpp:932-953, 1013-1030 -> closed form
  wedge                 Filtrans/tom_wedge.m:62-77 -> grid predicate, kz=0 kept
  bandpass_weight       Filtrans/tom_bandpass.m:43-52 -> grid predicate
  apply_weight          Filtrans/tom_apply_weight_function.m:56-62 -> W=1 identity,
                        Parseval, linearity
  tom_norm_mean0_1std   Sptrans/tom_norm.m:83-119 + Analysis/tom_dev.m:55-60
                        -> closed form (T - mean)/std_pop; constant -> 0
                        (test_tom_norm_mean0_1std_closed_form)
  norm_under_mask       statistics/tom_os3_normUnderMask.m:1-8 -> direct sums
  local_stats           statistics/tom_os3_mean.m:44-63, tom_os3_std.m:43-68
                        -> brute-force direct sums (orc_flcf_direct)
  flcf                  search/tom_os3_corr.m:77-192 (FLCF branch)
                        -> orc_flcf_direct; planted copy = 1; affine invariance
  find_template         search/tom_os3_findTemplate.m:157-307 (ccc-only) +
                        search/tom_os3_bestPeak.m:13-47,94-98
                        -> brute force over the angle list
  oscar                 Filtrans/tom_oscar.m:55-80 -> brute force
  peak / intpeak        Analysis/tom_peak.m:80-108, 437-525 -> the 5x5 KAT of
                        tom_peak.m:27-50 (value and spline position)
  find_max              Analysis/tom_peakinc.c:23-52 -> numpy argmax
  fourier               Filtrans/tom_fourier.m:42-50 -> the 4x4 KAT tom_fourier.m:15-22
  norm_mask             src/tom/core/volume_fcn.cpp:168-247 -> closed forms:
                        ramp under a binary mask, (-1)^x pattern under a soft
                        mask, constants (tests/test_subtomo.py)
  simple_wedge          src/tom/core/wedge.cpp:149-243 (cutoff branch, optional
                        rotation by the template angle, transform.c:99-157)
                        -> hand-checked frequencies, k -> -k symmetry, float64
                        definition; rotated: W(R^T k) in float64 away from
                        boundaries, 45-degree sense check
  subtomo_score         src/tom/corr/corr.cpp:969-1400, correlation_handler.cpp:96-120
                        -> planted copy scores 1 (no wedge; particle wedge with
                        the particle built from numpy FFTs); brute-force
                        circular correlation; template wedge: swap symmetry with
                        the particle-wedge branch and covariance under a
                        90-degree turn, each with a planted-mistake control
  mask_cc_sphere        CorrelationHandlerPeaks mask (init_spheremask + inverse
                        fftshift) -> lattice-point counts of the ball
  psr_ring / psr        src/tom/os3/os3_correlation.cpp:183-262 -> direct sums
                        under the ring (test_psr_direct_sums)
  erase_borders, peakc, return_picklist
                        Oscar3/matlabPicker/results/tom_os3_returnPicklist.m:40-95,
                        Analysis/tom_peakc.m:39-86 -> separated planted peaks come
                        back in descending order with their erase spheres
                        (test_picklist_planted_peaks)
  flcf_box              search/tom_os3_corr.m:77-192 at the template's size
                        -> brute-force direct sums (orc_flcf_direct); the
                        autocorrelation scores 1 at zero shift
  bestpeak              search/tom_os3_bestPeak.m:13-47, os3_results.cpp:127-150
                        -> the ccc rule reproduces find_template's maps; the
                        three-score rule: recorded maps are their angle's maps
                        and only strict improvements of every score update
  os3_scores            c++Picker/src/tom/os3/os3_picker.cpp:55-95 (composition of
                        the pinned rotate, flcf, psr and flcf_box)
  tom_bin               Sptrans/tom_bin.m:38-51 -> constant and linear-ramp
                        volumes (closed forms), odd-edge cropping, one 2x2x2
                        block summed by hand (test_tom_bin_closed_forms)

Deliberate deviations shared with the product (documented, not bugs):
  * Local variance is formed from the globally mean-shifted volume and shifted
    back with the exact correction terms (same mathematics, better FP64
    conditioning), and a local variance at or below 1e-12 x the global
    variance counts as zero (score 0), so constant regions give 0 instead of
    the reference's division by rounding noise.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

VAR_FLOOR_REL = 1e-12

# FFT backend: numpy's pocketfft; set_fft_workers(k) switches to scipy.fft with
# k threads (same algorithm family, used to time the CPU baseline).
_FFT_WORKERS = 1


def set_fft_workers(k: int) -> None:
    global _FFT_WORKERS
    _FFT_WORKERS = max(1, int(k))


def _fftn(a):
    if _FFT_WORKERS > 1:
        import scipy.fft
        return scipy.fft.fftn(a, workers=_FFT_WORKERS)
    return np.fft.fftn(a)


def _ifftn(a):
    if _FFT_WORKERS > 1:
        import scipy.fft
        return scipy.fft.ifftn(a, workers=_FFT_WORKERS)
    return np.fft.ifftn(a)


def lib():
    """Load (building if needed) the C part of the oracle."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        fp = ctypes.POINTER(ctypes.c_float)
        dp = ctypes.POINTER(ctypes.c_double)
        lp = ctypes.POINTER(ctypes.c_long)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_rotmat.argtypes = [fp, ctypes.c_int, fp]
        L.orc_rot3d.argtypes = [fp, fp, ctypes.c_long, ctypes.c_long, ctypes.c_long, fp,
                                ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float]
        L.orc_find_max.argtypes = [fp, ip, fp, ip]
        L.orc_rot2d.argtypes = [fp, fp, ctypes.c_long, ctypes.c_long, fp, ctypes.c_int, ctypes.c_float,
                                ctypes.c_float]
        L.orc_flcf_direct.argtypes = [fp, lp, dp, dp, lp, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, dp, dp, dp]
        _LIB = L
    return _LIB


def build() -> None:
    import subprocess
    src = os.path.join(_HERE, "tomoracle.c")
    out = os.path.join(_HERE, "liboracle.so")
    subprocess.check_call(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off",
                           "-o", out, src, "-lm"])


def _fptr(a, t=ctypes.c_float):
    return a.ctypes.data_as(ctypes.POINTER(t))


# ----------------------------------------------------------------------------
# angles
# ----------------------------------------------------------------------------
def angle_list(phi: Sequence[float], psi: Sequence[float], theta: Sequence[float]) -> np.ndarray:
    """[phi psi theta] rows, phi outermost, theta innermost (raw product, no dedup)."""
    out = [(a, b, c) for a in phi for b in psi for c in theta]
    return np.asarray(out, dtype=np.float64).reshape(-1, 3)


def angle_grid(step_phi: float, step_psi: float, step_theta: float) -> np.ndarray:
    """phi, psi in [0, 360) and theta in [0, 180] inclusive at the given steps."""
    phi = np.arange(0, 360, step_phi, dtype=np.float64)
    psi = np.arange(0, 360, step_psi, dtype=np.float64)
    the = np.arange(0, 180 + 1e-9, step_theta, dtype=np.float64)
    return angle_list(phi, psi, the)


# ----------------------------------------------------------------------------
# rotation (float-exact, C)
# ----------------------------------------------------------------------------
def rotmat(euler) -> np.ndarray:
    e = np.ascontiguousarray(np.asarray(euler, dtype=np.float32).ravel())
    rm = np.zeros(9, np.float32)
    lib().orc_rotmat(_fptr(e), int(e.size), _fptr(rm))
    return rm


def rotate(vol: np.ndarray, euler, center=None) -> np.ndarray:
    """tom_rotate(vol, euler, 'linear', center) with the default centre
    ceil(size/2) used as a 0-based coordinate (Sptrans/tom_rotate.m:105-108)."""
    v = np.asfortranarray(vol, dtype=np.float32)
    sx, sy, sz = v.shape
    if center is None:
        center = [int(np.ceil(s / 2)) for s in v.shape]
    e = np.ascontiguousarray(np.asarray(euler, dtype=np.float32).ravel())
    out = np.zeros_like(v, order="F")
    lib().orc_rot3d(_fptr(v), _fptr(out), sx, sy, sz, _fptr(e), int(e.size),
                    float(center[0]), float(center[1]), float(center[2]))
    return out


def rotate2d(img: np.ndarray, phi, center=None) -> np.ndarray:
    """tom_rotate(img, phi, 'linear', center) for a 2D image: rot2d of
    Sptrans/tom_rotatec.c:49-142 (phi degrees, or a 2x2 matrix given as
    rm00 rm10 rm01 rm11); default pivot ceil(size/2), 0-based."""
    v = np.asfortranarray(img, dtype=np.float32)
    sx, sy = v.shape
    if center is None:
        center = [int(np.ceil(s / 2)) for s in v.shape]
    e = np.ascontiguousarray(np.asarray(phi, dtype=np.float32).ravel())
    out = np.zeros_like(v, order="F")
    lib().orc_rot2d(_fptr(v), _fptr(out), sx, sy, _fptr(e), int(e.size), float(center[0]), float(center[1]))
    return out


# ----------------------------------------------------------------------------
# masks and weights
# ----------------------------------------------------------------------------
def _centred_grid(shape):
    ax = [np.arange(-(s // 2), -(s // 2) + s, dtype=np.float64) for s in shape]
    return np.meshgrid(*ax, indexing="ij")


def sphere_mask_os3(n: int) -> Tuple[np.ndarray, int]:
    """tom_os3_sphereMask for a cubic template: binary r < floor(n/2)-1 around
    floor(n/2) (blurring forced to 0, tom_os3_sphereMask.m:22)."""
    x, y, z = _centred_grid((n, n, n))
    r = np.sqrt(x * x + y * y + z * z)
    radius = n // 2 - 1
    m = (r < radius).astype(np.float32)
    return m, int((m > 0).sum())


def sphere_mask_binary(n: int, radius: float) -> np.ndarray:
    x, y, z = _centred_grid((n, n, n))
    r = np.sqrt(x * x + y * y + z * z)
    return (r < radius).astype(np.float32)


def sphere_mask_gauss(n: int, radius: float, sigma: float, max_radius: float = -1.0) -> np.ndarray:
    """tom::init_spheremask with sigma > 0 (float arithmetic as in the C++):
    1 inside radius, exp(-((r-radius)/sigma)^2) up to max_radius
    (default radius + sqrt(2 sigma)), 0 beyond; centre floor(n/2)."""
    f = np.float32
    if max_radius < radius:
        max_radius = radius + np.sqrt(2.0 * sigma)
    c = f(n // 2)
    ax = np.arange(n, dtype=np.float32) - c
    x, y, z = np.meshgrid(ax, ax, ax, indexing="ij")
    r = np.sqrt((x * x + (y * y + z * z)).astype(np.float32)).astype(np.float32)
    t = ((r - f(radius)) / f(sigma)).astype(np.float32)
    g = np.exp(-(t * t)).astype(np.float32)
    m = np.where(r <= f(radius), f(1), np.where(r > f(max_radius), f(0), g))
    return m.astype(np.float32)


def wedge(shape, semi_angle_deg: float) -> np.ndarray:
    """tom_wedge(image, angle): 0 where tan(angle) > |x|/|z| on the centred grid;
    the kz = 0 plane is always kept (|x|/0 is inf or NaN)."""
    x, _, z = _centred_grid(shape)
    a = semi_angle_deg * np.pi / 180.0
    with np.errstate(divide="ignore", invalid="ignore"):
        q = np.abs(x) / np.abs(z)
        miss = np.tan(a) > q
    w = np.ones(shape, np.float64)
    w[miss] = 0
    return w


def bandpass_weight(shape, low: float, hi: float) -> np.ndarray:
    """tom_bandpass fast version: keep low <= |k| <= hi on the centred grid."""
    x, y, z = _centred_grid(shape)
    ln = np.sqrt(x * x + y * y + z * z)
    return ((ln <= hi) & (ln >= low)).astype(np.float64)


def apply_weight(t: np.ndarray, W: np.ndarray) -> np.ndarray:
    """tom_apply_weight_function(in, W): real(ifftn(ifftshift(fftshift(fftn(in)).*W)))."""
    F = np.fft.fftshift(np.fft.fftn(np.asarray(t, np.float64)))
    return np.real(np.fft.ifftn(np.fft.ifftshift(F * W)))


def fourier(a: np.ndarray) -> np.ndarray:
    """tom_fourier: unscaled forward DFT with MATLAB's sign."""
    return np.fft.fftn(a)


# ----------------------------------------------------------------------------
# normalisation and statistics
# ----------------------------------------------------------------------------
def tom_norm_mean0_1std(T: np.ndarray) -> np.ndarray:
    """Template preparation of tom_os3_findTemplate.m:157:
    tom_norm(1e5*(T+1), 'mean0+1std') with tom_dev's population std."""
    a = 1e5 * (np.asarray(T, np.float64) + 1.0)
    a = a * 2 + 100
    mea = a.mean()
    st = a.std()
    return (a - mea) / st if st != 0 else (a - mea)


def norm_under_mask(vol: np.ndarray, mask: np.ndarray) -> Tuple[float, float]:
    """tom_os3_normUnderMask: mean over mask~=0 (and vol~=0), population std."""
    mk = mask != 0
    n = mk.sum()
    v = np.asarray(vol, np.float64)
    mea = (v * mk * (v != 0)).sum() / n
    st = np.sqrt((((~mk) * mea + v * mk - mea) ** 2).sum() / n)
    return float(mea), float(st)


def paste_center(shape, small: np.ndarray) -> np.ndarray:
    """tom_os3_pasteCenter into zeros(shape): start floor(N/2)-floor(n/2)."""
    out = np.zeros(shape, np.float64)
    st = [s // 2 - t // 2 for s, t in zip(shape, small.shape)]
    out[st[0]:st[0] + small.shape[0], st[1]:st[1] + small.shape[1],
        st[2]:st[2] + small.shape[2]] = small
    return out


def local_stats(V: np.ndarray, M: np.ndarray, m: float) -> Tuple[np.ndarray, np.ndarray]:
    """mu_I, sigma_I of tom_os3_mean / tom_os3_std (FFT convolution with the
    pasted soft mask, ifftshift, divide by the count m), FP64.

    Evaluated on the globally mean-shifted volume V - g for conditioning and
    shifted back exactly: with s = sum(M)/m, mu = mu_s + g s and
    E2 - mu^2 = (e2_s - mu_s^2) + (1 - s)(2 g mu_s + g^2 s), which is the
    reference's E[x^2] - E[x]^2 (the (1 - s) term vanishes for binary masks).
    A variance at or below VAR_FLOOR_REL x the global variance counts as 0."""
    V = np.asarray(V, np.float64)
    M = np.asarray(M, np.float64)
    g = V.mean()
    Vs = V - g
    s = float(M.sum()) / m
    FM = _fftn(paste_center(V.shape, M))
    mu_s = np.real(np.fft.ifftshift(_ifftn(_fftn(Vs) * FM))) / m
    e2 = np.real(np.fft.ifftshift(_ifftn(_fftn(Vs * Vs) * FM))) / m
    var = (e2 - mu_s * mu_s) + (1.0 - s) * (2.0 * g * mu_s + g * g * s)
    floor = VAR_FLOOR_REL * float((Vs * Vs).mean())
    var = np.where(var > floor, var, 0.0)
    return mu_s + g * s, np.sqrt(var)


def correlation_raw(V: np.ndarray, tm: np.ndarray, FV: Optional[np.ndarray] = None) -> np.ndarray:
    """real(ifftshift(ifftn(fftn(V) .* conj(fftn(pasteCenter(zeros, tm))))))."""
    if FV is None:
        FV = _fftn(np.asarray(V, np.float64))
    FT = _fftn(paste_center(V.shape, np.asarray(tm, np.float64)))
    return np.real(np.fft.ifftshift(_ifftn(FV * np.conj(FT))))


def flcf(V, tm, m, mu_T, sig_T, mu_I, sig_I, FV=None) -> np.ndarray:
    """Score map of the FLCF branch of tom_os3_corr for a masked template tm."""
    c = correlation_raw(V, tm, FV) / m
    den = sig_T * sig_I
    out = np.zeros(V.shape, np.float64)
    ok = den > 0
    out[ok] = (c[ok] - mu_I[ok] * mu_T) / den[ok]
    return out


def prepare_template(T, M, W=None):
    """Template normalisation and the fixed template statistics of
    tom_os3_findTemplate.m:157, 233-234."""
    Tn = tom_norm_mean0_1std(T)
    tmp = apply_weight(Tn, W) if W is not None else Tn
    mu_T, sig_T = norm_under_mask(tmp, M)
    return Tn, mu_T, sig_T


def flcf_angle(V, Tn, M, m, mu_T, sig_T, mu_I, sig_I, FV, angle, W=None):
    """One orientation of the findTemplate loop (tom_os3_findTemplate.m:264-270):
    rotate, mask, weight, re-mask, FLCF score map."""
    t = rotate(np.asarray(Tn, np.float32), angle).astype(np.float64) * M
    if W is not None:
        t = apply_weight(t, W)
    return flcf(V, t * M, m, mu_T, sig_T, mu_I, sig_I, FV)


def find_template(V, T, M, angles, W=None, return_all=False):
    """ccc-only tom_os3_findTemplate: returns (best score, 1-based best angle).

    Per angle: t = rotate(Tn, a) .* M; t = apply_weight(t, W); score =
    FLCF(V, t .* M); update where old < new (strict, first angle wins);
    maps start at -1000."""
    V = np.asarray(V, np.float64)
    M = np.asarray(M, np.float64)
    m = float((M > 0).sum())
    Tn, mu_T, sig_T = prepare_template(T, M, W)
    mu_I, sig_I = local_stats(V, M, m)
    FV = _fftn(V)
    ccc = np.full(V.shape, -1000.0)
    ang = np.full(V.shape, -1000.0)
    allmaps = []
    for i, a in enumerate(np.asarray(angles).reshape(-1, 3)):
        s = flcf_angle(V, Tn, M, m, mu_T, sig_T, mu_I, sig_I, FV, a, W)
        if return_all:
            allmaps.append(s)
        upd = ccc < s
        ccc[upd] = s[upd]
        ang[upd] = i + 1
    if return_all:
        return ccc, ang, np.stack(allmaps)
    return ccc, ang


def flcf_at_voxels(V, Tn, M, mu_T, sig_T, angles, voxels, W=None, chunk=256, workers=1):
    """FLCF scores of the findTemplate loop at chosen voxels only, FP64 direct
    sums (for volumes too large for full FP64 maps): for voxel y and
    orientation a, with t_a = remask(apply_weight(rotate(Tn, a) . M, W)) and
    the template centre floor(n/2) at y (pasteCenter + ifftshift, circular),
      c  = sum_u t_a(u) V(y + u - floor(n/2)) / m
      mu = sum_u M(u) V(...) / m,  e2 = sum_u M(u) V(...)^2 / m
      score = (c - mu mu_T) / (sig_T sqrt(e2 - mu^2))
    The same quantities as local_stats + flcf (tom_os3_corr.m:77-192,
    tom_os3_mean.m, tom_os3_std.m), without the global mean shift or the
    variance floor. Returns (n_voxels, n_angles) float64."""
    M = np.asarray(M, np.float64)
    n = M.shape
    N = V.shape
    m = float((M > 0).sum())
    sel = np.nonzero(M > 0)
    off = [sel[d] - n[d] // 2 for d in range(3)]
    vox = np.asarray(voxels, np.int64).reshape(-1, 3)
    idx = [(vox[:, d][:, None] + off[d][None, :]) % N[d] for d in range(3)]
    patch = np.asarray(V[idx[0], idx[1], idx[2]], np.float64)  # (nv, k)
    w = M[sel]
    mu = patch @ w / m
    e2 = (patch * patch) @ w / m
    sig = np.sqrt(np.maximum(e2 - mu * mu, 0.0))
    A = np.asarray(angles, np.float64).reshape(-1, 3)
    out = np.empty((vox.shape[0], A.shape[0]))
    Tn32 = np.asarray(Tn, np.float32)
    for a0 in range(0, A.shape[0], chunk):
        ts = []
        for a in A[a0:a0 + chunk]:
            ts.append(rotate(Tn32, a).astype(np.float64) * M)
        ts = np.stack(ts)
        if W is not None:
            if workers > 1:
                # ifftshift(fftshift(F) . W) = F . ifftshift(W)
                import scipy.fft
                F = scipy.fft.fftn(ts, axes=(1, 2, 3), workers=workers)
                F *= np.fft.ifftshift(W)[None]
                ts = np.real(scipy.fft.ifftn(F, axes=(1, 2, 3), workers=workers))
            else:
                ts = np.stack([apply_weight(t, W) for t in ts])
        tk = ts[:, sel[0], sel[1], sel[2]] * w[None]  # re-masked template at the mask's voxels
        c = tk @ patch.T / m  # (chunk, nv)
        out[:, a0:a0 + chunk] = ((c - mu[None] * mu_T) / (sig_T * sig[None])).T
    return out


def oscar(V, T, angles):
    """tom_oscar: unnormalised XCF, template pasted at N/2+1-n/2 (1-based),
    real(ifftshift(ifftn(fftn(temp).*conj(fftn(vol))))), strict '>' update.
    Returns (ccfmax, 0-based angle index map, -1 where never updated)."""
    V = np.asarray(V, np.float64)
    FVc = np.conj(np.fft.fftn(V))
    ccf = np.full(V.shape, -1000.0)
    idx = np.full(V.shape, -1, np.int64)
    for i, a in enumerate(np.asarray(angles).reshape(-1, 3)):
        t = rotate(np.asarray(T, np.float32), a).astype(np.float64)
        temp = paste_center(V.shape, t)
        c = np.real(np.fft.ifftshift(np.fft.ifftn(np.fft.fftn(temp) * FVc)))
        upd = c > ccf
        ccf[upd] = c[upd]
        idx[upd] = i
    return ccf, idx


# ----------------------------------------------------------------------------
# peaks
# ----------------------------------------------------------------------------
def find_max(a: np.ndarray):
    """tom_peakinc find_max: (value, 0-based [x y z]) of the first maximum in
    x-fastest order, start value -99999, strict '>'."""
    v = np.asfortranarray(a, dtype=np.float32)
    if v.ndim < 3:
        v = v.reshape(v.shape + (1,) * (3 - v.ndim), order="F")
    dims = np.asarray(v.shape, np.int32)
    mv = np.zeros(1, np.float32)
    pos = np.zeros(3, np.int32)
    lib().orc_find_max(_fptr(v), dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                       _fptr(mv), pos.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return float(mv[0]), pos.copy()


def _parabola_peak(p3):
    """interp1(1:3, p3, 1:0.001:3, 'spline') with 3 points is the parabola
    through them; max over the 0.001 grid (first on ties)."""
    xs = np.round(np.arange(1.0, 3.0 + 1e-9, 0.001), 3)
    a = (p3[0] - 2 * p3[1] + p3[2]) / 2.0
    b = (p3[2] - p3[0]) / 2.0
    y = p3[1] + b * (xs - 2) + a * (xs - 2) ** 2
    i = int(np.argmax(y))
    return float(y[i]), i


def peak(a: np.ndarray, spline: bool = False):
    """tom_peak(A) / tom_peak(A,'spline') for 2D/3D arrays (1-based output)."""
    a = np.asarray(a, np.float64)
    v = a.max()
    first = np.argwhere(np.asfortranarray(a) == v)
    # MATLAB find returns column-major order: first hit in x-fastest order
    order = np.lexsort(tuple(first[:, d] for d in range(first.shape[1])))
    s = first[order[0]] + 1
    if not spline:
        return s.astype(np.float64), float(v)
    c = s.astype(np.float64)
    vals = []
    for d in range(a.ndim):
        if s[d] == 1 or s[d] == a.shape[d]:
            c[d] = s[d]
            vals.append(v)
            continue
        idx = [int(x) - 1 for x in s]
        p3 = []
        for off in (-1, 0, 1):
            j = list(idx)
            j[d] += off
            p3.append(a[tuple(j)])
        val, pos = _parabola_peak(p3)
        c[d] = s[d] - 1 + pos * 0.001
        vals.append(val)
    return c, float(max(vals))


# ----------------------------------------------------------------------------
# subtomogram scoring (corr3d_batch / corr3d_single, 'peak' handler)
# ----------------------------------------------------------------------------
def norm_mask(v: np.ndarray, M: np.ndarray, boolean: bool):
    """tom::norm_mask(v, mask, NORM_NO_NORM, &variance, is_boolean_mask)
    (Oscar3/c++Picker/src/tom/core/volume_fcn.cpp:168-247) in FP64:
    mean over mask~=0; v2 = (v-mean)*mask (boolean: v-mean) inside, 0 outside;
    variance = (sum v2^2 - n mean2^2) / numel; v2 -= mean2 inside."""
    v = np.asarray(v, np.float64)
    B = M != 0
    n = int(B.sum())
    mean = v[B].sum() / n
    out = np.zeros_like(v)
    out[B] = (v[B] - mean) if boolean else (v[B] - mean) * M[B]
    s, s2 = out[B].sum(), (out[B] ** 2).sum()
    mean2 = s / n
    var = (s2 - n * mean2 * mean2) / v.size
    if mean2 != 0:
        out[B] -= mean2
    return out, var


# mex angle conversion of tom_mex_corr3d.cpp:247-249 (degrees times PI/180.)
_MEX_PI = 3.141592653589793238512808959406186204433


def _init_rotmat_zxz(phi, theta, psi):
    """tom_transf_init_rotmat_3d(P, 3, 3, {2, 0, 2}, {phi, theta, psi})
    (Oscar3/c++Picker/src/tom/core/transform.c:99-157) in double: starting
    from the identity, each rotation updates the columns of every row."""
    P = [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]
    import math
    for axis, ang in ((2, phi), (0, theta), (2, psi)):
        if ang - math.floor(ang / (_MEX_PI / 2.0)) * (2.0 * _MEX_PI) == 0.0:
            continue
        sa, ca = math.sin(ang), math.cos(ang)
        i0, i1 = {0: (1, 2), 1: (0, 2), 2: (0, 1)}[axis]
        for r in P:
            p0, p1 = r[i0], r[i1]
            if axis == 1:
                r[i0] = p0 * ca + p1 * sa
                r[i1] = -p0 * sa + p1 * ca
            else:
                r[i0] = p0 * ca - p1 * sa
                r[i1] = p0 * sa + p1 * ca
    return P


def simple_wedge(shape, semi_angle_deg: float, euler_deg=None):
    """tom::SimpleWedge(angle, cutoff = (min(size)-1)/2), the wedge that
    tom_mex_corr3d builds from a scalar WEDGE_REF / WEDGE_PART
    (tom_mex_corr3d.cpp:334,410), optionally rotated by [phi psi theta]
    degrees as corr3d_single does for the template's wedge (corr.cpp:1089,
    SimpleWedge::rotate). Restates init_wedge_volume's cutoff branch
    (src/tom/core/wedge.cpp:149-224) in its float arithmetic and order:
    P = init_rotmat(z,x,z; phi,theta,psi) in double with the offset
    -P.centre, cast to float; grid index x -> xs = (x + ceil(n/2)) mod n;
    keep = |P(xs)| <= cutoff and (|z'| < 1e-4 or tan(angle) <= |x'|/|z'|).
    The reference evaluates it on the r2c half (z index <= n/2); the full
    unshifted grid returned here takes the other half from the Hermitian
    partner, which is what the half-spectrum product means."""
    f = np.float32
    sx, sy, sz = shape
    if euler_deg is None:
        phi = psi = theta = 0.0
    else:
        phi, psi, theta = (float(a) * (_MEX_PI / 180.0) for a in euler_deg)
    cx, cy, cz = (float(np.ceil((s - 1) / 2.0)) for s in shape)
    Pl = _init_rotmat_zxz(phi, theta, psi)
    off = [-(r[0] * cx + r[1] * cy + r[2] * cz) for r in Pl]
    P = [[f(r[0]), f(r[1]), f(r[2]), f(o)] for r, o in zip(Pl, off)]
    maxdist = np.sqrt(f(cx * cx + cy * cy + cz * cz))
    tan_a = f(np.tan(semi_angle_deg * (_MEX_PI / 180.0)))
    c0 = float((min(shape) - 1) // 2)
    cutoff = f(c0) if (c0 > 0 and c0 < maxdist) else f(0)
    ox, oy, oz = (s // 2 + s % 2 for s in shape)
    xs = ((np.arange(sx) + ox) % sx).astype(np.float32)[:, None, None]
    ys = ((np.arange(sy) + oy) % sy).astype(np.float32)[None, :, None]
    zs = ((np.arange(sz // 2 + 1) + oz) % sz).astype(np.float32)[None, None, :]
    xt = P[0][2] * zs + P[0][3]
    yt = P[1][2] * zs + P[1][3]
    zt = P[2][2] * zs + P[2][3]
    xt = P[0][1] * ys + xt
    yt = P[1][1] * ys + yt
    zt = P[2][1] * ys + zt
    x_ = P[0][0] * xs + xt
    y_ = P[1][0] * xs + yt
    z_ = np.abs(P[2][0] * xs + zt)
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = (np.abs(x_) / z_).astype(np.float32)
        r = np.sqrt((x_ * x_ + y_ * y_) + z_ * z_).astype(np.float32)
        if cutoff > 0:
            half = (r <= cutoff) & ((z_ < f(1e-4)) | (tan_a <= ratio))
        else:
            half = (z_ < f(1e-4)) | (tan_a <= ratio)
    full = np.zeros(shape, np.float64)
    full[:, :, :sz // 2 + 1] = half
    for z in range(sz // 2 + 1, sz):
        full[:, :, z] = np.roll(half[::-1, ::-1, sz - z], 1, axis=(0, 1))
    return full


def simple_wedge_half(shape, semi_angle_deg: float):
    """The unrotated SimpleWedge weight on the full unshifted grid."""
    return simple_wedge(shape, semi_angle_deg)


def subtomo_score(particle, template, M, mask_cc, semi_angle_deg, angles, wedge_ref_deg=None):
    """One particle against one template over an angle list, the way
    corr3d_batch -> corr3d_single (reapply_mask_after_wedge = true) ->
    CorrelationHandlerPeaks do it (src/tom/corr/corr.cpp:969-1210, 1217-1400,
    src/tom/corr/correlation_handler.cpp:96-120), FP64:
      template -= mean(template); particle: norm_mask (soft), FFT, DC = 0
      per angle: t = rotate(T, a); norm_mask(t); Tf = FFT(t), DC = 0
                 template wedge (wedge_ref_deg not None, corr.cpp:1086-1111):
                   Pw = P . W_ref rotated by a, DC = 0, pw = c2r(Pw)
                   (unnormalised), norm_mask_bool(pw) -> var_p, Pw = FFT(pw)
                 particle wedge (semi_angle_deg not None, corr.cpp:1146-1171):
                   Tw = Tf . W_p, DC = 0, tw = c2r(Tw) (unnormalised),
                   norm_mask_bool(tw) -> var_t, Tw = FFT(tw)
                 C = conj(Tw).Pw, DC = 0; c = c2r(C)/(sqrt(var_p var_t) numel^2)
                 peak of c over mask_cc != 0 (first in x-fastest order)
      keep the strictly better value (first angle wins).
    A wedge semi-angle of None (or negative / NaN) is NoWedge ([] in the
    gateway); 0 is SimpleWedge(0), i.e. only the cutoff sphere.
    Returns (angle index or -1, peak [x y z] unshifted 0-based, value, maps)."""
    n = particle.size
    T = np.asarray(template, np.float64)
    T = T - T.mean()
    p, var_p0 = norm_mask(particle, M, False)
    P = np.fft.fftn(p)
    P.flat[0] = 0

    def active(a):
        return a is not None and np.isfinite(a) and a >= 0

    W = simple_wedge(particle.shape, semi_angle_deg) if active(semi_angle_deg) else None
    if mask_cc is None:
        mask_cc = np.ones(particle.shape)
    best = (-1, np.zeros(3, int), 0.0)
    vals = []
    for ia, a in enumerate(np.asarray(angles).reshape(-1, 3)):
        t = rotate(T.astype(np.float32), a).astype(np.float64)
        t, var_t = norm_mask(t, M, False)
        Tf = np.fft.fftn(t)
        Tf.flat[0] = 0
        var_p = var_p0
        if active(wedge_ref_deg):
            Pw = P * simple_wedge(particle.shape, wedge_ref_deg, a)
            Pw.flat[0] = 0
            pw = np.real(np.fft.ifftn(Pw)) * n
            pw, var_p = norm_mask(pw, M, True)
            Pw = np.fft.fftn(pw)
        else:
            Pw = P
        if W is not None:
            Tw = Tf * W
            Tw.flat[0] = 0
            tw = np.real(np.fft.ifftn(Tw)) * n  # FFTW c2r is unnormalised
            tw, var_t = norm_mask(tw, M, True)
            Tw = np.fft.fftn(tw)
        else:
            Tw = Tf
        C = np.conj(Tw) * Pw
        C.flat[0] = 0
        c = np.real(np.fft.ifftn(C)) * n / (np.sqrt(var_p * var_t) * n * n)
        sel = np.asfortranarray(mask_cc != 0)
        cf = np.asfortranarray(c)
        flat_c = cf.ravel(order="F")
        flat_m = sel.ravel(order="F")
        idxs = np.nonzero(flat_m)[0]
        j = idxs[np.argmax(flat_c[idxs])]
        val = flat_c[j]
        pos = np.array([j % c.shape[0], (j // c.shape[0]) % c.shape[1], j // (c.shape[0] * c.shape[1])])
        vals.append(val)
        if ia == 0 or val > best[2]:
            best = (ia, pos, val)
    return best[0], best[1], best[2], np.asarray(vals)


def mask_cc_sphere(shape, radius: float) -> np.ndarray:
    """Peak-search mask in unshifted correlation coordinates: shifts s with
    |s| <= radius (wrap-around), as a mask_cc for the 'peak' handler."""
    ax = [np.fft.fftfreq(s, 1.0 / s) for s in shape]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    return (np.sqrt(X * X + Y * Y + Z * Z) <= radius).astype(np.float32)


# ----------------------------------------------------------------------------
# secondary scores
# ----------------------------------------------------------------------------
def psr_ring(n: int) -> np.ndarray:
    """Ring mask of os3_psr (Oscar3/c++Picker/src/tom/os3/os3_correlation.cpp:
    203-243): binary sphere of radius n/2 minus the binary sphere of radius
    max(n/4, 5), both about floor(n/2)."""
    return sphere_mask_binary(n, n // 2) - sphere_mask_binary(n, max(n // 4, 5))


def psr(xcf: np.ndarray, n: int) -> np.ndarray:
    """Peak-to-sidelobe ratio map (os3_correlation.cpp:245-259): (xcf - mean) /
    std under the ring, FP64 local statistics, 0 where the std is 0."""
    M = psr_ring(n).astype(np.float64)
    m = float((M > 0).sum())
    x = np.asarray(xcf, np.float64)
    mu, sig = local_stats(x, M, m)
    out = np.zeros_like(x)
    nz = sig > 0
    out[nz] = (x[nz] - mu[nz]) / sig[nz]
    return out


def flcf_box(search, tmpl, M=None):
    """tom_os3_corr(search, tmpl) for a search volume of the template's size
    (Oscar3/matlabPicker/search/tom_os3_corr.m:77-192: image statistics under
    the mask, template statistics tom_os3_normUnderMask, every voxel divided,
    circular). M None: tom_os3_sphereMask(template)."""
    S = np.asarray(search, np.float64)
    T = np.asarray(tmpl, np.float64)
    if M is None:
        M, _ = sphere_mask_os3(T.shape[0])
    M = np.asarray(M, np.float64)
    m = float((M > 0).sum())
    mu_T, sig_T = norm_under_mask(T, M)
    mu_I, sig_I = local_stats(S, M, m)
    c = correlation_raw(S, T * M) / m
    with np.errstate(divide="ignore", invalid="ignore"):
        return (c - mu_I * mu_T) / (sig_T * sig_I)


def bestpeak(res, ccc, psr, soc, angle, rule):
    """tom_os3_bestPeak (search/tom_os3_bestPeak.m:13-47; the three-score rule
    is os3_results::update, Oscar3/c++Picker/src/tom/os3/os3_results.cpp:
    127-150): where every selected score strictly improves, copy the maps and
    the angle. res: dict of float64 maps (start -1000) and 'angle'."""
    up = np.ones(np.shape(ccc), bool)
    if "ccc" in rule:
        up &= res["ccc"] < ccc
    if "psr" in rule:
        up &= res["psr"] < psr
    if "soc" in rule:
        up &= res["soc"] < soc
    res["ccc"][up] = np.asarray(ccc)[up]
    if psr is not None:
        res["psr"][up] = np.asarray(psr)[up]
    if soc is not None:
        res["soc"][up] = np.asarray(soc)[up]
    res["angle"][up] = angle
    return res


def os3_scores(V, T, M, angle):
    """One orientation of the correlation picker with secondary scores
    (Oscar3/c++Picker/src/tom/os3/os3_picker.cpp:55-95): t = rotate(T).M,
    ccc = tom_os3_corr(V, t, M) (template statistics of t), psr = the PSR map
    of ccc, soc = tom_os3_corr(ccc, flcf_box(t, t, M), M). FP64."""
    M = np.asarray(M, np.float64)
    m = float((M > 0).sum())
    t = rotate(np.asarray(T, np.float32), angle).astype(np.float64) * M
    V = np.asarray(V, np.float64)
    mu_t, sig_t = norm_under_mask(t, M)
    mu_I, sig_I = local_stats(V, M, m)
    ccc = flcf(V, t * M, m, mu_t, sig_t, mu_I, sig_I)
    pr = psr(ccc, T.shape[0])
    auto = flcf_box(t, t, M)
    mu_a, sig_a = norm_under_mask(auto, M)
    mu_c, sig_c = local_stats(ccc, M, m)
    soc = flcf(ccc, auto * M, m, mu_a, sig_a, mu_c, sig_c)
    return ccc, pr, soc


def tom_bin(im: np.ndarray, nbin: int = 1) -> np.ndarray:
    """tom_bin for a 3D volume (Sptrans/tom_bin.m:38-51): nbin rounds of the
    single-precision sum of the eight 2x2x2 samples in the order the .m file
    writes them ((1,1,1) (1,2,1) (2,1,1) (2,2,1), then the same at z + 1,
    MATLAB (x, y, z) indices), divided by 8; odd trailing samples dropped."""
    a = np.asarray(im, np.float32)
    for _ in range(int(nbin)):
        X, Y, Z = (d // 2 * 2 for d in a.shape)
        a = a[:X, :Y, :Z]
        terms = [a[0::2, 0::2, 0::2], a[0::2, 1::2, 0::2], a[1::2, 0::2, 0::2], a[1::2, 1::2, 0::2],
                 a[0::2, 0::2, 1::2], a[0::2, 1::2, 1::2], a[1::2, 0::2, 1::2], a[1::2, 1::2, 1::2]]
        acc = terms[0].copy()
        for t in terms[1:]:
            acc = (acc + t).astype(np.float32)
        a = (acc / np.float32(8)).astype(np.float32)
    return a


# ----------------------------------------------------------------------------
# picklist (tom_os3_returnPicklist -> tom_peakc)
# ----------------------------------------------------------------------------
def erase_borders(volume, tsize):
    """tom_os3_eraseBorders (tools/tom_os3_eraseBorders.m): volume .* a mask
    that keeps the centred box of size N - (n+1); a product, so NaN stays NaN
    and negative border values become -0."""
    shape = np.shape(volume)
    inner = [max(s - (t + 1), 0) for s, t in zip(shape, tsize)]
    if len(shape) == 3 and shape[2] == 1:
        inner[2] = 1  # an image (one z plane): borders in x and y only
    keep = paste_center(shape, np.ones(inner, np.float32)).astype(np.float32)
    return np.asarray(volume, np.float32) * keep


def _peakc_region(c, radius, N):
    """The cut-out and paste of tom_peakc along one axis (1-based peak c):
    tom_cut_out (Sptrans/tom_cut_out.m) clamps the start c - r to 1 and keeps
    the size 2r, clipped at the far end; tom_paste (Sptrans/tom_paste.m) puts
    the masked box back at the unclamped start c - r, clipped to the volume.
    Returns (target first, target last, source shift, box size), 1-based."""
    coord = c - radius
    pos = max(coord, 1)
    bound = pos + 2 * radius - 1 if pos + 2 * radius <= N else N
    dims = bound - pos + 1
    t0, t1 = max(coord, 1), min(coord + dims - 1, N)
    return t0, t1, pos - coord, dims, coord


def peakc(a, radius, inplace=False):
    """tom_peakc(in, radius) (Analysis/tom_peakc.m:39-86) on a 3D map (a 2D
    map is (nx, ny, 1)): the first maximum c (1-based, x-fastest, tom_peakinc),
    then in = tom_paste(in, out .* (sphere == 0), c - r) with out =
    tom_cut_out(in, c - r, [2r 2r 2r], 'nofill') and the sphere of
    tom_spheremask (Sptrans/tom_spheremask.m: 1 where the distance to the
    box centre size/2 + 1 is below r; for a 2D map the z centre is 1.5).
    Near the low border the cut-out starts at 1 while the paste starts at
    c - r, so the kept box voxels land shifted by 1 - (c - r). A map with a
    single z plane keeps z fixed (the reference's 2D branch). Returns
    (1-based position, value, map); the map is float32, Fortran-ordered."""
    a = np.asarray(a, np.float32)
    if a.ndim == 2:
        a = a[:, :, None]
    if not inplace or not a.flags.f_contiguous:
        a = np.array(a, np.float32, order="F", copy=True)
    v, p = find_max(a)
    c = [int(q) + 1 for q in p]
    N = a.shape
    reg = [_peakc_region(c[d], radius, N[d]) for d in range(3)]
    if N[2] == 1:
        reg[2] = (1, 1, 0, 1, 1)
    dims = [r[3] for r in reg]
    cen = [d / 2.0 + 1.0 for d in dims]
    I, J, K = np.meshgrid(*[np.arange(d) for d in dims], indexing="ij")
    dist = np.sqrt((I + 1 - cen[0]) ** 2 + (J + 1 - cen[1]) ** 2 + (K + 1 - cen[2]) ** 2)
    keep = (dist >= radius).astype(np.float32)
    tgt = tuple(slice(r[0] - 1, r[1]) for r in reg)
    src = tuple(slice(r[0] - 1 + r[2], r[1] + r[2]) for r in reg)
    box = tuple(slice(r[0] - r[4], r[1] - r[4] + 1) for r in reg)
    if all(r[1] >= r[0] for r in reg):
        a[tgt] = a[src] * keep[box]
    return np.asarray(c), v, a


def return_picklist(peaks, angles, tsize, n, erase=None):
    """tom_os3_returnPicklist rows [x y z value angle] (no binning). The loop
    runs while std(peaks(:)) > 0, i.e. while the map holds only finite
    values (a NaN or Inf makes the std NaN) that are not all equal. erase:
    None = the reference's rule (borders erased unless the map has the
    template's size), False = never."""
    r = int(np.ceil(tsize[0] / 2))
    pk = np.asarray(peaks, np.float32)
    if pk.ndim == 2:
        pk = pk[:, :, None]
    if erase is not False and tuple(pk.shape) != tuple(tsize):
        pk = erase_borders(pk, tsize)
    pk = np.array(pk, np.float32, order="F", copy=True)
    angs = None if angles is None else np.asarray(angles)
    if angs is not None and angs.ndim == 2:
        angs = angs[:, :, None]
    rows = []
    while len(rows) < n and np.isfinite(pk).all() and pk.max() > pk.min():
        c, v, pk = peakc(pk, r, inplace=True)
        ang = 0 if angs is None else angs[c[0] - 1, c[1] - 1, c[2] - 1]
        rows.append([c[0], c[1], c[2], v, ang])
    return np.asarray(rows, np.float64).reshape(-1, 5)
