"""MATLAB-shaped host interface over libtomhip.so.

Arrays follow MATLAB indexing: a numpy array of shape (nx, ny, nz) with
``a[x, y, z]`` == MATLAB ``a(x+1, y+1, z+1)``; it is passed to the library
x-fastest (Fortran order). A torch CUDA tensor is taken as already resident in
device memory with shape (nz, ny, nx), C-contiguous (the same x-fastest bytes).
Every compute call goes to the GPU library; there is no CPU path.

Function                 reference counterpart
tom_rotate               Sptrans/tom_rotate.m:89-128 -> tom_rotatec MEX
tom_rotate_batch         loop of tom_rotate over an angle list
tom_peakinc / tom_peak   Analysis/tom_peakinc.c:23-77, Analysis/tom_peak.m:80-108
tom_os3_sphereMask       Oscar3/matlabPicker/search/tom_os3_sphereMask.m:4-36
tom_spheremask_gauss     src/tom/core/volume_fcn.cpp:1013-1030 (init_spheremask)
tom_wedge / weight       Filtrans/tom_wedge.m:62-77, Filtrans/tom_bandpass.m:43-52
prepare_template         tom_os3_findTemplate.m:157, 233-234
tom_os3_corr             Oscar3/matlabPicker/search/tom_os3_corr.m:77-192 (FLCF)
tom_os3_findTemplate     search/tom_os3_findTemplate.m:157-307, ccc-only
tom_oscar                Filtrans/tom_oscar.m:1-82
SearchPlan               the per-job state of tom_os3_findTemplate (stats once,
                         angle batches, reduction) exposed step by step
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence, Tuple

import numpy as np

from ._lib import TomHipError, load

try:  # torch is only plumbing for device memory and streams
    import torch
except Exception:  # pragma: no cover
    torch = None


# ---------------------------------------------------------------- context
class Context:
    """One libtomhip context (device + stream)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = ctypes.c_void_p()
        rc = self.lib.tomhip_init(device, ctypes.byref(h))
        if rc != 0:
            raise TomHipError(rc, self.lib.tomhip_last_error(None).decode())
        self.h = h
        self.device = device

    def check(self, rc: int):
        if rc != 0:
            raise TomHipError(rc, self.lib.tomhip_last_error(self.h).decode())

    def use_torch_stream(self):
        if torch is not None and torch.cuda.is_available():
            s = torch.cuda.current_stream(self.device).cuda_stream
            self.check(self.lib.tomhip_set_stream(self.h, ctypes.c_void_p(s)))

    def close(self):
        if self.h:
            self.lib.tomhip_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


_default: Optional[Context] = None


def context() -> Context:
    global _default
    if _default is None:
        _default = Context(0)
    return _default


# ---------------------------------------------------------------- buffers
def _is_torch(a) -> bool:
    return torch is not None and isinstance(a, torch.Tensor)


class _Buf:
    """Pointer + MATLAB dims of an input or output array."""

    def __init__(self, a, dtype=np.float32):
        if _is_torch(a):
            tdt = {np.float32: torch.float32, np.int32: torch.int32}[dtype]
            if a.dtype != tdt or not a.is_contiguous():
                a = a.to(tdt).contiguous()
            self.obj = a
            self.ptr = ctypes.c_void_p(a.data_ptr())
            self.dims = tuple(reversed(a.shape))
            self.torch = True
        else:
            a = np.asfortranarray(a, dtype=dtype)
            self.obj = a
            self.ptr = ctypes.c_void_p(a.ctypes.data)
            self.dims = a.shape
            self.torch = False


def _dims3(dims) -> np.ndarray:
    d = list(dims) + [1] * (3 - len(dims))
    return np.asarray(d[:3], dtype=np.int64)


def _p(arr: np.ndarray):
    return ctypes.c_void_p(arr.ctypes.data)


def _empty_like(buf: _Buf, dims, dtype=np.float32):
    if buf.torch:
        tdt = {np.float32: torch.float32, np.int32: torch.int32}[dtype]
        return torch.empty(tuple(reversed(tuple(dims))), dtype=tdt, device=buf.obj.device)
    return np.zeros(tuple(dims), dtype=dtype, order="F")


def _ptr(a):
    return ctypes.c_void_p(a.data_ptr()) if _is_torch(a) else ctypes.c_void_p(a.ctypes.data)


def _prep_ctx(ctx, *bufs) -> Context:
    ctx = ctx or context()
    if any(b is not None and b.torch for b in bufs):
        ctx.use_torch_stream()
    return ctx


# ---------------------------------------------------------------- rotation
def tom_rotate(vol, euler, interp: str = "linear", center=None, ctx: Optional[Context] = None):
    """out = tom_rotate(in, [phi psi theta] | 3x3 matrix, 'linear', [center]).

    center is the 0-based pivot handed to tom_rotatec (default ceil(size/2)).
    A 2D image (numpy (nx, ny) or torch (ny, nx)) takes one angle phi in
    degrees or a 2x2 matrix and goes through rot2d (tom_rotatec.c:49-142)."""
    if interp != "linear":
        raise ValueError("only 'linear' interpolation exists (tom_rotatec.c:424-430)")
    b = _Buf(vol)
    ctx = _prep_ctx(ctx, b)
    if len(b.dims) == 2:
        e = np.asarray(euler, dtype=np.float32)
        if e.size == 4 and e.ndim == 2:
            e = np.asfortranarray(e).ravel(order="F")
        e = np.ascontiguousarray(e.ravel())
        if e.size not in (1, 4):
            raise ValueError("a 2D rotation takes one angle or a 2x2 matrix")
        c = np.asarray([np.ceil(d / 2) for d in b.dims] if center is None else np.asarray(center).ravel()[:2],
                       np.float32)
        c = np.ascontiguousarray(c)
        out = _empty_like(b, b.dims)
        ctx.check(ctx.lib.tomhip_rotate2d_f32(ctx.h, b.ptr, _ptr(out), int(b.dims[0]), int(b.dims[1]), _p(e),
                                              int(e.size), _p(c)))
        return out
    e = np.asarray(euler, dtype=np.float32)
    neuler = 9 if e.size == 9 else 3
    if e.size == 9 and e.ndim == 2:
        e = np.asfortranarray(e)
        e = e.ravel(order="F")
    e = np.ascontiguousarray(e.ravel())
    if e.size not in (3, 9):
        raise ValueError("euler must hold 3 angles or 9 matrix entries")
    d = _dims3(b.dims)
    out = _empty_like(b, b.dims)
    c = None if center is None else np.ascontiguousarray(np.asarray(center, np.float32).ravel()[:3])
    ctx.check(ctx.lib.tomhip_rotate3d_f32(ctx.h, b.ptr, _ptr(out), int(d[0]), int(d[1]), int(d[2]), _p(e), neuler,
                                          None if c is None else _p(c)))
    return out


def tom_rotate_batch(vol, angles, center=None, mask=None, ctx: Optional[Context] = None):
    """Rotations of one volume by every [phi psi theta] row of angles; the
    result has a trailing angle axis (nx, ny, nz, A) (torch: (A, nz, ny, nx))."""
    b = _Buf(vol)
    mb = _Buf(mask) if mask is not None else None
    ctx = _prep_ctx(ctx, b, mb)
    ang = np.ascontiguousarray(np.asarray(angles, np.float32).reshape(-1, 3))
    A = ang.shape[0]
    d = _dims3(b.dims)
    if b.torch:
        out = torch.empty((A, int(d[2]), int(d[1]), int(d[0])), dtype=torch.float32, device=b.obj.device)
    else:
        out = np.zeros((int(d[0]), int(d[1]), int(d[2]), A), np.float32, order="F")
    c = None if center is None else np.ascontiguousarray(np.asarray(center, np.float32).ravel()[:3])
    ctx.check(ctx.lib.tomhip_rotate3d_batch_f32(ctx.h, b.ptr, _ptr(out), int(d[0]), int(d[1]), int(d[2]), _p(ang),
                                                A, None if c is None else _p(c), None if mb is None else mb.ptr))
    return out


# ---------------------------------------------------------------- peaks
def tom_peakinc(a, ctx: Optional[Context] = None) -> Tuple[float, np.ndarray]:
    """(max value, 0-based [x y z]) of the first maximum in x-fastest order."""
    b = _Buf(a)
    ctx = _prep_ctx(ctx, b)
    d = np.asarray(_dims3(b.dims), np.int32)
    val = np.zeros(1, np.float32)
    pos = np.zeros(3, np.int32)
    ctx.check(ctx.lib.tomhip_peak_f32(ctx.h, b.ptr, _p(d), _p(val), _p(pos)))
    return float(val[0]), pos


def tom_peak(a, r=None, flag=None, ctx: Optional[Context] = None):
    """tom_peak in its MATLAB forms (Analysis/tom_peak.m), positions 1-based:
      [c val]   = tom_peak(A)            every position holding the maximum,
                                         one row each, in find order (x fastest)
      [c val]   = tom_peak(A, 'spline')  first maximum refined per axis to 0.001
      [c val M] = tom_peak(A, R)         first maximum; M = A with |d| <= R zeroed
      [c val M] = tom_peak(A, R, 'spline')
    A single maximum is returned as a vector (as MATLAB's one-row c)."""
    b = _Buf(a)
    ctx = _prep_ctx(ctx, b)
    nd = max(len(b.dims), 2)
    N = _dims3(b.dims)
    if isinstance(r, str):
        r, flag = None, r
    spline = flag == "spline"
    if flag is not None and not spline:
        raise ValueError("the only flag is 'spline'")
    if r is None and not spline:
        cap = 1 << 16
        pos = np.zeros((cap, 3), np.int32)
        val = np.zeros(1, np.float32)
        cnt = np.zeros(1, np.int64)
        ctx.check(ctx.lib.tomhip_peak_all_f32(ctx.h, b.ptr, _p(N), _p(val), _p(pos), cap, _p(cnt)))
        c = (pos[:int(cnt[0]), :nd] + 1).astype(np.float64)
        return (c[0] if len(c) == 1 else c), float(val[0])
    c = np.zeros(3, np.float64)
    v = np.zeros(1, np.float64)
    out = None
    if r is not None:
        out = _empty_like(b, b.dims)
    ctx.check(ctx.lib.tomhip_peak_radius_f32(ctx.h, b.ptr, _p(N), int(r) if r is not None else 0, int(spline),
                                             _p(c), _p(v), None if out is None else _ptr(out)))
    c = c[:nd] + 1
    if r is None:
        return c, float(v[0])
    return c, float(v[0]), out


def tom_os3_bestPeak(res, ccc, psr=None, soc=None, angle: int = 0, rule: str = "ccc", ctx: Optional[Context] = None):
    """In-place tom_os3_bestPeak update of res = dict(ccc=, psr=, soc=, angle=)
    result maps (torch tensors on the device or numpy arrays; start them at
    -1000): rule is any '+'-joined subset of 'ccc', 'psr', 'soc' (the
    reference's options.analysis flags; all three is os3_results::update)."""
    bits = {"ccc": 1, "psr": 2, "soc": 4}
    code = 0
    for k in rule.split("+"):
        code |= bits[k]
    cb = _Buf(ccc)
    ctx = _prep_ctx(ctx, cb)
    N = _dims3(cb.dims)

    def ptr(x):
        return None if x is None else _ptr(x)
    for k in ("ccc", "psr", "soc", "angle"):
        x = res.get(k)
        if x is not None and not _is_torch(x) and not (x.flags.f_contiguous or x.ndim < 2):
            raise ValueError(f"res['{k}'] must be a contiguous device tensor or a Fortran-ordered array")
    # keep the staged psr / soc arrays alive until the call returns
    pb = None if psr is None else _Buf(psr)
    qb = None if soc is None else _Buf(soc)
    ctx.check(ctx.lib.tomhip_os3_bestpeak_f32(ctx.h, _p(N), code, int(angle), cb.ptr,
                                              None if pb is None else pb.ptr,
                                              None if qb is None else qb.ptr, ptr(res["ccc"]),
                                              ptr(res.get("psr")) if psr is not None else None,
                                              ptr(res.get("soc")) if soc is not None else None,
                                              ptr(res["angle"])))
    return res


def flcf_box(search, template, mask=None, ctx: Optional[Context] = None):
    """tom_os3_corr(search, template) for a search volume of the template's
    size (circular; with search = template the autocorrelation of the
    second-order correlation)."""
    sb, tb = _Buf(search), _Buf(template)
    mb = _Buf(mask) if mask is not None else None
    ctx = _prep_ctx(ctx, sb, tb, mb)
    n = _dims3(tb.dims)
    out = _empty_like(tb, tb.dims)
    ctx.check(ctx.lib.tomhip_flcf_box_f32(ctx.h, sb.ptr, tb.ptr, None if mb is None else mb.ptr, _p(n), _ptr(out)))
    return out


def os3_picker(volume, template, mask, angles, rule: str = "ccc+psr+soc", ctx: Optional[Context] = None):
    """The per-orientation loop of the reference's correlation picker with
    its secondary scores (Oscar3/c++Picker/src/tom/os3/os3_picker.cpp:29-121;
    MATLAB tom_os3_findTemplate.m:255-307 with options.analysis set): for each
    orientation a, t = tom_rotate(T, a) .* M; ccc = tom_os3_corr(V, t, M);
    psr = the PSR map of ccc; soc = tom_os3_corr(ccc, flcf_box(t, t, M), M);
    then tom_os3_bestPeak(rule). volume: a (Z, Y, X) float32 CUDA tensor.
    Returns dict(ccc, psr, soc, angle (1-based, 0 = never updated)) as tensors."""
    ctx = ctx or context()
    dev = volume.device
    n = tuple(np.asarray(template).shape)
    Td = torch.from_numpy(np.ascontiguousarray(np.asarray(template, np.float32).transpose(2, 1, 0))).to(dev)
    Md = torch.from_numpy(np.ascontiguousarray(np.asarray(mask, np.float32).transpose(2, 1, 0))).to(dev)
    res = {k: torch.full(volume.shape, -1000.0, dtype=torch.float32, device=dev) for k in ("ccc", "psr", "soc")}
    res["angle"] = torch.zeros(volume.shape, dtype=torch.int32, device=dev)
    need_psr = "psr" in rule
    need_soc = "soc" in rule
    for i, a in enumerate(np.asarray(angles, np.float32).reshape(-1, 3)):
        t = tom_rotate(Td, a, ctx=ctx) * Md
        ccc = tom_os3_corr(volume, t, Md, ctx=ctx)
        psr = tom_os3_psr(ccc, n, ctx=ctx) if need_psr else None
        soc = None
        if need_soc:
            auto = flcf_box(t, t, Md, ctx=ctx)
            soc = tom_os3_corr(ccc, auto, Md, ctx=ctx)
        tom_os3_bestPeak(res, ccc, psr, soc, angle=i + 1, rule=rule, ctx=ctx)
    return res


# ---------------------------------------------------------------- masks/weights
def tom_os3_sphereMask(template_shape, ctx: Optional[Context] = None):
    ctx = ctx or context()
    d = _dims3(template_shape)
    out = np.zeros(tuple(int(x) for x in d), np.float32, order="F")
    cnt = np.zeros(1, np.int64)
    ctx.check(ctx.lib.tomhip_spheremask_os3_f32(ctx.h, _p(d), _p(out), _p(cnt)))
    return out, int(cnt[0])


def tom_spheremask_gauss(shape, radius: float, sigma: float, max_radius: float = -1.0,
                         ctx: Optional[Context] = None):
    ctx = ctx or context()
    d = _dims3(shape)
    out = np.zeros(tuple(int(x) for x in d), np.float32, order="F")
    ctx.check(ctx.lib.tomhip_spheremask_f32(ctx.h, _p(d), float(radius), float(sigma), float(max_radius),
                                            _p(out)))
    return out


def weight(shape, semi_angle_deg: float = 0.0, low: float = 0.0, high: float = 0.0,
           ctx: Optional[Context] = None):
    """Centred wedge (tom_wedge semi-angle) times bandpass annulus."""
    ctx = ctx or context()
    d = _dims3(shape)
    out = np.zeros(tuple(int(x) for x in d), np.float32, order="F")
    ctx.check(ctx.lib.tomhip_weight_f32(ctx.h, _p(d), float(semi_angle_deg), float(low), float(high), _p(out)))
    return out


def tom_wedge(shape, semi_angle_deg: float, ctx: Optional[Context] = None):
    return weight(shape, semi_angle_deg, 0.0, 0.0, ctx)


# ---------------------------------------------------------------- template prep
def prepare_template(template, mask, W=None, ctx: Optional[Context] = None):
    """(Tn, mu_T, sig_T): tom_norm(1e5*(T+1),'mean0+1std') and the fixed
    template statistics of the (W-filtered) normalised template."""
    tb, mb = _Buf(template), _Buf(mask)
    wb = _Buf(W) if W is not None else None
    ctx = _prep_ctx(ctx, tb, mb, wb)
    d = _dims3(tb.dims)
    out = _empty_like(tb, tb.dims)
    mu = np.zeros(1)
    sg = np.zeros(1)
    ctx.check(ctx.lib.tomhip_prepare_template_f32(ctx.h, tb.ptr, _p(d), mb.ptr, None if wb is None else wb.ptr,
                                                  _ptr(out), _p(mu), _p(sg)))
    return out, float(mu[0]), float(sg[0])


# ---------------------------------------------------------------- search
class SearchPlan:
    """Exhaustive search over one volume (FLCF or tom_oscar XCF)."""

    def __init__(self, volume, template_shape, mask=None, mode: str = "flcf", batch: int = 4,
                 ctx: Optional[Context] = None):
        vb = _Buf(volume)
        mb = _Buf(mask) if mask is not None else None
        self.ctx = _prep_ctx(ctx, vb, mb)
        self.N = _dims3(vb.dims)
        self.n = _dims3(template_shape)
        self._torch = vb.torch
        self._device = vb.obj.device if vb.torch else None
        self.mode = {"flcf": 0, "xcf": 1}[mode]
        h = ctypes.c_void_p()
        self.ctx.check(self.ctx.lib.tomhip_search_create(self.ctx.h, vb.ptr, _p(self.N),
                                                         None if mb is None else mb.ptr, _p(self.n), self.mode,
                                                         int(batch), ctypes.byref(h)))
        self.h = h

    def set_template(self, tmpl, W=None, mu_t: float = 0.0, sig_t: float = 1.0, remask: bool = True):
        tb = _Buf(tmpl)
        wb = _Buf(W) if W is not None else None
        self.ctx.check(self.ctx.lib.tomhip_search_set_template(self.h, tb.ptr, None if wb is None else wb.ptr,
                                                               float(mu_t), float(sig_t), int(remask)))

    def run(self, angles, first_index: int = 0):
        ang = np.ascontiguousarray(np.asarray(angles, np.float32).reshape(-1, 3))
        self.ctx.check(self.ctx.lib.tomhip_search_run(self.h, _p(ang), ang.shape[0], int(first_index)))

    def finalize(self):
        """(score, 0-based angle index) maps, numpy (nx,ny,nz) or torch (nz,ny,nx)."""
        dims = tuple(int(x) for x in self.N)
        if self._torch:
            sc = torch.empty(tuple(reversed(dims)), dtype=torch.float32, device=self._device)
            an = torch.empty(tuple(reversed(dims)), dtype=torch.int32, device=self._device)
        else:
            sc = np.zeros(dims, np.float32, order="F")
            an = np.zeros(dims, np.int32, order="F")
        self.ctx.check(self.ctx.lib.tomhip_search_finalize(self.h, _ptr(sc), _ptr(an)))
        return sc, an

    def refine(self, volume, group_angles, cand, centroid=None):
        """Exact FLCF around candidates: group_angles (A, G, 3) [phi psi theta]
        (A orientations for each of G groups), cand (P, 4) int rows
        x y z group (0-based). centroid (3,) x y z: the candidates locate the
        rotated template point R d + p instead of the centre p (chunks of 4
        orientations are scored around cand - round(mean R d)). Returns
        (score[P], angle index[P] into the group's list, position[P, 3] of
        the best centre, x y z 0-based)."""
        vb = _Buf(volume)
        _prep_ctx(self.ctx, vb)
        ga = np.ascontiguousarray(np.asarray(group_angles, np.float32))
        if ga.ndim != 3 or ga.shape[2] != 3:
            raise ValueError("group_angles must be (A, G, 3)")
        A, G = ga.shape[0], ga.shape[1]
        c = np.ascontiguousarray(np.asarray(cand, np.int32).reshape(-1, 4))
        P = c.shape[0]
        sc = np.zeros(P, np.float32)
        an = np.zeros(P, np.int32)
        pos = np.zeros((P, 3), np.int32)
        cen = None if centroid is None else np.ascontiguousarray(np.asarray(centroid, np.float32).reshape(3))
        self.ctx.check(self.ctx.lib.tomhip_search_refine(self.h, vb.ptr, _p(ga), G, A, _p(c), P,
                                                         None if cen is None else _p(cen), _p(sc), _p(an), _p(pos)))
        return sc, an, pos

    def set_tiles(self, tile_z: int = 16, tile_y: int = 16):
        """Column-tile widths of the z- and y-FFT passes (8, 16 or 32)."""
        self.ctx.check(self.ctx.lib.tomhip_search_set_tiles(self.h, int(tile_z), int(tile_y)))

    def set_zchunk(self, planes: int = 0):
        """Run the y-FFT and c2r/max passes over z-chunks (0 = whole volume)."""
        self.ctx.check(self.ctx.lib.tomhip_search_set_zchunk(self.h, int(planes)))

    def set_graph(self, enable: bool):
        """Replay full angle batches through a captured hipGraph (default)."""
        self.ctx.check(self.ctx.lib.tomhip_search_set_graph(self.h, int(enable)))

    def profile(self, enable: bool):
        """Enable/disable per-kernel event timing; returns (ms[8], counts[8])
        accumulated since the previous call."""
        ms = np.zeros(8)
        cnt = np.zeros(8, np.int64)
        self.ctx.check(self.ctx.lib.tomhip_search_profile(self.h, int(enable), _p(ms), _p(cnt)))
        return ms, cnt

    def close(self):
        if self.h:
            self.ctx.lib.tomhip_search_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def tom_os3_findTemplate(volume, template, mask=None, angles=None, W=None, batch: int = 4,
                         ctx: Optional[Context] = None):
    """ccc-only tom_os3_findTemplate: (ccc map, 1-based angle index map)."""
    vb, tb = _Buf(volume), _Buf(template)
    mb = _Buf(mask) if mask is not None else None
    wb = _Buf(W) if W is not None else None
    ctx = _prep_ctx(ctx, vb, tb, mb, wb)
    ang = np.ascontiguousarray(np.asarray(angles, np.float32).reshape(-1, 3))
    N, n = _dims3(vb.dims), _dims3(tb.dims)
    sc = _empty_like(vb, vb.dims)
    an = _empty_like(vb, vb.dims, np.int32)
    ctx.check(ctx.lib.tomhip_flcf_search_f32(ctx.h, vb.ptr, _p(N), tb.ptr, _p(n), None if mb is None else mb.ptr,
                                             None if wb is None else wb.ptr, _p(ang), ang.shape[0], int(batch),
                                             _ptr(sc), _ptr(an)))
    return sc, an + 1


def tom_os3_corr(img, template, mask=None, ctx: Optional[Context] = None):
    """FLCF score map of tom_os3_corr(img, template, options)."""
    vb, tb = _Buf(img), _Buf(template)
    mb = _Buf(mask) if mask is not None else None
    ctx = _prep_ctx(ctx, vb, tb, mb)
    N, n = _dims3(vb.dims), _dims3(tb.dims)
    out = _empty_like(vb, vb.dims)
    ctx.check(ctx.lib.tomhip_os3_corr_f32(ctx.h, vb.ptr, _p(N), tb.ptr, _p(n), None if mb is None else mb.ptr,
                                          _ptr(out)))
    return out


def _angle_list_1d(start: float, end: float, inc: float):
    """One axis of tools/tom_os3_angleList.m:8-39: start:inc:end mod 360 when
    start < end, else start+1:inc:end+359 mod 360 (wrap-around range)."""
    if start < end:
        return [float(np.mod(v, 360)) for v in _colon(start, inc, end)]
    return [float(np.mod(v, 360)) for v in _colon(start + 1, inc, end + 360 - 1)]


def _rotmatrix(phi, psi, theta):
    """tom_angles2rotmatrix -> tom_sum_rotation (Sptrans/tom_sum_rotation.m:115-124):
    Rz(psi) Rx(theta) Rz(phi) in double."""
    f, p, t = np.deg2rad([phi, psi, theta])
    rz_p = np.array([[np.cos(p), -np.sin(p), 0], [np.sin(p), np.cos(p), 0], [0, 0, 1.0]])
    rx_t = np.array([[1.0, 0, 0], [0, np.cos(t), -np.sin(t)], [0, np.sin(t), np.cos(t)]])
    rz_f = np.array([[np.cos(f), -np.sin(f), 0], [np.sin(f), np.cos(f), 0], [0, 0, 1.0]])
    return (rz_p @ rx_t) @ rz_f


def tom_os3_angleList(start, end, increment, dedup: bool = True) -> np.ndarray:
    """Angle list of Oscar3/matlabPicker/tools/tom_os3_angleList.m:41-95 (3D):
    phi outermost, psi, theta innermost; with dedup a triple is dropped when
    floor(1e4 * its rotation matrix) equals that of an earlier kept triple.
    Returns an (A, 3) float32 array of [phi psi theta] rows."""
    axes = [_angle_list_1d(start[i], end[i], increment[i]) for i in range(3)]
    out, seen = [], set()
    for ph in axes[0]:
        for ps in axes[1]:
            for th in axes[2]:
                if dedup:
                    key = tuple(np.floor(_rotmatrix(ph, ps, th) * 10000).astype(np.int64).ravel())
                    if key in seen:
                        continue
                    seen.add(key)
                out.append((ph, ps, th))
    return np.asarray(out, np.float32).reshape(-1, 3)


def subtomo_score_batch(boxes, particles, particle_box, wedge_semi_deg, angles, ctx: Optional[Context] = None):
    """Batched local angular search of subtomograms (tom_av3_corr3d 'peak' /
    corr3d_batch semantics, one persistent launch).

    boxes: list of dicts {"template": (B,B,B), "mask": (B,B,B), "mask_cc":
    (B,B,B) or None, "wedge_ref": semi-angle of the template's missing wedge in
    degrees or None}, one per box edge B in {32, 48, 64} (MATLAB indexing).
    particles: list of (B,B,B) arrays, or one contiguous float32 torch tensor
    holding them packed in order (x fastest); particle_box[p] indexes boxes;
    wedge_semi_deg[p]: the particle's wedge semi-angle (negative or NaN: none,
    0: the cutoff sphere only, as tom_mex_corr3d treats a scalar 0);
    angles: (P, A, 3) degrees.
    Returns (best angle index 0-based (-1 none), peak (P,3) 0-based unshifted,
    ccc (P,))."""
    ctx = ctx or context()
    edges = np.asarray([np.asarray(b["template"]).shape[0] for b in boxes], np.int32)
    T = np.concatenate([np.asfortranarray(b["template"], np.float32).ravel(order="F") for b in boxes])
    M = np.concatenate([np.asfortranarray(b["mask"], np.float32).ravel(order="F") for b in boxes])
    if any(b.get("mask_cc") is not None for b in boxes):
        Mcc = np.concatenate([np.asfortranarray(b["mask_cc"] if b.get("mask_cc") is not None
                                                else np.ones(b["template"].shape), np.float32).ravel(order="F")
                              for b in boxes])
    else:
        Mcc = None
    wref = None
    if any(b.get("wedge_ref") is not None for b in boxes):
        wref = np.asarray([np.nan if b.get("wedge_ref") is None else float(b["wedge_ref"]) for b in boxes], np.float32)
    pb = np.ascontiguousarray(particle_box, np.int32)
    P = pb.shape[0]
    sizes = [int(edges[i]) ** 3 for i in pb]
    off = np.zeros(P, np.int64)
    off[1:] = np.cumsum(sizes)[:-1]
    if _is_torch(particles):
        # already packed on the device: particle p at off[p], x fastest
        if particles.dtype != torch.float32 or not particles.is_contiguous() or particles.numel() < sum(sizes):
            raise ValueError("packed particles must be a contiguous float32 tensor holding every particle")
        packed = particles
        ctx.use_torch_stream()
    else:
        packed = np.concatenate([np.asfortranarray(v, np.float32).ravel(order="F") for v in particles])
    semi = np.ascontiguousarray(wedge_semi_deg, np.float32)
    ang = np.ascontiguousarray(np.asarray(angles, np.float32).reshape(P, -1, 3))
    A = ang.shape[1]
    best = np.zeros(P, np.int32)
    peak = np.zeros((P, 3), np.int32)
    ccc = np.zeros(P, np.float32)
    ctx.check(ctx.lib.tomhip_subtomo_score_batch(ctx.h, len(boxes), _p(edges), _p(T), _p(M),
                                                 None if Mcc is None else _p(Mcc),
                                                 None if wref is None else _p(wref), P, _p(pb), _ptr(packed),
                                                 _p(off), _p(semi), _p(ang), A, _p(best), _p(peak), _p(ccc)))
    return best, peak, ccc


def tom_os3_returnPicklist(peaks, template_size, n, angles=None, erase_borders: bool = True,
                           ctx: Optional[Context] = None) -> np.ndarray:
    """Picklist rows [x y z (1-based) value angle] of tom_os3_returnPicklist
    (no binning): iterated tom_peakc with radius ceil(template_size/2)."""
    pb = _Buf(peaks)
    ab = _Buf(angles, np.int32) if angles is not None else None
    ctx = _prep_ctx(ctx, pb, ab)
    N = _dims3(pb.dims)
    t = _dims3(template_size if np.ndim(template_size) else (template_size,) * 3)
    out = np.zeros((max(int(n), 1), 5), np.float32)
    cnt = np.zeros(1, np.int64)
    ctx.check(ctx.lib.tomhip_picklist_f32(ctx.h, pb.ptr, _p(N), None if ab is None else ab.ptr, _p(t), int(n),
                                          int(erase_borders), _p(out), _p(cnt)))
    return out[:int(cnt[0])]


_EM_HEADER = np.dtype([("machine", "i1"), ("byte2", "i1"), ("byte3", "i1"), ("type", "i1"), ("dims", "<u4", 3),
                       ("comment", "S80"), ("emdata", "<i4", 40), ("userdata", "S256")])


def tom_emread(path: str):
    """(volume (nx, ny, nz) float32 in MATLAB indexing, header record) of an
    EM file (IOfun/tom_emread.m), read by the library's C++ reader."""
    lib = load()
    hdr = np.zeros(1, _EM_HEADER)
    rc = lib.tomhip_em_read_header(path.encode(), _p(hdr))
    if rc != 0:
        raise TomHipError(rc, lib.tomhip_em_last_error().decode())
    dims = tuple(int(d) for d in hdr["dims"][0])
    out = np.zeros(dims, np.float32, order="F")
    rc = lib.tomhip_em_read_f32(path.encode(), _p(out), out.size, _p(hdr))
    if rc != 0:
        raise TomHipError(rc, lib.tomhip_em_last_error().decode())
    return out, hdr[0]


def tom_emwrite(path: str, vol, header=None) -> None:
    """Write a float32 PC-format EM file (tom_emwrite)."""
    lib = load()
    v = np.asfortranarray(vol, np.float32)
    d = _dims3(v.shape)
    h = None
    if header is not None:
        h = np.zeros(1, _EM_HEADER)
        h[0] = header
    rc = lib.tomhip_em_write_f32(path.encode(), _p(v), _p(d), None if h is None else _p(h))
    if rc != 0:
        raise TomHipError(rc, lib.tomhip_em_last_error().decode())


def _colon(a: float, step: float, b: float) -> np.ndarray:
    """MATLAB a:step:b."""
    if step == 0 or (step > 0 and a > b) or (step < 0 and a < b):
        return np.zeros(0)
    n = int(np.floor((b - a) / step + 1e-10)) + 1
    return a + step * np.arange(n)


def tom_oscar(vol, template, phi_start, phi_end, phi_incr, psi_start, psi_end, psi_incr, the_start, the_end,
              the_incr, ctx: Optional[Context] = None):
    """[ccfmax, phi_max, psi_max, theta_max] = tom_oscar(...) (int16 angle
    maps, -1 where no angle improved on the initial -1000)."""
    vb, tb = _Buf(vol), _Buf(template)
    ctx = _prep_ctx(ctx, vb, tb)
    ang = np.asarray([(p, s, t) for p in _colon(phi_start, phi_incr, phi_end)
                      for s in _colon(psi_start, psi_incr, psi_end)
                      for t in _colon(the_start, the_incr, the_end)], np.float32).reshape(-1, 3)
    ang = np.ascontiguousarray(ang)
    N, n = _dims3(vb.dims), _dims3(tb.dims)
    sc = _empty_like(vb, vb.dims)
    an = _empty_like(vb, vb.dims, np.int32)
    ctx.check(ctx.lib.tomhip_oscar_f32(ctx.h, vb.ptr, _p(N), tb.ptr, _p(n), _p(ang),
                                       ang.shape[0], _ptr(sc), _ptr(an)))
    idx = an.cpu().numpy() if _is_torch(an) else an
    ccf = sc.cpu().numpy() if _is_torch(sc) else sc
    # tom_oscar keeps -1000 / -1 where nothing beat the initial value
    never = (idx < 0) | (ccf <= -1000)
    safe = np.where(never, 0, idx)
    maps = [np.where(never, -1, ang[safe, k]).astype(np.int16) for k in range(3)]
    ccf = np.where(never, -1000.0, ccf)
    return ccf, maps[0], maps[1], maps[2]


# ---------------------------------------------------------------- learned proposal
class Unet:
    """3D U-Net proposal network (tomhip_unet_*): per-voxel particle
    probability and coarse-orientation bin. `dims` = (nx, ny, nz) of the
    network input (multiples of 4, >= 8); `train=False` allocates no gradient
    buffers. Inputs may be numpy (F-order (nx,ny,nz)) or torch ((nz,ny,nx))."""

    HEAD = 80

    def __init__(self, dims, nbin: int = 64, seed: int = 0, train: bool = True, ctx: Optional[Context] = None):
        self.ctx = ctx or context()
        self.dims = _dims3(dims)
        self.nbin = int(nbin)
        self.V = int(np.prod(self.dims))
        h = ctypes.c_void_p()
        self.ctx.check(self.ctx.lib.tomhip_unet_create(self.ctx.h, _p(self.dims), self.nbin, int(seed), int(train),
                                                       ctypes.byref(h)))
        self.h = h

    def num_params(self) -> int:
        n = np.zeros(1, np.int64)
        self.ctx.check(self.ctx.lib.tomhip_unet_num_params(self.h, _p(n)))
        return int(n[0])

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.num_params(), np.float32)
        self.ctx.check(self.ctx.lib.tomhip_unet_get_params(self.h, _p(out)))
        return out

    def set_params(self, theta):
        t = np.ascontiguousarray(np.asarray(theta, np.float32))
        if t.size != self.num_params():
            raise ValueError("parameter vector has the wrong length")
        self.ctx.check(self.ctx.lib.tomhip_unet_set_params(self.h, _p(t)))

    def get_optimizer(self):
        """Adam state (m, v, step) of a training network."""
        n = self.num_params()
        m = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        st = np.zeros(1, np.int64)
        self.ctx.check(self.ctx.lib.tomhip_unet_get_optimizer(self.h, _p(m), _p(v), _p(st)))
        return m, v, int(st[0])

    def set_optimizer(self, m, v, step: int):
        n = self.num_params()
        m = np.ascontiguousarray(np.asarray(m, np.float32))
        v = np.ascontiguousarray(np.asarray(v, np.float32))
        if m.size != n or v.size != n:
            raise ValueError("optimizer state has the wrong length")
        self.ctx.check(self.ctx.lib.tomhip_unet_set_optimizer(self.h, _p(m), _p(v), int(step)))

    def forward(self, vol, aux=None):
        """vol (and the optional aux map): normalised inputs of the network size."""
        vb = _Buf(vol)
        ab = _Buf(aux) if aux is not None else None
        for b in (vb, ab):
            if b is not None and tuple(int(x) for x in b.dims) != tuple(int(x) for x in self.dims):
                raise ValueError("input does not match the network input size")
        _prep_ctx(self.ctx, vb, ab)
        self.ctx.check(self.ctx.lib.tomhip_unet_forward(self.h, vb.ptr, None if ab is None else ab.ptr))

    def logits(self) -> np.ndarray:
        """float32 (nz, ny, nx, 80)."""
        out = np.zeros((self.V, self.HEAD), np.float32)
        self.ctx.check(self.ctx.lib.tomhip_unet_get_logits(self.h, _p(out)))
        nx, ny, nz = (int(x) for x in self.dims)
        return out.reshape(nz, ny, nx, self.HEAD)

    def backward(self, tprob, tbin, wpos: float = 1.0, toff=None):
        """Gradients of BCE(tprob) + CE(tbin) [+ offset squared error against
        toff, whose memory is [V][4] = (dx/8, dy/8, dz/8, weight): a torch
        (nz, ny, nx, 4) tensor or a numpy (4, nx, ny, nz) F-order array];
        returns (bce, ce, offset)."""
        pb, bb = _Buf(tprob), _Buf(tbin, np.int32)
        ob = _Buf(toff) if toff is not None else None
        _prep_ctx(self.ctx, pb, bb, ob)
        loss = np.zeros(3)
        self.ctx.check(self.ctx.lib.tomhip_unet_backward(self.h, pb.ptr, bb.ptr, None if ob is None else ob.ptr,
                                                         float(wpos), _p(loss)))
        return float(loss[0]), float(loss[1]), float(loss[2])

    def get_grads(self) -> np.ndarray:
        out = np.zeros(self.num_params(), np.float32)
        self.ctx.check(self.ctx.lib.tomhip_unet_get_grads(self.h, _p(out)))
        return out

    def adam(self, lr: float):
        self.ctx.check(self.ctx.lib.tomhip_unet_adam(self.h, float(lr)))

    def _offsets_like(self, buf, dims):
        if buf is not None and buf.torch:
            return torch.empty(tuple(reversed(tuple(dims))) + (3,), dtype=torch.float32, device=buf.obj.device)
        return np.zeros((3,) + tuple(dims), np.float32, order="F")

    def predict(self, like=None, offsets: bool = False):
        """(prob, bin[, offsets]) of the last forward pass, shaped like the
        input; offsets: torch (nz, ny, nx, 3) or numpy (3, nx, ny, nz), the
        predicted (dx, dy, dz) / 8 to the particle centre."""
        ref = _Buf(like) if like is not None else None
        dims = tuple(int(x) for x in self.dims)
        if ref is not None and ref.torch:
            prob = torch.empty(tuple(reversed(dims)), dtype=torch.float32, device=ref.obj.device)
            b = torch.empty(tuple(reversed(dims)), dtype=torch.int32, device=ref.obj.device)
        else:
            prob = np.zeros(dims, np.float32, order="F")
            b = np.zeros(dims, np.int32, order="F")
        o = self._offsets_like(ref, dims) if offsets else None
        self.ctx.check(self.ctx.lib.tomhip_unet_predict(self.h, _ptr(prob), _ptr(b), None if o is None else _ptr(o)))
        return (prob, b, o) if offsets else (prob, b)

    def infer(self, vol, halo: int = 24, aux=None, logit: bool = False, offsets: bool = False):
        """Whole-volume (prob, bin[, offsets]) maps in overlapping tiles of the
        network size; the volume (and the aux map) are normalised by their
        global mean and std. logit: return the particle logit instead of its
        sigmoid; offsets as in predict()."""
        vb = _Buf(vol)
        ab = _Buf(aux) if aux is not None else None
        ctx = _prep_ctx(self.ctx, vb, ab)
        N = _dims3(vb.dims)
        prob = _empty_like(vb, vb.dims)
        b = _empty_like(vb, vb.dims, np.int32)
        o = self._offsets_like(vb, vb.dims) if offsets else None
        ctx.check(ctx.lib.tomhip_unet_infer(self.h, vb.ptr, None if ab is None else ab.ptr, _p(N), int(halo),
                                            1 if logit else 0, _ptr(prob), _ptr(b), None if o is None else _ptr(o)))
        return (prob, b, o) if offsets else (prob, b)

    def close(self):
        if self.h:
            self.ctx.lib.tomhip_unet_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def local_maxima(peaks, radius: int, threshold: float, capacity: int = 1 << 20, ctx: Optional[Context] = None):
    """Local maxima of a map above threshold (cube of half-edge radius, ties
    to the lowest index): (P, 4) float32 rows x y z (0-based) value, sorted by
    value (descending, then index). Raises if more than capacity are found."""
    mb = _Buf(peaks)
    ctx = _prep_ctx(ctx, mb)
    N = _dims3(mb.dims)
    out = np.zeros((int(capacity), 4), np.float32)
    cnt = np.zeros(1, np.int64)
    ctx.check(ctx.lib.tomhip_local_maxima_f32(ctx.h, mb.ptr, _p(N), int(radius), float(threshold), int(capacity),
                                              _p(out), _p(cnt)))
    k = int(cnt[0])
    if k > capacity:
        raise RuntimeError(f"{k} local maxima exceed the capacity {capacity}; raise the threshold or the capacity")
    out = out[:k]
    lin = (out[:, 2].astype(np.int64) * int(N[1]) + out[:, 1].astype(np.int64)) * int(N[0]) + out[:, 0].astype(np.int64)
    order = np.lexsort((lin, -out[:, 3]))
    return out[order]


def tom_os3_psr(xcf, template_size, ctx: Optional[Context] = None):
    """Peak-to-sidelobe ratio map of a correlation map (os3_psr): local
    normalisation under the ring of the template's size."""
    xb = _Buf(xcf)
    ctx = _prep_ctx(ctx, xb)
    N = _dims3(xb.dims)
    n = _dims3(template_size if np.ndim(template_size) else (template_size,) * 3)
    out = _empty_like(xb, xb.dims)
    ctx.check(ctx.lib.tomhip_psr_f32(ctx.h, xb.ptr, _p(N), _p(n), _ptr(out)))
    return out


def tom_bin(im, nbin: int = 1, ctx: Optional[Context] = None):
    """tom_bin(im, nbin) of a 3D volume: nbin rounds of 2x2x2 averaging
    (odd trailing samples dropped), shaped like the input's kind."""
    ib = _Buf(im)
    ctx = _prep_ctx(ctx, ib)
    N = _dims3(ib.dims)
    out_dims = tuple(int(d) >> int(nbin) for d in N)
    out = _empty_like(ib, out_dims)
    ctx.check(ctx.lib.tomhip_bin_f32(ctx.h, ib.ptr, _p(N), int(nbin), _ptr(out)))
    return out


def snap_maxima(peaks, cand_xyz, radius: int, ctx: Optional[Context] = None):
    """Move each (x, y, z) row to the largest map value within the cube of
    half-edge radius around it; returns the new (P, 3) int32 rows."""
    mb = _Buf(peaks)
    ctx = _prep_ctx(ctx, mb)
    N = _dims3(mb.dims)
    c = np.array(np.asarray(cand_xyz, np.int32).reshape(-1, 3), dtype=np.int32, order="C", copy=True)
    if c.shape[0]:
        ctx.check(ctx.lib.tomhip_snap_maxima_f32(ctx.h, mb.ptr, _p(N), int(radius), _p(c), c.shape[0], 3))
    return c
