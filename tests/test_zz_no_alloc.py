"""The allocation-free global max (tomhip_peak_f32_noalloc) through the C ABI:
equal to tomhip_peak_f32 (bit-exact value, same first position in x-fastest
order) on two calls in a row, the second on a smaller map, on ties, and on an
all-below-start map; a host map is refused with -1."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(lib, ctx, fn, m, dims):
    d = np.asarray(dims, np.int32)
    v = np.zeros(1, np.float32)
    pos = np.full(3, -7, np.int32)
    rc = getattr(lib, fn)(ctx.h, ctypes.c_void_p(m.data_ptr()), ctypes.c_void_p(d.ctypes.data),
                          ctypes.c_void_p(v.ctypes.data), ctypes.c_void_p(pos.ctypes.data))
    assert rc == 0, lib.tomhip_last_error(ctx.h)
    return float(v[0]), pos.tolist()


def test_noalloc_peak_matches_the_allocating_entry_point():
    import torch

    from tom_amd import api
    print("peak noalloc vs peak", flush=True)
    ctx = api.context()
    lib = ctx.lib
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(17)
    cases = []
    big = torch.randn((128, 256, 256), generator=g, device=dev)  # (z, y, x)
    cases.append((big, (256, 256, 128)))
    small = torch.randn((11, 23, 37), generator=g, device=dev)
    cases.append((small, (37, 23, 11)))
    ties = torch.zeros((9, 10, 11), device=dev)
    ties[3, 4, 5] = ties[7, 1, 2] = ties[3, 4, 6] = 2.5
    cases.append((ties, (11, 10, 9)))
    low = torch.full((5, 6, 7), -1e6, device=dev)
    cases.append((low, (7, 6, 5)))
    for m, dims in cases:
        a = _call(lib, ctx, "tomhip_peak_f32", m, dims)
        b = _call(lib, ctx, "tomhip_peak_f32_noalloc", m, dims)
        print(f"  dims {dims}: {a} {b}", flush=True)
        assert a == b
    # reference semantics on the tie case: first in x-fastest order (z = 3, y = 4, x = 5)
    assert _call(lib, ctx, "tomhip_peak_f32_noalloc", ties, (11, 10, 9)) == (2.5, [5, 4, 3])
    # nothing above the -99999 start: value -99999, positions untouched
    assert _call(lib, ctx, "tomhip_peak_f32_noalloc", low, (7, 6, 5)) == (-99999.0, [-7, -7, -7])
    # the map must already be on the device
    h = np.ones(8, np.float32)
    d = np.asarray([2, 2, 2], np.int32)
    v = np.zeros(1, np.float32)
    pos = np.zeros(3, np.int32)
    rc = lib.tomhip_peak_f32_noalloc(ctx.h, ctypes.c_void_p(h.ctypes.data), ctypes.c_void_p(d.ctypes.data),
                                     ctypes.c_void_p(v.ctypes.data), ctypes.c_void_p(pos.ctypes.data))
    assert rc == -1 and b"device memory" in lib.tomhip_last_error(ctx.h)
