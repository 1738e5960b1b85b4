"""EM file I/O (host, no GPU) and the GPU picklist against the oracle."""
import os

import numpy as np
import pytest


def test_em_roundtrip_and_layout(tmp_path):
    from tom_amd import api
    rng = np.random.default_rng(0)
    v = rng.standard_normal((7, 5, 3)).astype(np.float32)
    path = str(tmp_path / "a.em")
    api.tom_emwrite(path, v)
    raw = open(path, "rb").read()
    assert len(raw) == 512 + 4 * v.size
    assert raw[0] == 6 and raw[3] == 5  # PC, float (io.h:49-63)
    assert np.frombuffer(raw[4:16], "<u4").tolist() == [7, 5, 3]
    np.testing.assert_array_equal(np.frombuffer(raw[512:], "<f4"), v.ravel(order="F"))
    w, h = api.tom_emread(path)
    np.testing.assert_array_equal(w, v)
    assert int(h["type"]) == 5


@pytest.mark.parametrize("machine,code,dt", [(3, 2, ">i2"), (5, 9, ">f8"), (6, 4, "<i4"), (0, 1, "i1")])
def test_em_read_types_and_endianness(tmp_path, machine, code, dt):
    """Big-endian machines 0/3/5 and the integer/double types of tom_emread.m:160-215."""
    from tom_amd import api
    rng = np.random.default_rng(code)
    vals = rng.integers(-100, 100, size=(4, 3, 2)).astype(dt)
    hdr = bytearray(512)
    hdr[0], hdr[3] = machine, code
    bo = ">" if machine in (0, 3, 5) else "<"
    hdr[4:16] = np.asarray([4, 3, 2], bo + "u4").tobytes()
    path = str(tmp_path / "b.em")
    open(path, "wb").write(bytes(hdr) + vals.ravel(order="F").tobytes())
    w, h = api.tom_emread(path)
    np.testing.assert_array_equal(w, vals.astype(np.float32))


def test_em_read_refuses_malformed_headers(tmp_path):
    """dims of 0xFFFFFFFF (their product overflows 64 bits) and a header that
    claims more data than the file holds are refused with -1 and a message,
    not a crash or an exception across the C ABI."""
    import ctypes

    from tom_amd import api
    lib = api.load()
    out = np.zeros(64, np.float32)
    for dims, extra in (([0xFFFFFFFF] * 3, 16), ([4, 4, 4], 100), ([4, 4, 4], 252)):
        hdr = bytearray(512)
        hdr[0], hdr[3] = 6, 5
        hdr[4:16] = np.asarray(dims, "<u4").tobytes()
        path = str(tmp_path / "bad.em")
        open(path, "wb").write(bytes(hdr) + bytes(extra))
        rc = lib.tomhip_em_read_f32(path.encode(), ctypes.c_void_p(out.ctypes.data), out.size, None)
        assert rc == -1, dims
        assert lib.tomhip_em_last_error()


@pytest.mark.gpu
def test_picklist_matches_oracle(oracle):
    from tom_amd import api
    rng = np.random.default_rng(3)
    N = (48, 40, 36)
    m = (rng.standard_normal(N) * 0.05).astype(np.float32)
    x = np.arange(N[0])[:, None, None]
    y = np.arange(N[1])[None, :, None]
    z = np.arange(N[2])[None, None, :]
    for _ in range(12):
        c = [rng.integers(8, s - 8) for s in N]
        m += rng.uniform(0.3, 1.0) * np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / 4.0).astype(np.float32)
    ang = rng.integers(0, 1000, N).astype(np.int32)
    got = api.tom_os3_returnPicklist(m, (8, 8, 8), 20, angles=ang)
    ref = oracle.return_picklist(m, ang, (8, 8, 8), 20)
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got[:, :3], ref[:, :3])
    np.testing.assert_allclose(got[:, 3], ref[:, 3], rtol=0, atol=1e-7)
    np.testing.assert_array_equal(got[:, 4], ref[:, 4])
