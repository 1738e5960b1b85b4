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


@pytest.mark.gpu
def test_picklist_cases_match_oracle(oracle):
    """The brick-table pick kernel row for row against the oracle: no border
    erase (peaks at the low borders take tom_paste's shifted write-back),
    all-negative maps (erased zeros become the maximum), integer ties (first
    in x-fastest order), an image (one z plane), a map that runs flat before
    n picks, and maps that give no rows (constant; a NaN)."""
    from tom_amd import api
    rng = np.random.default_rng(17)
    N = (37, 35, 20)
    ang = rng.integers(0, 5000, N).astype(np.int32)
    cases = [("uniform, no border erase", rng.random(N).astype(np.float32), (8, 8, 8), 150, False),
             ("all negative", -rng.random(N).astype(np.float32) - 0.5, (8, 8, 8), 40, False),
             ("integer ties", rng.integers(0, 4, N).astype(np.float32), (6, 6, 6), 200, False),
             ("normal, borders erased", rng.standard_normal(N).astype(np.float32), (10, 10, 10), 300, True),
             ("spikes run flat", np.zeros(N, np.float32), (8, 8, 8), 50, True)]
    cases[-1][1][[10, 20, 15], [12, 20, 9], [8, 10, 11]] = [3.0, 1.0, 2.0]
    for name, a, t, n, erase in cases:
        print(f"  picklist: {name}, map {N}, template {t}, n {n}", flush=True)
        got = api.tom_os3_returnPicklist(a, t, n, angles=ang, erase_borders=erase)
        ref = oracle.return_picklist(a, ang, t, n, erase=None if erase else False)
        assert got.shape == ref.shape, name
        np.testing.assert_array_equal(got[:, :3], ref[:, :3], err_msg=name)
        np.testing.assert_array_equal(got[:, 3], ref[:, 3].astype(np.float32), err_msg=name)
        np.testing.assert_array_equal(got[:, 4], ref[:, 4], err_msg=name)
    assert len(api.tom_os3_returnPicklist(cases[-1][1], (8, 8, 8), 50)) == 3
    img = rng.random((40, 33)).astype(np.float32)
    print("  picklist: image 40x33", flush=True)
    got = api.tom_os3_returnPicklist(img, (8, 8, 1), 80, erase_borders=False)
    ref = oracle.return_picklist(img, None, (8, 8, 1), 80, erase=False)
    assert got.shape == ref.shape and np.array_equal(got[:, :4], ref[:, :4].astype(np.float32))
    flat = np.full(N, 0.25, np.float32)
    assert api.tom_os3_returnPicklist(flat, N, 10).shape == (0, 5)
    nan = rng.random(N).astype(np.float32)
    nan[18, 17, 10] = np.nan
    assert api.tom_os3_returnPicklist(nan, (8, 8, 8), 10).shape == (0, 5)
    # tom_peakc's 1D branch (size(in,1) == 1) is refused, not guessed
    from tom_amd._lib import TomHipError
    with pytest.raises(TomHipError):
        api.tom_os3_returnPicklist(rng.random((1, 30, 1)).astype(np.float32), (8, 8, 1), 5, erase_borders=False)


@pytest.mark.gpu
def test_picklist_large_map_matches_oracle(oracle):
    """A 160x150x140 map with 300 picks (bricks of many rewrites) against
    the oracle."""
    from tom_amd import api
    rng = np.random.default_rng(23)
    N = (160, 150, 140)
    a = (0.1 * rng.standard_normal(N)).astype(np.float32)
    x, y, z = [np.arange(s) for s in N]
    for _ in range(200):
        c = [rng.integers(0, s) for s in N]
        a[max(c[0] - 3, 0):c[0] + 4, max(c[1] - 3, 0):c[1] + 4, max(c[2] - 3, 0):c[2] + 4] += rng.uniform(0.2, 1.0)
    print(f"  picklist: map {N}, 300 picks", flush=True)
    got = api.tom_os3_returnPicklist(a, (24, 24, 24), 300)
    ref = oracle.return_picklist(a, None, (24, 24, 24), 300)
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got[:, :3], ref[:, :3])
    np.testing.assert_array_equal(got[:, 3], ref[:, 3].astype(np.float32))
