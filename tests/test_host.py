"""Host-side logic of the product that runs without a GPU."""
import numpy as np


def test_angle_list_raw_product_order():
    from tom_amd import api
    a = api.tom_os3_angleList([0, 0, 30], [330, 330, 150], [30, 30, 30])
    # theta never 0 or 180: no two triples share a rotation -> raw product
    assert a.shape == (12 * 12 * 5, 3)
    assert list(a[0]) == [0, 0, 30] and list(a[1]) == [0, 0, 60] and list(a[5]) == [0, 30, 30]


def test_angle_list_dedup_degenerate_theta():
    """Dedup only ever drops degenerate rows (theta = 0 or 180, where the
    rotation depends on phi+psi or psi-phi alone), keeps the first member of
    each key group, and keeps every other row. How many degenerate rows
    survive depends on last-ulp rounding of floor(1e4*R) at integer
    boundaries (tom_os3_angleList.m:76-88): parity of that count is unpinned
    without MATLAB, so only the bounds are checked."""
    from tom_amd import api
    a = api.tom_os3_angleList([0, 0, 0], [330, 330, 180], [30, 30, 30])
    raw = api.tom_os3_angleList([0, 0, 0], [330, 330, 180], [30, 30, 30], dedup=False)
    assert raw.shape == (1008, 3)
    for th in (30, 60, 90, 120, 150):
        assert (a[:, 2] == th).sum() == 144
    for th in (0, 180):
        k = int((a[:, 2] == th).sum())
        assert 12 <= k <= 144
    # order is preserved: the kept rows are a subsequence of the raw list
    rows = [tuple(r) for r in raw]
    pos = [rows.index(tuple(r)) for r in a]
    assert pos == sorted(pos)
    # theta = 0: the survivors cover every phi+psi class at least once
    t0 = a[a[:, 2] == 0]
    assert set(((t0[:, 0] + t0[:, 1]) % 360).tolist()) == set(range(0, 360, 30))


def test_angle_list_wrap_range():
    """start >= end enumerates start+1 : inc : end+359 (tom_os3_angleList.m:27-37)."""
    from tom_amd import api
    vals = api._angle_list_1d(350, 20, 10)
    assert vals[0] == 351 % 360 and vals[-1] == (351 + 10 * (len(vals) - 1)) % 360
    assert len(vals) == len(range(351, 20 + 360, 10))


def test_bench_one_gpu_only():
    """bench.py measures one GPU: any other --gpus exits with status 2 and a
    message, before touching a device."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=root)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_bench_byte_model_matches_the_unfused_reference_model():
    """The per-launch algorithmic bytes of the three FLCF passes at C2 and the
    unfused per-orientation model (64 N_c + 16 N) the roofline also reports."""
    import importlib.util
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench", os.path.join(root, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    m = bench.byte_model(bench.CONFIGS["c2"], 16)
    q = 256 * 512 * 256 * 8
    assert m["passB"] == 2 * 16 * q == 8589934592
    assert m["passA"] == 16 * 48 * 512 * 257 * 8 + 256 * 512 * 257 * 8 + 16 * q
    assert m["passC"] == 16 * q + 512 * 512 * 256 * 4
    X, Y, Z = bench.CONFIGS["c2"]["N"]
    assert 64 * X * Y * (Z // 2 + 1) + 16 * X * Y * Z == 3238002688
