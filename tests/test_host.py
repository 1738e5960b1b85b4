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
