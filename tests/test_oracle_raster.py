"""The oracle's rasterizer against the reference's own rasterizer.cpp.

Fixtures in tests/golden/raster_golden.npz were produced by the reference code
(compiled from src/puresoft3d/rasterizer.cpp into oracle/_ref); when that
library is present the test also cross-checks fresh random triangles live.
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(HERE, "golden", "raster_golden.npz"))


def test_oracle_matches_reference_fixtures(oracle, golden):
    g = golden
    L = oracle.lib()
    n = len(g["ok"])
    assert n >= 2000
    for i in range(n):
        W, H = (int(x) for x in g["sizes"][i])
        ok, f, l, rows, vv = oracle.push_triangle(L.or_push_triangle, W, H, *g["verts"][i])
        assert (ok, f, l) == (int(g["ok"][i]), int(g["first"][i]), int(g["last"][i])), i
        np.testing.assert_array_equal(vv.view(np.uint32), g["verts_out"][i].view(np.uint32))
        blk = g["rows"][g["row_off"][i]: g["row_off"][i + 1]]
        if len(blk):
            np.testing.assert_array_equal(rows[f: l + 1], blk)
            untouched = np.ones(H, bool)
            untouched[f: l + 1] = False
            assert (rows[untouched, 0] == 0x7FFFFFFF).all()
        else:
            assert (rows[:, 0] == 0x7FFFFFFF).all()


def test_oracle_matches_reference_live(oracle):
    R = oracle.ref()
    if R is None:
        pytest.skip("oracle/_ref/libref_raster.so not built (reference tree absent)")
    L = oracle.lib()
    rng = np.random.default_rng(99)
    for k in range(3000):
        W, H = [(640, 480), (97, 61), (1920, 1080)][k % 3]
        v = np.zeros((3, 4), np.float32)
        v[:, :2] = rng.uniform(-1.2, 1.2, (3, 2)) * (1.0 if k % 5 else 40.0)
        if k % 7 == 0:
            v[1, 1] = v[0, 1]
        v[:, 3] = 1
        a = oracle.push_triangle(L.or_push_triangle, W, H, *v)
        b = oracle.push_triangle(R.ref_push_triangle, W, H, *v)
        assert a[:3] == b[:3]
        np.testing.assert_array_equal(a[3], b[3])


def test_span_invariants(oracle):
    """leftClamped >= 0, rightClamped <= W-1, rows written == [first, last],
    zero-height triangles rejected (rasterizer.cpp:100-111, 165-168)"""
    L = oracle.lib()
    rng = np.random.default_rng(3)
    W, H = 200, 150
    for _ in range(2000):
        v = np.zeros((3, 4), np.float32)
        v[:, :2] = rng.uniform(-1.5, 1.5, (3, 2))
        v[:, 3] = 1
        ok, f, l, rows, vv = oracle.push_triangle(L.or_push_triangle, W, H, *v)
        ys = vv[1::2]
        if ok:
            assert f != l
            if f <= l:
                r = rows[f: l + 1]
                assert (r[:, 1] >= 0).all() and (r[:, 5] <= W - 1).all()
                assert (r[:, 1] == np.maximum(r[:, 0], 0)).all()
                assert (r[:, 5] == np.where(r[:, 4] >= W, W - 1, r[:, 4])).all()
        else:
            assert f == l      # pushTriangle returns false only for zero height


def test_shared_row_belongs_to_lower_half(oracle):
    """processStandingTriangle writes the upper half first and the lower half
    overwrites the third vertex's row (rasterizer.cpp:126-139)"""
    L = oracle.lib()
    W, H = 64, 64
    # third vertex exactly on row 32 (screen y = 32.0), pointing left
    v = np.array([[0.5, 0.9, 0, 1], [0.6, -0.9, 0, 1], [-0.7, 0.0, 0, 1]], np.float32)
    ok, f, l, rows, vv = oracle.push_triangle(L.or_push_triangle, W, H, *v)
    assert ok and vv[5] == 32.0
    # lower half's left edge is (third, bottom) = (2, 1); upper half's is (top, third) = (0, 2)
    assert tuple(rows[32, 2:4]) == (2, 1)
    assert tuple(rows[33, 2:4]) == (0, 2)
    assert tuple(rows[31, 2:4]) == (2, 1)
