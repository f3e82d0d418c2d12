"""Pins the oracle's clipper against the reference's own Clipper.cpp.

oracle/_ref/libswr_ref.so is src/Device/Clipper.cpp compiled unmodified;
the oracle's restatement must give bit-identical polygons.  Where the
reference build is absent, the test falls back to what the mathematics
fixes: every output vertex lies on the inside of each plane it was clipped
against (up to rounding), and fully inside triangles come back unchanged.
"""
import ctypes

import numpy as np
import pytest

from oracle_runner import oracle, reference_lib

CLIP_RIGHT, CLIP_TOP, CLIP_FAR, CLIP_LEFT, CLIP_BOTTOM, CLIP_NEAR, CLIP_FINITE = 1, 2, 4, 8, 16, 32, 128


def clip_flags(v):
    x, y, z, w = [np.float32(c) for c in v]
    f = 0
    f |= CLIP_RIGHT if w < x else 0
    f |= CLIP_TOP if w < y else 0
    f |= CLIP_LEFT if not (-w <= x) else 0
    f |= CLIP_BOTTOM if not (-w <= y) else 0
    f |= CLIP_FAR if w < z else 0
    f |= CLIP_NEAR if not (0 <= z) else 0
    return f | CLIP_FINITE


def run(fn, tri, flags):
    fp = ctypes.POINTER(ctypes.c_float)
    vs = [np.ascontiguousarray(v, np.float32) for v in tri]
    out = np.zeros(64, np.float32)
    n = ctypes.c_int(0)
    r = fn(*[v.ctypes.data_as(fp) for v in vs], flags, out.ctypes.data_as(fp), ctypes.byref(n))
    return r, out[:4 * n.value].reshape(-1, 4).copy()


def random_triangles(seed, count):
    rng = np.random.default_rng(seed)
    for _ in range(count):
        w = rng.uniform(-1.0, 3.0, 3)
        xyz = rng.uniform(-3.0, 3.0, (3, 3)) * np.abs(w)[:, None]
        yield [np.array([*xyz[i], w[i]], np.float32) for i in range(3)]


def test_clipper_bit_exact_vs_compiled_reference():
    ref = reference_lib()
    if ref is None:
        pytest.skip("oracle/_ref/libswr_ref.so not built (needs the reference sources)")
    L = oracle()
    checked = 0
    for tri in random_triangles(1, 3000):
        flags = clip_flags(tri[0]) | clip_flags(tri[1]) | clip_flags(tri[2])
        if flags == CLIP_FINITE:
            continue
        r1, p1 = run(L.swo_clip_triangle, tri, flags)
        r2, p2 = run(ref.ref_clip_triangle, tri, flags)
        assert r1 == r2
        assert p1.shape == p2.shape
        assert np.array_equal(p1.view(np.uint32), p2.view(np.uint32))
        checked += 1
    assert checked > 1000


def test_clipper_invariants():
    L = oracle()
    for tri in random_triangles(2, 1500):
        flags = clip_flags(tri[0]) | clip_flags(tri[1]) | clip_flags(tri[2])
        r, poly = run(L.swo_clip_triangle, tri, flags)
        if flags == CLIP_FINITE:
            assert r == 1 and np.array_equal(poly, np.stack(tri))
            continue
        if not r:
            continue
        assert 3 <= len(poly) <= 9
        x, y, z, w = poly.T.astype(np.float64)
        scale = 1e-5 * (1.0 + np.abs(poly).max())
        if flags & CLIP_NEAR:
            assert (z >= -scale).all()
        if flags & CLIP_FAR:
            assert (w - z >= -scale).all()
        if flags & CLIP_LEFT:
            assert (w + x >= -scale).all()
        if flags & CLIP_RIGHT:
            assert (w - x >= -scale).all()
        if flags & CLIP_BOTTOM:
            assert (w + y >= -scale).all()
        if flags & CLIP_TOP:
            assert (w - y >= -scale).all()


def test_clipper_bit_exact_vs_golden_vectors():
    """The committed vectors of the compiled reference (tests/golden/make_ref_vectors.py)."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "clipper_ref.npz")
    z = np.load(path)
    L = oracle()
    for tri, f, vis, n, poly in zip(z["triangles"], z["flags"], z["visible"], z["count"], z["polygon"]):
        r, p = run(L.swo_clip_triangle, list(tri), int(f))
        assert r == vis
        assert len(p) == n
        assert np.array_equal(p.view(np.uint32), poly[:n].view(np.uint32))
