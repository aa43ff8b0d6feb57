"""Pins of the oracle's primitives against closed forms, independent
restatements of a different form, and statistical properties."""
import numpy as np
import pytest

from oracle import oracle as orc


def _interleave(x, y, z):
    """Bit-by-bit Morton interleave (x highest of each triple)."""
    m = 0
    for b in range(10):
        m |= ((x >> b) & 1) << (3 * b + 2)
        m |= ((y >> b) & 1) << (3 * b + 1)
        m |= ((z >> b) & 1) << (3 * b)
    return m


def test_morton_matches_bitwise_interleave():
    rng = np.random.default_rng(0)
    for x, y, z in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1023, 1023, 1023), (512, 3, 700)]:
        assert orc.morton3(x, y, z) == _interleave(x, y, z)
    for x, y, z in rng.integers(0, 1024, (2000, 3)):
        assert orc.morton3(int(x), int(y), int(z)) == _interleave(int(x), int(y), int(z))
    assert orc.morton3(1023, 1023, 1023) == (1 << 30) - 1


def test_pcg_is_a_permutation_on_a_range_and_mixes():
    # pcg (LCG step + RXS-M-XS output) is a bijection on 32-bit words: no collisions
    vals = np.array([orc.pcg(i) for i in range(1 << 16)], np.uint64)
    assert len(np.unique(vals)) == 1 << 16
    # avalanche: flipping one input bit flips ~half of the output bits
    rng = np.random.default_rng(1)
    flips = []
    for v in rng.integers(0, 1 << 32, 500, dtype=np.uint64):
        for b in range(0, 32, 5):
            d = orc.pcg(int(v)) ^ orc.pcg(int(v) ^ (1 << b))
            flips.append(bin(d).count("1"))
    assert 14.0 < np.mean(flips) < 18.0


def test_pcg4d_samples_uniform_and_decorrelated():
    n = 256
    us, vs = [], []
    for y in range(n):
        for x in range(0, n, 4):
            u, v = orc.sample2(x, y, 0, 3)
            us.append(u)
            vs.append(v)
    us, vs = np.array(us), np.array(vs)
    assert us.min() >= 0.0 and us.max() < 1.0 and vs.min() >= 0.0 and vs.max() < 1.0
    # values lie on the 2^-24 grid
    assert np.all(np.floor(us * 2**24) == us * 2**24)
    # chi-square over 16 bins
    for a in (us, vs):
        h, _ = np.histogram(a, bins=16, range=(0, 1))
        e = len(a) / 16
        assert ((h - e) ** 2 / e).sum() < 45.0  # 15 dof, p ~ 1e-4
    assert abs(np.corrcoef(us, vs)[0, 1]) < 0.05
    # different slots and frames give different samples
    assert not np.array_equal(orc.sample2(5, 7, 0, 0), orc.sample2(5, 7, 0, 1))
    assert not np.array_equal(orc.sample2(5, 7, 0, 0), orc.sample2(5, 7, 1, 0))


def test_pcg4d_avalanche():
    rng = np.random.default_rng(2)
    flips = []
    for v in rng.integers(0, 1 << 32, (200, 4), dtype=np.uint64):
        base = orc.pcg4d(v.astype(np.uint32))
        for lane in range(4):
            w = v.copy()
            w[lane] ^= 1 << int(rng.integers(0, 32))
            d = orc.pcg4d(w.astype(np.uint32))
            flips.append(sum(bin(int(a) ^ int(b)).count("1") for a, b in zip(base[:2], d[:2])))
    assert 28.0 < np.mean(flips) < 36.0


def test_sincos_turns_against_numpy():
    w = np.concatenate([np.linspace(0, 1, 20001, endpoint=False, dtype=np.float32),
                        np.random.default_rng(3).random(5000, dtype=np.float32)])
    s = np.empty_like(w)
    c = np.empty_like(w)
    for i, x in enumerate(w):
        s[i], c[i] = orc.sincos_turns(float(x))
    ang = 2 * np.pi * w.astype(np.float64)
    assert np.abs(s - np.sin(ang)).max() < 4e-7
    assert np.abs(c - np.cos(ang)).max() < 4e-7
    assert orc.sincos_turns(0.0) == (0.0, 1.0)
    assert orc.sincos_turns(0.25) == (1.0, 0.0) or abs(orc.sincos_turns(0.25)[1]) < 1e-7


def test_triangle_closed_forms():
    A, B, Cc = (0, 0, 0), (1, 0, 0), (0, 1, 0)
    hit, t, bary = orc.intersect_tri((0.25, 0.25, 1), (0, 0, -1), 0.0, 10.0, A, B, Cc)
    assert hit and t == 1.0
    np.testing.assert_allclose(bary, [0.5, 0.25, 0.25], atol=1e-7)
    # scaled direction: t is parametric in direction units
    hit, t, _ = orc.intersect_tri((0.25, 0.25, 1), (0, 0, -2), 0.0, 10.0, A, B, Cc)
    assert hit and t == 0.5
    # outside the triangle
    assert not orc.intersect_tri((0.75, 0.75, 1), (0, 0, -1), 0.0, 10.0, A, B, Cc)[0]
    # closed interval [tmin, tmax]
    assert orc.intersect_tri((0.25, 0.25, 1), (0, 0, -1), 0.0, 1.0, A, B, Cc)[0]
    assert not orc.intersect_tri((0.25, 0.25, 1), (0, 0, -1), 0.0, 0.999, A, B, Cc)[0]
    assert orc.intersect_tri((0.25, 0.25, 1), (0, 0, -1), 1.0, 2.0, A, B, Cc)[0]
    assert not orc.intersect_tri((0.25, 0.25, 1), (0, 0, -1), 1.001, 2.0, A, B, Cc)[0]
    # both faces (no culling), and behind the origin is a miss
    assert orc.intersect_tri((0.25, 0.25, -1), (0, 0, 1), 0.0, 10.0, A, B, Cc)[0]
    assert not orc.intersect_tri((0.25, 0.25, 1), (0, 0, 1), 0.0, 10.0, A, B, Cc)[0]
    # vertex and edge hits count
    assert orc.intersect_tri((0, 0, 1), (0, 0, -1), 0.0, 10.0, A, B, Cc)[0]
    assert orc.intersect_tri((0.5, 0, 1), (0, 0, -1), 0.0, 10.0, A, B, Cc)[0]


def _cube():
    v = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float32)
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for a, b, c, d in faces:
        tris += [(a, b, c), (a, c, d)]
    return v, tris


def test_watertight_no_leaks_through_edges_and_vertices():
    v, tris = _cube()
    rng = np.random.default_rng(4)
    targets = list(v) + [(v[a] + v[b]) / 2 for t in tris for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0]))]
    targets += list(rng.uniform(-1, 1, (300, 3)) * np.array([1, 1, 0]) + np.array([0, 0, 1]))
    origins = [np.zeros(3), np.array([0.1, -0.2, 0.3]), np.array([0.3, 0.3, -0.3])]
    for o in origins:
        for tgt in targets:
            d = np.asarray(tgt, np.float32) - o.astype(np.float32)
            hits = sum(orc.intersect_tri(o, d, 0.0, 10.0, v[a], v[b], v[c])[0] for a, b, c in tris)
            assert hits >= 1, (o, tgt)


def test_box_test_is_conservative_for_triangle_hits():
    rng = np.random.default_rng(5)
    n_hit = 0
    for _ in range(3000):
        tri = rng.uniform(-3, 3, (3, 3)).astype(np.float32)
        o = rng.uniform(-6, 6, 3).astype(np.float32)
        tgt = (tri * rng.dirichlet([1, 1, 1])[:, None]).sum(0)
        d = (tgt - o).astype(np.float32) * np.float32(rng.uniform(0.5, 2.0))
        hit, t, _ = orc.intersect_tri(o, d, 0.0, 100.0, *tri)
        if not hit:
            continue
        n_hit += 1
        m = np.abs(tri).max() * np.float32(2.0**-16)
        lo, hi = tri.min(0) - m, tri.max(0) + m
        assert orc.box_test(o, d, 0.0, 100.0, lo, hi)
        # the box test also honours the interval end at the hit distance
        assert orc.box_test(o, d, 0.0, t, lo, hi)
    assert n_hit > 2500
    # misses: a box entirely behind the ray, and one beyond tmax
    assert not orc.box_test((0, 0, 0), (1, 0, 0), 0.0, 100.0, (-5, -1, -1), (-4, 1, 1))
    assert not orc.box_test((0, 0, 0), (1, 0, 0), 0.0, 3.0, (4, -1, -1), (5, 1, 1))
    # axis-parallel ray inside a slab (d = 0 components)
    assert orc.box_test((0, 0, 0), (1, 0, 0), 0.0, 10.0, (2, -1, -1), (3, 1, 1))
    assert orc.box_test((0, 1, 0), (1, 0, 0), 0.0, 10.0, (2, 1, -1), (3, 2, 1))  # on the face
    assert not orc.box_test((0, 2.5, 0), (1, 0, 0), 0.0, 10.0, (2, -1, -1), (3, 2, 1))


def test_invert_3x4_matches_numpy():
    rng = np.random.default_rng(6)
    for _ in range(50):
        m = rng.normal(size=(3, 4)).astype(np.float32)
        inv = orc.invert_3x4(m).reshape(3, 4).astype(np.float64)
        full = np.eye(4)
        full[:3] = m
        ref = np.linalg.inv(full)[:3]
        np.testing.assert_allclose(inv, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())
