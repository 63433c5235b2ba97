"""Synthetic workload generators: PCG32 against its scalar definition,
mesh sizes, watertightness and ranges."""
import numpy as np

from bvhtracer_amd import scenes


def test_pcg32_vectorised_matches_scalar():
    for seed, stream in [(42, 54), (0x5EED0001, 7), (1, 0)]:
        rng = scenes.Pcg32(seed, np.array([stream], np.uint64))
        got = [int(rng.next_u32()[0]) for _ in range(6)]
        assert got == scenes.pcg32_scalar_sequence(seed, stream, 6)
    # published first outputs of pcg32 with initstate 42, initseq 54 (pcg32-demo)
    assert scenes.pcg32_scalar_sequence(42, 54, 3) == [0xa15c02b7, 0x7b47f409, 0xba1d3330]


def test_soup_ranges():
    t = scenes.triangle_soup()
    assert t.shape == (12582, 9) and t.dtype == np.float32
    v0 = t[:, 0:3]
    assert v0.min() >= -5 and v0.max() < 5
    for k in (3, 6):
        off = t[:, k:k + 3] - v0
        assert off.min() >= 0 and off.max() <= 1.0001
    assert np.array_equal(t, scenes.triangle_soup())  # deterministic


def test_icosphere_counts_and_watertight():
    for n in (1, 3, 8):
        t = scenes.icosphere(n, radius=10.0)
        assert t.shape == (20 * n * n, 9)
        e = t.reshape(-1, 3, 3)
        edges = {}
        for a, b in ((0, 1), (1, 2), (2, 0)):
            for p, q in zip(e[:, a], e[:, b]):
                key = tuple(sorted((p.tobytes(), q.tobytes())))
                edges[key] = edges.get(key, 0) + 1
        assert set(edges.values()) == {2}  # every edge shared by exactly two faces
        r = np.linalg.norm(e.reshape(-1, 3).astype(np.float64), axis=1)
        assert r.min() > 10 * 0.84 and r.max() < 10 * 1.16
    assert scenes.icosphere_frequency(1310720) == 256 and scenes.icosphere_frequency(200000) == 100


def test_random_affines_are_inverse_pairs():
    M, Minv = scenes.random_affines(16, seed=9)
    for k in range(16):
        A = M[k].reshape(4, 4).T.astype(np.float64)
        B = Minv[k].reshape(4, 4).T.astype(np.float64)
        assert np.allclose(A @ B, np.eye(4), atol=1e-5)
        s = np.linalg.norm(A[:3, 0])
        assert 0.5 <= s <= 1.5 and np.abs(A[:3, 3]).max() <= 20
