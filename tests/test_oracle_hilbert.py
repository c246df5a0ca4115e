"""Oracle pin: 63-bit Hilbert keys (Cornerstone hilbert3D, restated from its published algorithm).

The Cornerstone bit pattern itself cannot be compared (its source is not available), so the
keys are pinned by what any valid 3-D Hilbert curve of this construction must satisfy:
  - bijection: the independent decoder inverts the encoder;
  - locality: consecutive keys are face-adjacent cells (Manhattan distance 1);
  - nesting: every aligned level-l cube maps to one contiguous key range of length 8^(21-l);
  - the level-1 octant order follows mortonToHilbert {0,1,3,2,7,6,4,5} from the origin;
and the quantisation / dead-particle rules of include/tree.cuh:89-105.
"""
import numpy as np

from oracle import oracle as O

M = (1 << 21) - 1


def test_bijection_random():
    rng = np.random.default_rng(1)
    p = rng.integers(0, 1 << 21, size=(200000, 3), dtype=np.uint32)
    k = O.hilbert_encode(p[:, 0], p[:, 1], p[:, 2])
    assert np.all(k < (1 << 63))
    np.testing.assert_array_equal(O.hilbert_decode(k), p)


def test_consecutive_keys_face_adjacent():
    rng = np.random.default_rng(2)
    k = rng.integers(0, (1 << 63) - 1, size=100000, dtype=np.uint64)
    a = O.hilbert_decode(k).astype(np.int64)
    b = O.hilbert_decode(k + np.uint64(1)).astype(np.int64)
    np.testing.assert_array_equal(np.abs(a - b).sum(axis=1), 1)


def test_aligned_cubes_are_contiguous_key_ranges():
    rng = np.random.default_rng(3)
    for level in (1, 3, 7, 12, 20):
        side = 1 << (21 - level)
        for _ in range(20):
            corner = rng.integers(0, 1 << level, size=3) * side
            pts = corner + rng.integers(0, side, size=(500, 3))
            k = O.hilbert_encode(pts[:, 0], pts[:, 1], pts[:, 2])
            pref = k >> np.uint64(3 * (21 - level))
            assert np.all(pref == pref[0])


def test_level1_octant_order():
    h = 1 << 20
    digits = []
    for o in range(8):
        x, y, z = (o >> 2) & 1, (o >> 1) & 1, o & 1
        k = O.hilbert_encode(np.array([x * h]), np.array([y * h]), np.array([z * h]))
        digits.append(int(k[0] >> np.uint64(60)))
    assert digits == [0, 1, 3, 2, 7, 6, 4, 5]
    assert int(O.hilbert_encode(np.array([0]), np.array([0]), np.array([0]))[0]) == 0


def test_quantisation_and_dead_keys():
    box = np.array([0.0, 1.0, -1.0, 1.0, 2.0, 3.0])
    x = np.array([0.0, 1.0, 0.5, 0.25])
    y = np.array([-1.0, 1.0, 0.0, 0.0])
    z = np.array([2.0, 3.0, 2.5, 2.5])
    h = np.array([0.1, 0.1, 0.1, -0.1])
    k = O.keys(x, y, z, h, box)
    assert k[3] == np.uint64(0xFFFFFFFFFFFFFFFF)
    c = O.hilbert_decode(k[:3])
    np.testing.assert_array_equal(c[0], [0, 0, 0])          # box minimum -> cell 0
    np.testing.assert_array_equal(c[1], [M, M, M])          # box maximum -> clamped
    np.testing.assert_array_equal(c[2], [1 << 20, 1 << 20, 1 << 20])
    # open box: x1.00001 padding of the largest span, centred (src/tree.cu:56-71)
    b = O.open_box(np.array([0.0, 2.0]), np.array([0.0, 1.0]), np.array([5.0, 5.5]))
    span = 2.0 * 1.00001
    np.testing.assert_allclose(b, [1 - span / 2, 1 + span / 2, 0.5 - span / 2, 0.5 + span / 2,
                                   5.25 - span / 2, 5.25 + span / 2], rtol=0, atol=1e-15)
