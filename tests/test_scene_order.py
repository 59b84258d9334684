"""Host restatement of Main.cpp:184-190's front-to-back order
(scenes.front_to_back), pinned by known answers: ascending squared distance,
ties in index order, and the float32 summation order of DPPS."""

from __future__ import annotations

import numpy as np

from rasterizer_amd import scenes


class _C:
    def __init__(self, c):
        self.m_center = np.array([*c, 1.0], np.float32)


def test_order_is_ascending_distance():
    rng = np.random.default_rng(0)
    dist = rng.permutation(50).astype(np.float32) + 1.0
    dirs = rng.normal(size=(50, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    occ = [_C(np.float32(d) * u) for d, u in zip(dist, dirs)]
    order = scenes.front_to_back(occ, np.zeros(3, np.float32))
    assert np.array_equal(order, np.argsort(dist))


def test_ties_keep_index_order():
    occ = [_C((3, 0, 0)), _C((0, 3, 0)), _C((1, 0, 0)), _C((0, 0, -3)), _C((0, 1, 0))]
    assert scenes.front_to_back(occ, np.zeros(3, np.float32)).tolist() == [2, 4, 0, 1, 3]


def test_dpps_summation_order():
    """(dx^2 + dy^2) + dz^2: for d = (1, 2^-12, 2^-12) that is exactly 1, a tie
    with d = (1, 0, 0) that index order breaks; dx^2 + (dy^2 + dz^2) would
    give 1 + 2^-23 and the reverse order."""
    t = 2.0 ** -12
    occ = [_C((1, t, t)), _C((1, 0, 0))]
    assert scenes.front_to_back(occ, np.zeros(3, np.float32)).tolist() == [0, 1]
    x2, y2 = np.float32(1), np.float32(t * t)
    assert (x2 + y2) + y2 == np.float32(1) and x2 + (y2 + y2) > np.float32(1)
