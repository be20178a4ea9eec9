"""Cluster tree and block tree of the hierarchical path: the library's plan
equals the CPU restatement index for index, and the block structure is a
valid partition with the stated admissibility."""
import math

import numpy as np
import pytest

import bemhip as bh
import hmatrix_model as hm


def _centroids(nodes, tri):
    return nodes[:, tri - 1].mean(axis=1)  # 3 x n


@pytest.mark.parametrize("case", ["random", "sphere", "bent_cylinder"])
def test_plan_matches_restatement(case):
    if case == "random":
        pts = np.random.default_rng(4).standard_normal((3, 3000))
        near = 0.05
    elif case == "sphere":
        pts = _centroids(*bh.icosphere_arrays(4, 0.01))
        near = 2 * 0.001
    else:
        pts = _centroids(*bh.cylinder_arrays(bh.Cylinder(0.5, 2.0, 2.0, 0.5), 64))
        near = 0.2
    perm, nn, blocks = bh.hmatrix_plan(pts, 64, 2.0, near)
    p2, nn2, b2 = hm.plan(pts.T, 64, 2.0, near)
    assert np.array_equal(perm, p2)
    assert nn == nn2
    assert np.array_equal(blocks, b2)


def test_blocks_partition_the_matrix_and_respect_admissibility():
    nodes, tri = bh.icosphere_arrays(3, 1.0)
    pts = _centroids(nodes, tri)
    n = pts.shape[1]
    perm, _, blocks = bh.hmatrix_plan(pts, 32, 2.0, 0.3)
    assert sorted(perm) == list(range(n))
    cover = np.zeros((n, n), dtype=np.int32)
    P = pts[:, perm]
    for lr, r0, m, c0, k in blocks:
        cover[r0:r0 + m, c0:c0 + k] += 1
        if lr:
            a, b = P[:, r0:r0 + m], P[:, c0:c0 + k]
            gap = np.maximum(0, np.maximum(b.min(1) - a.max(1), a.min(1) - b.max(1)))
            dist = math.sqrt((gap ** 2).sum())
            da = np.linalg.norm(a.max(1) - a.min(1))
            db = np.linalg.norm(b.max(1) - b.min(1))
            assert dist > 0.3 and min(da, db) < 2.0 * dist
        else:
            assert m <= 32 and k <= 32
    assert (cover == 1).all()
    assert (blocks[:, 0] == 1).sum() > 0


def test_bent_cylinder_is_closed_with_the_right_volume():
    body = bh.Cylinder(0.5, 2.0, 2.0, 0.5)
    nodes, tri = bh.cylinder_arrays(body, 96)
    bh.validate_surface(nodes, tri)
    p = nodes[:, tri - 1]
    vol = (np.einsum("ij,ij->j", p[:, 0], np.cross(p[:, 1].T, p[:, 2].T).T) / 6).sum()
    exact = math.pi * 0.5 ** 2 * 2.0 + 4 / 3 * math.pi * 0.5 ** 2 * 0.5  # Pappus tube + two half-ellipsoid caps
    assert vol == pytest.approx(exact, rel=2e-3)
    # centerline ends at angle +-L/(2R) on a circle of radius R in the xy plane
    straight = bh.cylinder_arrays(bh.Cylinder(0.5, 2.0, math.inf, 0.5), 96)[0]
    assert straight[0].max() == pytest.approx(1.5, rel=1e-9) and straight[0].min() == pytest.approx(-1.5, rel=1e-9)
