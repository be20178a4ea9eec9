"""CPU tests of the product library's curved-element Nystrom host code (no
GPU needed): the curved icosphere mesher bit-exact against the oracle's
restatement of the same rule, the quadrature nodes against the oracle's
independent numpy evaluation, and the curved topology contract of the
reference (src/geometry/surface.jl:271-298,341-419) with its messages.
"""
import numpy as np
import pytest

import bemhip as bh
from oracle import nystrom as NY


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("level", [0, 1, 2, 3])
def test_curved_icosphere_bit_exact(level, order):
    n1, t1 = bh.curved_icosphere_arrays(level, 0.02, 0.01, 0.01, order=order)
    n2, t2 = NY.curved_icosphere(level, (0.02, 0.01, 0.01), order)
    assert np.array_equal(t1 - 1, t2)
    assert np.array_equal(n1, n2)
    assert t1.shape == ({1: 3, 2: 6, 3: 10}[order], 20 * 4 ** level)


@pytest.mark.parametrize("order,qorder", [(1, 4), (2, 1), (2, 2), (2, 3), (2, 4), (2, 5), (2, 6), (2, 8),
                                          (3, 4), (3, 7)])
def test_quadrature_matches_oracle(order, qorder):
    nodes, tri = bh.curved_icosphere_arrays(1, 0.02, 0.01, 0.015, order=order)
    xyz, nr, w = bh.nystrom_quadrature(nodes, tri, qorder)
    Q = NY.Quadrature(nodes, tri - 1, qorder)
    assert xyz.shape == Q.coords.shape
    assert np.allclose(xyz, Q.coords, rtol=0, atol=1e-16)
    assert np.allclose(nr, Q.normals, rtol=0, atol=1e-14)
    assert np.allclose(w, Q.weights, rtol=1e-13, atol=0)


def test_quadrature_of_supplied_flat_mesh():
    nodes, tri = bh.icosphere_arrays(1, 1.0)
    xyz, nr, w = bh.nystrom_quadrature(nodes, tri, 4)
    assert len(w) == 6 * tri.shape[1]
    Q = NY.Quadrature(nodes, tri - 1, 4)
    assert np.allclose(w, Q.weights, rtol=1e-13)


def test_curved_checks_and_messages():
    nodes, tri = bh.curved_icosphere_arrays(0, 1.0, order=2)
    bad = tri.copy()
    bad[3, 0] = bad[0, 0]
    with pytest.raises(ValueError, match="triangle 1 repeats a node"):
        bh.nystrom_quadrature(nodes, bad, 4)
    # an edge node that the neighbour does not share
    extra = np.hstack([nodes, nodes[:, [int(tri[3, 0]) - 1]]])
    bad = tri.copy()
    bad[3, 0] = extra.shape[1]
    with pytest.raises(ValueError, match=r"nonconforming curved nodes at edge \("):
        bh.nystrom_quadrature(extra, bad, 4)
    with pytest.raises(ValueError, match="triangles must have 3, 6 or 10 rows"):
        bh.nystrom_quadrature(nodes, tri[:5], 4)
    with pytest.raises(ValueError, match="qorder must be between 1 and 8"):
        bh.nystrom_quadrature(nodes, tri, 9)
    bad = tri.copy()
    bad[5, 2] = nodes.shape[1] + 1
    with pytest.raises(ValueError, match="invalid node index"):
        bh.nystrom_quadrature(nodes, bad, 4)
    # the flat checks still apply on the corners: flip one triangle
    bad = tri.copy()
    bad[[1, 2], 0] = bad[[2, 1], 0]
    bad[[3, 5], 0] = bad[[5, 3], 0]
    with pytest.raises(ValueError, match="inconsistent orientation"):
        bh.nystrom_quadrature(nodes, bad, 4)


def test_curved_icosphere_sizes():
    import ctypes
    L = bh.load_library()
    nv, nt = ctypes.c_int64(), ctypes.c_int64()
    # level 2: 162 vertices, 480 edges, 320 faces
    for order, count in ((1, 162), (2, 162 + 480), (3, 162 + 2 * 480 + 320)):
        assert L.bem_curved_icosphere_size(2, order, ctypes.byref(nv), ctypes.byref(nt)) == 0
        assert (nv.value, nt.value) == (count, 320)
    assert L.bem_curved_icosphere_size(2, 4, ctypes.byref(nv), ctypes.byref(nt)) != 0
    for order in (1, 2, 3):
        n, t = bh.curved_icosphere_arrays(2, 1.0, order=order)
        assert L.bem_curved_icosphere_size(2, order, ctypes.byref(nv), ctypes.byref(nt)) == 0
        assert (nv.value, nt.value) == (n.shape[1], t.shape[1])
