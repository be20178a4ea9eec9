"""Reference mesh generators for the oracle (TEST INFRASTRUCTURE ONLY).

These restate, in numpy, the deterministic mesh rules the product library's
host code implements, so tests can check the product's connectivity bit-exactly.

Icosphere rule (shared with the product): the 12 icosahedron vertices
(+-1, +-phi, 0) cyclic, normalised, 20 faces in a fixed order; each level
splits every face (in face order) at edge midpoints, a midpoint vertex is
created at first encounter of its edge, `m = 0.5*(a+b)` is projected to the
unit sphere, and each face (a,b,c) becomes (a,ab,ca),(b,bc,ab),(c,ca,bc),(ab,bc,ca).

Conventions followed: nodes are 3 x Nv (x,y,z per column) and triangles
3 x N with outward orientation, as the reference's supplied-mesh contract
(src/geometry/surface.jl:271-298); the spheroid symmetry axis is x
(src/geometry/canonical_mesh.jl:24-27, src/geometry/spheroid.jl:7-28).
Triangles here are 0-based; the public API takes 1-based like the reference.
"""
import math

import numpy as np

_PHI = (1.0 + math.sqrt(5.0)) / 2.0

_ICO_V = [(-1, _PHI, 0), (1, _PHI, 0), (-1, -_PHI, 0), (1, -_PHI, 0),
          (0, -1, _PHI), (0, 1, _PHI), (0, -1, -_PHI), (0, 1, -_PHI),
          (_PHI, 0, -1), (_PHI, 0, 1), (-_PHI, 0, -1), (-_PHI, 0, 1)]
_ICO_F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
          (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
          (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
          (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]


def _normalise(p):
    x, y, z = p
    n = math.sqrt(x * x + y * y + z * z)
    return (x / n, y / n, z / n)


def icosphere(level):
    """Unit icosphere: (nodes 3 x Nv float64, triangles 3 x N int64, 0-based)."""
    verts = [_normalise(v) for v in _ICO_V]
    faces = list(_ICO_F)
    for _ in range(level):
        cache = {}
        new_faces = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            idx = cache.get(key)
            if idx is None:
                pa, pb = verts[a], verts[b]
                m = (0.5 * (pa[0] + pb[0]), 0.5 * (pa[1] + pb[1]), 0.5 * (pa[2] + pb[2]))
                verts.append(_normalise(m))
                idx = len(verts) - 1
                cache[key] = idx
            return idx

        for (a, b, c) in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new_faces.extend([(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)])
        faces = new_faces
    nodes = np.asarray(verts, dtype=np.float64).T.copy()
    tri = np.asarray(faces, dtype=np.int64).T.copy()
    return nodes, tri


def sphere(radius, level):
    nodes, tri = icosphere(level)
    return nodes * radius, tri


def spheroid(a, b, level):
    """Prolate/oblate spheroid with semi-axis `a` along x and `b` along y, z."""
    nodes, tri = icosphere(level)
    nodes = nodes.copy()
    nodes[0] *= a
    nodes[1] *= b
    nodes[2] *= b
    return nodes, tri


def tetrahedron():
    """The reference's 4-node supplied-mesh fixture (test/basic/bem/formulations.jl:229-231 shape):
    a right tetrahedron with outward faces."""
    nodes = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]).T.copy()
    tri = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]], dtype=np.int64).T.copy()
    return nodes, tri
