"""CPU tests of the product library's host side (no GPU needed).

libbemhip.so loads without a device; these tests cover its exported symbols,
the host mesh generators (bit-exact against the oracle's restatement of the
same rule) and the supplied-mesh topology contract of the reference
(src/geometry/surface.jl:341-405), including its error messages.
"""
import math
import os
import re

import numpy as np
import pytest

import bemhip as bh
from oracle import meshes
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = bh.load_library()
    header = open(os.path.join(ROOT, "include", "bemhip.h")).read()
    declared = sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(bem_\w+)\s*\(", header, re.M)))
    assert len(declared) >= 30
    assert sorted(declared) == sorted(bh.EXPORTED)
    for name in declared:
        assert hasattr(L, name), name
    assert L.bem_version().startswith(b"bemhip")


@pytest.mark.parametrize("level", [0, 1, 2, 3, 4, 5])
def test_icosphere_connectivity_bit_exact(level):
    n1, t1 = bh.icosphere_arrays(level, 0.02, 0.01, 0.01)
    n2, t2 = meshes.spheroid(0.02, 0.01, level)
    assert np.array_equal(t1 - 1, t2)
    assert np.array_equal(n1, n2)
    assert t1.shape[1] == 20 * 4 ** level
    assert n1.shape[1] == 10 * 4 ** level + 2


def test_icosphere_sizes():
    L = bh.load_library()
    import ctypes
    nv, nt = ctypes.c_int64(), ctypes.c_int64()
    assert L.bem_icosphere_size(5, ctypes.byref(nv), ctypes.byref(nt)) == 0
    assert (nv.value, nt.value) == (10242, 20480)
    assert L.bem_icosphere_size(-1, ctypes.byref(nv), ctypes.byref(nt)) != 0


def test_validate_accepts_closed_meshes():
    for level in (0, 2, 3):
        n, t = bh.icosphere_arrays(level, 1.0)
        res = bh.validate_surface(n, t)
        assert res > 0
        assert O.validate(n, t - 1)[0] == 0
    n, t = meshes.tetrahedron()
    assert bh.validate_surface(n, t + 1) == pytest.approx(math.sqrt(2))


def _msg(nodes, tri1):
    with pytest.raises(ValueError) as e:
        bh.validate_surface(nodes, tri1)
    code, omsg, _ = O.validate(nodes, tri1 - 1)
    assert code != 0
    assert str(e.value) == omsg
    return str(e.value)


def test_validate_rejects_open_surface():
    n, t = bh.icosphere_arrays(1, 1.0)
    assert "incident triangles; expected two" in _msg(n, t[:, 1:])


def test_validate_rejects_inconsistent_orientation():
    n, t = bh.icosphere_arrays(1, 1.0)
    t = t.copy()
    t[[1, 2], 5] = t[[2, 1], 5]
    assert "inconsistent orientation" in _msg(n, t)


def test_validate_rejects_inward_orientation():
    n, t = bh.icosphere_arrays(1, 1.0)
    t = t[[0, 2, 1], :]
    assert "positive volume" in _msg(n, t)


def test_validate_rejects_disconnected():
    n1, t1 = bh.icosphere_arrays(1, 1.0)
    n2 = n1 + np.array([[5.0], [0], [0]])
    nodes = np.hstack([n1, n2])
    tri = np.hstack([t1, t1 + n1.shape[1]])
    assert "connected surface" in _msg(nodes, tri)


def test_validate_rejects_repeated_node_and_degenerate():
    n, t = meshes.tetrahedron()
    t1 = t.copy() + 1
    t1[1, 0] = t1[0, 0]
    assert "repeats a node" in _msg(n, t1)
    n2 = n.copy()
    n2[:, 3] = n2[:, 1]  # collapse the tetrahedron's apex onto a base vertex
    with pytest.raises(ValueError):
        bh.validate_surface(n2, t + 1)


def test_validate_rejects_bad_index():
    n, t = meshes.tetrahedron()
    t1 = t.copy() + 1
    t1[0, 0] = 99
    with pytest.raises(ValueError, match="invalid node index"):
        bh.validate_surface(n, t1)


def test_revolution_mesh_is_closed_and_outward():
    s = np.linspace(0.0, 1.0, 33)
    a, aspect = 0.01, 2.0
    xs = -aspect * a * np.cos(np.pi * s)
    rs = a * np.sin(np.pi * s)
    rs[0] = rs[-1] = 0.0
    n, t = bh.revolution_arrays(xs, rs, 48)
    assert n.shape[1] == 31 * 48 + 2 and t.shape[1] == 2 * 31 * 48
    bh.validate_surface(n, t)
    assert O.validate(n, t - 1)[0] == 0
    # polyhedral volume approaches the spheroid's (4/3) pi a^2 (aspect a)
    a0, b0, c0 = n[:, t[0] - 1], n[:, t[1] - 1], n[:, t[2] - 1]
    vol = np.sum(np.einsum("ij,ij->j", a0, np.cross(b0.T, c0.T).T)) / 6
    assert abs(vol / (4 / 3 * math.pi * a * a * aspect * a) - 1) < 0.02


def test_problem_defaults_follow_reference():
    p = bh.make_problem(bh.FluidFilled(1.2, 1.1), 3.0)
    assert p.formulation == bh.BEM_MULLER and p.equilibrate == 1 and p.precision == bh.BEM_DOUBLE
    assert p.refine_tolerance == 1e-13 and p.refine_maxiter == 8
    assert p.condition_guard == pytest.approx(10 * np.finfo(np.float32).eps)
    r = bh.make_problem(bh.Rigid(), 3.0)
    assert r.formulation == bh.BEM_BURTON_MILLER
    with pytest.raises(ValueError):
        bh.make_problem(bh.Rigid(), 3.0, precision="half")
    with pytest.raises(ValueError):
        bh.make_problem(bh.Rigid(), 3.0, refinement={"tolerance": 1e-10})
    with pytest.raises(ValueError):
        bh.make_problem(bh.FluidFilled(1, 1), 3.0, precision="mixed", refinement={"bogus": 1})
