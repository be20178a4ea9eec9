"""GPU: the compressed path's solver layer.

  * Two-level preconditioner (reference compression/fluid.jl:166-301): block
    Jacobi on the diagonal leaf blocks plus an orthonormal coarse space of the
    degree <= 2 harmonic polynomials on each trace.  Every preconditioner
    (none, block Jacobi, two-level) reaches the same solution (the dense
    LU's, to the compression accuracy 2e-5), and the two-level form needs no
    more iterations than block Jacobi.
  * Recycled solutions (compression/fluid.jl:303-419): solving the same
    incidences again needs no GMRES iteration, nearby incidences need fewer
    than a cold start and give the same amplitudes, and a space that does not
    fit is dropped.
"""
import math
import sys

import numpy as np
import pytest

import bemhip as bh
from oracle import meshes

pytestmark = pytest.mark.gpu

FLUID = (1.2, 1.1)
K_F = 500.0


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    say("[gpu] creating context on device 0")
    c = bh.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def spheroid3(ctx):
    nodes, tri = meshes.spheroid(0.02, 0.01, 3)
    return bh.Mesh(nodes, tri + 1, ctx=ctx)


@pytest.mark.parametrize("body", ["fluid", "rigid"])
def test_preconditioners_agree_and_two_level_helps(ctx, spheroid3, body):
    gm = spheroid3
    bc = bh.FluidFilled(*FLUID) if body == "fluid" else bh.Rigid()
    angles = [(0.0, 0.0), (1.1, 0.4), (2.5, 0.1)]
    dense = bh.Operator(gm, bc, K_F).solve(angles)
    iters = {}
    for prec in ("none", "block_jacobi", "twolevel"):
        say(f"[gpu] {body} spheroid L3, preconditioner {prec}")
        sol = bh.HMatrix(gm, bc, K_F, preconditioner=prec, recycle_dimension=0).solve(angles)
        iters[prec] = sol.diagnostics(0)["iterations"]
        for r in range(len(angles)):
            d = sol.diagnostics(r)
            assert d["converged"] and d["relative_residual"] <= 1e-8
            f, f0 = sol.far_field(r)[0], dense.far_field(r)[0]
            assert abs(f - f0) / abs(f0) < 2e-5
    say(f"[gpu]   GMRES iterations {iters}")
    assert iters["twolevel"] <= iters["block_jacobi"]


def test_recycled_solutions(ctx, spheroid3):
    gm = spheroid3
    bc = bh.FluidFilled(*FLUID)
    first = [(0.2 * r, 0.1) for r in range(4)]
    near = [(0.2 * r + 0.05, 0.1) for r in range(4)]
    H = bh.HMatrix(gm, bc, K_F, recycle_dimension=8)
    s1 = H.solve(first)
    it1 = s1.diagnostics(0)["iterations"]
    s2 = H.solve(first)
    it2 = max(s2.diagnostics(r)["iterations"] for r in range(4))
    s3 = H.solve(near)
    it3 = s3.diagnostics(0)["iterations"]
    cold = bh.HMatrix(gm, bc, K_F, recycle_dimension=0).solve(near)
    itc = cold.diagnostics(0)["iterations"]
    say(f"[gpu] recycling: first {it1}, repeat {it2}, nearby {it3}, nearby cold {itc}")
    assert it2 == 0
    assert it3 < itc
    for r in range(4):
        assert s2.diagnostics(r)["converged"] and s2.diagnostics(r)["relative_residual"] <= 1e-8
        assert s3.diagnostics(r)["converged"] and s3.diagnostics(r)["relative_residual"] <= 1e-8
        f1, f2 = s1.far_field(r)[0], s2.far_field(r)[0]
        assert abs(f2 - f1) <= 1e-7 * abs(f1)
        f3, fc = s3.far_field(r)[0], cold.far_field(r)[0]
        assert abs(f3 - fc) <= 1e-6 * abs(fc)
    # a very different right-hand side still converges (the space is dropped
    # or improved, never trusted blindly)
    s4 = H.solve([(math.pi, 2.0)])
    assert s4.diagnostics(0)["converged"] and s4.diagnostics(0)["relative_residual"] <= 1e-8
    with pytest.raises(ValueError):
        bh.HMatrix(gm, bc, K_F, recycle_dimension=-1)
    with pytest.raises(ValueError):
        bh.HMatrix(gm, bc, K_F, preconditioner="ilu")
