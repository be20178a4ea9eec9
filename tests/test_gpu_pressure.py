"""GPU: pressure at points of a full-3D BEM solution (the reference's
pressure(solution, points; field), src/fields/pressure.jl:404-472) through
bem_pressure, against the CPU FP64 oracle evaluating the same surface
traces (1e-9 relative: same quadrature, FP64), including points close
enough to the surface for the subdivided near rule, interior points of a
fluid body, the field/region errors of the reference (pressure.jl:496-501),
and a hierarchical-matrix solution against the dense one (2e-5, the
compressed path's bar)."""
import math
import sys

import numpy as np
import pytest

import bemhip as bh
from oracle import meshes
from oracle import oracle as O

pytestmark = pytest.mark.gpu

K1 = 2 * math.pi * 38000.0 / 1477.4
A1 = 0.01
FLUID1 = (1028.9 / 1026.8, 1480.3 / 1477.4)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    say("[gpu] creating context on device 0")
    c = bh.Context(0)
    yield c
    c.close()


def _dirs(n, seed):
    v = np.random.default_rng(seed).standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


@pytest.mark.parametrize("bcname", ["rigid", "soft", "fluid"])
def test_pressure_matches_oracle(ctx, bcname):
    say(f"[gpu] pressure at points, sphere L3, {bcname}")
    nodes, tri = meshes.sphere(A1, 3)
    gm, om = bh.Mesh(nodes, tri + 1, ctx=ctx), O.Mesh(nodes, tri)
    bc = {"rigid": bh.Rigid(), "soft": bh.PressureRelease(), "fluid": bh.FluidFilled(*FLUID1)}[bcname]
    beta, alpha = 0.6, 0.3
    sol = bh.bem(gm, bc, K1, incidence_angle=beta, incidence_azimuth=alpha)
    p, q = sol.p_scat, sol.dpdn_scat
    fluid = FLUID1 if bcname == "fluid" else None
    # outside: far, moderate, and within twice a panel's edge of the surface
    radii = np.concatenate([np.full(8, 3.0), np.full(8, 1.5), np.full(8, 1.03)]) * A1
    pts = _dirs(24, 7) * radii[:, None]
    for field in ("total", "scattered", "incident"):
        v = bh.pressure(sol, pts.T, field=field)
        ref, reg, _ = om.pressure(K1, pts, p, q, beta=beta, alpha=alpha, field=field, fluid=fluid)
        err = np.max(np.abs(v - ref)) / np.max(np.abs(ref))
        say(f"[gpu]   {field} outside: rel err {err:.2e}")
        assert np.all(reg == 0) and err < 1e-9
    one = bh.pressure(sol, pts[0], field="scattered")
    assert isinstance(one, complex)
    inside = _dirs(6, 8) * (np.array([0.2, 0.4, 0.6, 0.8, 0.9, 0.95]) * A1)[:, None]
    if bcname == "fluid":
        for field in ("total", "interior"):
            v = bh.pressure(sol, inside, field=field)
            ref, reg, _ = om.pressure(K1, inside, p, q, beta=beta, alpha=alpha, field=field, fluid=fluid)
            err = np.max(np.abs(v - ref)) / np.max(np.abs(ref))
            say(f"[gpu]   {field} inside: rel err {err:.2e}")
            assert np.all(reg == 1) and err < 1e-9
        with pytest.raises(ValueError, match="no acoustic pressure"):
            bh.pressure(sol, inside, field="scattered")
        with pytest.raises(ValueError, match="no acoustic pressure"):
            bh.pressure(sol, pts, field="interior")
    else:
        for field in ("total", "scattered", "interior"):
            with pytest.raises(ValueError, match="no acoustic pressure"):
                bh.pressure(sol, inside, field=field)
    with pytest.raises(ValueError):
        bh.pressure(sol, np.array([[np.nan, 0.0, 0.0]]).T)
    with pytest.raises(ValueError):
        bh.pressure(sol, pts, field="shell")


def test_pressure_of_hierarchical_solution_matches_dense(ctx):
    say("[gpu] pressure from a hierarchical-matrix solution vs the dense solution")
    gm = bh.mesh(bh.Sphere(A1), level=4, ctx=ctx)
    H = bh.HMatrix(gm, bh.Rigid(), K1, formulation="cbie")
    hs = H.solve([(0.9, 0.0)])
    op = bh.Operator(gm, bh.Rigid(), K1, formulation="cbie")
    ds = op.solve([(0.9, 0.0)])
    pts = (_dirs(16, 11) * (np.linspace(1.2, 4.0, 16) * A1)[:, None]).T
    vh = hs.pressure(pts, "scattered")
    vd = ds.pressure(pts, "scattered")
    err = np.max(np.abs(vh - vd)) / np.max(np.abs(vd))
    say(f"[gpu]   hierarchical vs dense: {err:.2e}")
    assert err < 2e-5
