"""GPU parity: libbemhip (through its C ABI) against the CPU FP64 oracle.

Tolerances: FP64 assembly entries 1e-12 relative to the matrix max; FP64
solutions 1e-9 relative in the scattering amplitude (the plain GPU form's
parity bar); the mixed form |dTS| <= 0.01 dB against the FP64 form and
backward errors <= 1e-13 (src/solvers/bem/fluid_system.jl acceptance).  The
reference's mixed-versus-double amplitude bar is 1e-8 (test/basic/bem/
formulations.jl:431-436) because its refinement runs against the FP64 matrix;
here regular panel pairs are assembled in FP32 and the refinement converges
to that FP32-assembled system, which sits ~1e-7 from the FP64 one, so the
amplitude bar is 1e-5 (the refinement's own 1e-9 parity against the oracle's
solve of the same FP32-assembled matrix is test_mixed_refinement_reaches_fp32_
system_solution).  Physics against the sphere modal series at the reference's
0.1 dB / 1 % bar (test/extended/bem/sphere/full3d.jl:139-140).
"""
import math
import sys

import numpy as np
import pytest

import scipy.linalg as sla

import bemhip as bh
from oracle import meshes, modal
from oracle import oracle as O

pytestmark = pytest.mark.gpu

K_SPHERE38 = 2 * math.pi * 38000.0 / 1477.4
A_SPHERE38 = 0.01
FLUID_SPHERE38 = (1028.9 / 1026.8, 1480.3 / 1477.4)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    say("[gpu] creating context on device 0")
    c = bh.Context(0)
    yield c
    c.close()


def _sphere(ctx, level, radius=A_SPHERE38):
    nodes, tri = meshes.sphere(radius, level)
    return bh.Mesh(nodes, tri + 1, ctx=ctx), O.Mesh(nodes, tri)


def _reconstruct(LU, piv):
    n = LU.shape[0]
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    return L @ U, piv


def _permute(A, piv):
    PA = A.copy()
    for i, p in enumerate(piv):
        if p - 1 != i:
            PA[[i, p - 1]] = PA[[p - 1, i]]
    return PA


@pytest.mark.parametrize("n", [8, 64, 65, 128, 130, 256, 300, 1000, 2600])
def test_device_lu(ctx, n):
    """Device LU against LAPACK zgetrf: P A = L U to FP64 rounding, the solve
    against numpy, and (n <= 256: one 256-row tournament leaf, which is plain
    partial pivoting) the identical pivot rows and factors."""
    say(f"[gpu] device LU n={n}")
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    LU, piv, x = bh.debug_lu(A, b=b, ctx=ctx)
    LUp, _ = _reconstruct(LU, piv)
    assert np.abs(_permute(A, piv) - LUp).max() < 1e-12 * np.abs(A).max() * max(1.0, n / 100)
    xs = np.linalg.solve(A, b)
    assert np.abs(x - xs).max() < 1e-9 * np.abs(xs).max()
    if n <= 256:
        lu0, piv0 = sla.lu_factor(A)
        assert np.array_equal(piv, piv0 + 1)
        assert np.abs(LU - lu0).max() < 1e-10 * np.abs(lu0).max()
    LU32, piv32, x32 = bh.debug_lu(A, single=True, b=b, ctx=ctx)
    lu32 = sla.lu_factor(A.astype(np.complex64))
    x_lapack32 = sla.lu_solve(lu32, b.astype(np.complex64))
    ref32 = np.abs(x_lapack32 - xs).max()
    assert np.abs(x32 - xs).max() < 20 * ref32 + 1e-6 * np.abs(xs).max()
    # conjugate-transpose solves with the same factors (the rcond estimator's B^H products)
    _, _, xh = bh.debug_lu(A, b=b, ctx=ctx, adjoint=True)
    xhs = np.linalg.solve(A.conj().T, b)
    assert np.abs(xh - xhs).max() < 1e-9 * np.abs(xhs).max()
    _, _, xh32 = bh.debug_lu(A, single=True, b=b, ctx=ctx, adjoint=True)
    refh32 = np.abs(sla.lu_solve(lu32, b.astype(np.complex64), trans=2) - xhs).max()
    assert np.abs(xh32 - xhs).max() < 20 * refh32 + 1e-6 * np.abs(xhs).max()


def test_device_lu_aligned_block_and_edge_strips(ctx):
    """n = 6500: the first outer trailing update is 5988 x 5988 with K = 512,
    whose 128-aligned block (46 x 46 = 2116 tiles) runs in the full-tile
    form and its 100-wide right and bottom strips in the general form (FP32
    3M and FP64 products).  A strip written twice, missed or offset would
    wreck the factors, so the solves are checked against numpy / LAPACK."""
    n = 6500
    say(f"[gpu] device LU n={n} (aligned block + edge strips)")
    rng = np.random.default_rng(6500)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    xs = np.linalg.solve(A, b)
    _, _, x = bh.debug_lu(A, b=b, ctx=ctx)
    assert np.abs(x - xs).max() < 1e-9 * np.abs(xs).max()
    _, _, x32 = bh.debug_lu(A, single=True, b=b, ctx=ctx)
    ref32 = np.abs(sla.lu_solve(sla.lu_factor(A.astype(np.complex64)), b.astype(np.complex64)) - xs).max()
    assert np.abs(x32 - xs).max() < 20 * ref32 + 1e-6 * np.abs(xs).max()


@pytest.mark.parametrize("n", [300, 1000, 2600, 4100])
def test_device_lu_tournament_pivots(ctx, n):
    """Multi-leaf panels: the device's pivot sequence is the tournament's
    (tests/tslu_model.py), index for index, in FP64."""
    import tslu_model as tm
    say(f"[gpu] device LU tournament pivots n={n}")
    rng = np.random.default_rng(100 + n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    _, piv = bh.debug_lu(A, ctx=ctx)
    assert np.array_equal(piv - 1, tm.blocked_lu_pivots(A))


def test_device_lu_singular_column_reports_zero_pivot(ctx):
    say("[gpu] device LU on a singular matrix")
    A = np.eye(100, dtype=np.complex128)
    A[:, 37] = 0
    with pytest.raises(RuntimeError, match="singular"):
        bh.debug_lu(A, ctx=ctx)


CASES = [
    ("rigid_bm", bh.Rigid(), O.RIGID_BM, None),
    ("soft_bm", bh.PressureRelease(), O.SOFT_BM, None),
    ("rigid_cbie", bh.Rigid(), O.RIGID_CBIE, "cbie"),
    ("soft_cbie", bh.PressureRelease(), O.SOFT_CBIE, "cbie"),
    ("fluid_muller", bh.FluidFilled(*FLUID_SPHERE38), O.FLUID_MULLER, None),
    ("fluid_cbie", bh.FluidFilled(1.2, 1.1), O.FLUID_CBIE, "cbie"),
]


@pytest.mark.parametrize("name,bc,kind,form", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_assembled_matrix_matches_oracle(ctx, name, bc, kind, form, precision):
    say(f"[gpu] assembly {name} {precision} L2 sphere")
    gm, om = _sphere(ctx, 2)
    op = bh.Operator(gm, bc, K_SPHERE38, formulation=form, precision=precision)
    A = op.matrix()
    g, h = (bc.density_contrast, bc.soundspeed_contrast) if isinstance(bc, bh.FluidFilled) else (1.0, 1.0)
    R = om.assemble(kind, K_SPHERE38, g, h)
    err = np.abs(A - R).max() / np.abs(R).max()
    tol = 1e-12 if precision == "double" else 2e-6
    assert err < tol, f"{name} {precision}: rel err {err:.3e}"


@pytest.mark.parametrize("bcname", ["rigid", "soft", "fluid"])
def test_sphere38_solution_matches_oracle_fp64(ctx, bcname):
    say(f"[gpu] 38 kHz sphere {bcname} FP64 vs oracle")
    gm, om = _sphere(ctx, 3)
    if bcname == "fluid":
        bc = bh.FluidFilled(*FLUID_SPHERE38)
        p0, q0, _ = O.solve_fluid(om, *FLUID_SPHERE38, K_SPHERE38, beta=0.0)
    else:
        bc = bh.Rigid() if bcname == "rigid" else bh.PressureRelease()
        p0, q0, _ = O.solve_boundary(om, bcname, K_SPHERE38, beta=0.0)
    p, q, _, diag = bh.solve_full_bem(bc, K_SPHERE38, gm, incidence_angle=0.0, return_diagnostics=True)
    assert np.abs(p - p0).max() / np.abs(p0).max() < 1e-9
    assert np.abs(q - q0).max() / np.abs(q0).max() < 1e-9
    f0 = om.far_field(K_SPHERE38, -O.direction(0.0), p0, q0)[0]
    sol = bh.bem(gm, bc, K_SPHERE38, incidence_angle=0.0)
    f = bh.scattering_amplitude(sol)
    assert abs(f - f0) / abs(f0) < 1e-9
    assert diag["relative_residual"] < 1e-12


@pytest.mark.parametrize("bcname,level", [("rigid", 3), ("fluid", 3), ("rigid", 4), ("soft", 4), ("fluid", 4)])
def test_sphere_against_modal_series(ctx, bcname, level):
    say(f"[gpu] modal {bcname} L{level}")
    gm = bh.mesh(bh.Sphere(A_SPHERE38), level=level, ctx=ctx)
    bc = {"rigid": bh.Rigid(), "soft": bh.PressureRelease(), "fluid": bh.FluidFilled(*FLUID_SPHERE38)}[bcname]
    mbc = {"rigid": ("rigid",), "soft": ("soft",), "fluid": ("fluid",) + FLUID_SPHERE38}[bcname]
    sol = bh.bem(gm, bc, K_SPHERE38, incidence_angle=0.0)
    f = bh.scattering_amplitude(sol)
    fm = modal.form_function(mbc, K_SPHERE38, A_SPHERE38)
    assert abs(bh.target_strength(f) - modal.target_strength(fm)) < 0.1
    assert abs(f - fm) / abs(fm) < 0.01


def test_mixed_matches_double_sphere38_fluid(ctx):
    say("[gpu] mixed vs double, 38 kHz fluid sphere L3")
    gm, _ = _sphere(ctx, 3)
    bc = bh.FluidFilled(*FLUID_SPHERE38)
    sd = bh.bem(gm, bc, K_SPHERE38, incidence_angle=0.0)
    sm = bh.bem(gm, bc, K_SPHERE38, incidence_angle=0.0, precision="mixed")
    fd, fm = bh.scattering_amplitude(sd), bh.scattering_amplitude(sm)
    d = sm.diagnostics
    assert d["factor_precision"] == "single"
    assert d["converged"]
    assert d["normwise_backward_error"] <= 1e-13 and d["componentwise_backward_error"] <= 1e-13
    assert abs(bh.target_strength(fm) - bh.target_strength(fd)) <= 0.01
    # the mixed form solves the FP32-assembled system; its distance to the
    # FP64 system bounds the amplitude difference
    assert abs(fm - fd) / abs(fd) < 1e-5


def test_mixed_refinement_reaches_fp32_system_solution(ctx):
    say("[gpu] mixed refinement vs oracle solve of the same FP32-assembled matrix")
    gm, om = _sphere(ctx, 2)
    bc = bh.FluidFilled(*FLUID_SPHERE38)
    op = bh.Operator(gm, bc, K_SPHERE38, precision="mixed")
    A = op.matrix()  # FP32-assembled system, widened to FP64
    p0, q0, dd = O.solve_fluid(om, *FLUID_SPHERE38, K_SPHERE38, A=A)
    sol = op.solve([(0.0, 0.0)])
    p, q = sol.traces(0)
    assert np.abs(p - p0).max() / np.abs(p0).max() < 1e-9
    assert np.abs(q - q0).max() / np.abs(q0).max() < 1e-9


def test_forced_promotion(ctx):
    say("[gpu] forced promotion to FP64")
    gm, _ = _sphere(ctx, 2)
    bc = bh.FluidFilled(*FLUID_SPHERE38)
    s = bh.bem(gm, bc, K_SPHERE38, incidence_angle=0.0, precision="mixed", _force_promotion=True)
    d = s.diagnostics
    assert d["factor_precision"] == "double" and d["refinement_reason"] == "condition"
    sd = bh.bem(gm, bc, K_SPHERE38, incidence_angle=0.0)
    assert abs(bh.scattering_amplitude(s) - bh.scattering_amplitude(sd)) / abs(bh.scattering_amplitude(sd)) < 1e-5


def test_condition_guard_forces_promotion(ctx):
    """refinement=(condition_guard=1.0,) promotes with reason :condition
    (test/basic/bem/formulations.jl:424-425)."""
    say("[gpu] condition guard 1.0")
    gm, _ = _sphere(ctx, 2)
    s = bh.bem(gm, bh.FluidFilled(*FLUID_SPHERE38), K_SPHERE38, incidence_angle=0.0, precision="mixed",
               refinement={"condition_guard": 1.0})
    d = s.diagnostics
    assert d["factor_precision"] == "double" and d["refinement_reason"] == "condition"


def test_reciprocal_condition_estimate_matches_lapack(ctx):
    """The single-precision rcond estimate against LAPACK cgecon('I') on the
    same equilibrated FP32 matrix (fluid_system.jl:218-222)."""
    say("[gpu] rcond estimate vs cgecon")
    from scipy.linalg import lapack
    gm, _ = _sphere(ctx, 2)
    bc = bh.FluidFilled(*FLUID_SPHERE38)
    op = bh.Operator(gm, bc, K_SPHERE38, precision="mixed")
    A = op.matrix()
    rows, cols = O.equilibrate(A)
    low = (A / rows[:, None] / cols[None, :]).astype(np.complex64)
    anorm = float(np.abs(low).sum(axis=1).max())
    lu, piv, info = lapack.cgetrf(low)
    rc_ref, _ = lapack.cgecon(lu, anorm, norm="I")
    s = bh.bem(gm, bc, K_SPHERE38, incidence_angle=0.0, precision="mixed")
    rc = s.diagnostics["single_reciprocal_condition"]
    assert np.isfinite(rc) and 0 < rc <= 1
    assert 0.1 < rc / rc_ref < 10


def test_iteration_limit_promotes(ctx):
    say("[gpu] maxiter=0 forces iteration_limit promotion")
    gm, _ = _sphere(ctx, 2)
    s = bh.bem(gm, bh.FluidFilled(*FLUID_SPHERE38), K_SPHERE38, incidence_angle=0.0, precision="mixed",
               refinement={"maxiter": 0})
    d = s.diagnostics
    assert d["factor_precision"] == "double"
    assert d["refinement_reason"] == "iteration_limit"


def test_matched_medium_is_invisible(ctx):
    say("[gpu] matched medium g=h=1, L4")
    gm = bh.mesh(bh.Sphere(1.0), level=4, ctx=ctx)
    s = bh.bem(gm, bh.FluidFilled(1.0, 1.0), 1.0)
    assert abs(bh.scattering_amplitude(s)) < 1e-12


def test_spheroid_L4_fluid_matches_oracle(ctx):
    say("[gpu] spheroid L4 fluid Mueller (2N=10240) vs oracle")
    nodes, tri = meshes.spheroid(0.02, 0.01, 4)
    gm = bh.Mesh(nodes, tri + 1, ctx=ctx)
    om = O.Mesh(nodes, tri)
    k = 500.0
    g, h = 1.0357, 1.0279
    p0, q0, _ = O.solve_fluid(om, g, h, k, beta=0.0)
    f0 = om.far_field(k, -O.direction(0.0), p0, q0)[0]
    s = bh.bem(gm, bh.FluidFilled(g, h), k, incidence_angle=0.0)
    f = bh.scattering_amplitude(s)
    assert abs(f - f0) / abs(f0) < 1e-9
    sm = bh.bem(gm, bh.FluidFilled(g, h), k, incidence_angle=0.0, precision="mixed")
    assert abs(bh.target_strength(bh.scattering_amplitude(sm)) - bh.target_strength(f)) <= 0.01


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_low_frequency_mueller_calderon_blocks(ctx, precision):
    """At max(k, k_int) * radius <= pi/2 the Mueller derivative blocks come
    from the Calderon identities (full.jl:201-207,319-339): the device's
    system equals the oracle's, and so do the traces."""
    say(f"[gpu] low-frequency Mueller (Calderon blocks), spheroid L2, {precision}")
    nodes, tri = meshes.spheroid(0.02, 0.01, 2)
    gm = bh.Mesh(nodes, tri + 1, ctx=ctx)
    om = O.Mesh(nodes, tri)
    g, h, k = 1.2, 1.1, 50.0
    assert O.calderon_applies(om, k, h, "muller")
    A = bh.Operator(gm, bh.FluidFilled(g, h), k, precision=precision).matrix()
    R = O.assemble_fluid(om, g, h, k)
    err = np.abs(A - R).max() / np.abs(R).max()
    assert err < (1e-10 if precision == "double" else 2e-6), err
    p, q, _, d = bh.solve_full_bem(bh.FluidFilled(g, h), k, gm, incidence_angle=0.4, precision=precision,
                                   return_diagnostics=True)
    # the mixed path refines against its FP32-stored system: solve that one
    p0, q0, _ = O.solve_fluid(om, g, h, k, beta=0.4, A=None if precision == "double" else A)
    assert np.abs(p - p0).max() / np.abs(p0).max() < 1e-9
    assert np.abs(q - q0).max() / np.abs(q0).max() < 1e-9
    assert d["derivative_evaluation"] == {"exterior": "calderon", "interior": "calderon"}
    _, _, _, d2 = bh.solve_full_bem(bh.FluidFilled(g, h), 2000.0, gm, return_diagnostics=True)
    assert d2["derivative_evaluation"] == {"exterior": "direct", "interior": "direct"}


def test_incidence_sweep_matches_single_solves(ctx):
    say("[gpu] incidence-angle sweep vs single solves")
    gm, _ = _sphere(ctx, 2)
    angles = [0.0, 0.7, 1.9, math.pi]
    amps, ts = bh.incidence_angle_sweep(gm, bh.Rigid(), K_SPHERE38, angles)
    for a, f in zip(angles, amps):
        s = bh.bem(gm, bh.Rigid(), K_SPHERE38, incidence_angle=a)
        assert abs(bh.scattering_amplitude(s) - f) <= 1e-12 * abs(f)


def test_batched_sweep_matches_single_solves(ctx):
    say("[gpu] bem_sweep vs single solves")
    m2 = bh.mesh(bh.Spheroid(0.02, 0.01), level=2, ctx=ctx)
    m3 = bh.mesh(bh.Spheroid(0.02, 0.01), level=3, ctx=ctx)
    fl = bh.FluidFilled(1.0357, 1.0279)
    ks = [50.0, 200.0, 400.0]
    ms = [m2, m3, m2]
    probs = [bh.make_problem(fl, k, precision="mixed") for k in ks]
    amps, lat, diags = bh.batched_sweep(ms, probs, [(0.0, 0.0)] * 3)
    for m, k, f in zip(ms, ks, amps):
        s = bh.bem(m, fl, k, incidence_angle=0.0, precision="mixed")
        assert abs(bh.scattering_amplitude(s) - f) <= 1e-12 * abs(f)
    assert np.all(lat > 0)


def test_sweep_heavy_lane_limit_keeps_results(ctx):
    """Solves with 16384 or more unknowns are admitted at most `heavy lanes`
    at a time while lighter ones fill the other lanes; the schedule must not
    change any amplitude."""
    say("[gpu] bem_sweep heavy-lane limit 1 vs 8")
    heavy = bh.mesh(bh.Sphere(0.01), level=5, ctx=ctx)  # 20480 unknowns
    light = bh.mesh(bh.Sphere(0.01), level=3, ctx=ctx)
    ms = [heavy, light, heavy, light, light]
    probs = [bh.make_problem(bh.Rigid(), k, precision="mixed") for k in (200.0, 300.0, 100.0, 150.0, 250.0)]
    out = {}
    try:
        for lim in (1, 8):
            ctx.set_heavy_lanes(lim)
            out[lim], _, diags = bh.batched_sweep(ms, probs, [(0.0, 0.0)] * len(ms), ctx=ctx)
            assert all(d["converged"] for d in diags)
        with pytest.raises(ValueError):
            ctx.set_heavy_lanes(0)
        with pytest.raises(ValueError):
            ctx.set_heavy_lanes(9)
    finally:
        ctx.set_heavy_lanes(1)
    assert np.all(np.abs(out[1] - out[8]) <= 1e-12 * np.abs(out[8]))


def test_argument_errors(ctx):
    say("[gpu] argument errors")
    gm, _ = _sphere(ctx, 1, 1.0)
    for k in (0.0, -1.0, math.inf, math.nan):
        with pytest.raises(ValueError):
            bh.solve_full_bem(bh.Rigid(), k, gm)
    for bad in (bh.FluidFilled(0.0, 1.0), bh.FluidFilled(1.0, -1.0), bh.FluidFilled(math.inf, 1.0),
                bh.FluidFilled(1.0, math.nan)):
        with pytest.raises(ValueError):
            bh.solve_full_bem(bad, 1.0, gm)
    with pytest.raises(ValueError):
        bh.solve_full_bem(bh.FluidFilled(1.2, 1.1), 1.0, gm, formulation="burton_miller")
    with pytest.raises(ValueError):
        bh.solve_full_bem(bh.Rigid(), 1.0, gm, formulation="muller")
    with pytest.raises(ValueError):
        bh.solve_full_bem(bh.FluidFilled(1.2, 1.1), 1.0, gm, precision="mixed", equilibrate=False)
