"""CPU tests that pin the FP64 oracle before it is trusted as the checker.

Each check is independent of the oracle's own formulas: published golden
values from the reference's tests, closed-form identities of the polyhedral
surface, an independent (Duffy-transformed) quadrature of the self-panel
integrals, convergence of the near-field subdivision, scipy's special
functions and LAPACK.
"""
import math

import numpy as np
import pytest
from scipy import special

from oracle import meshes, modal
from oracle import oracle as O

K1 = 2 * math.pi * 38000.0 / 1477.4
A1 = 0.01
FLUID1 = (1028.9 / 1026.8, 1480.3 / 1477.4)


# ---------------------------------------------------------------------------
# modal series (test/basic/modal/sphere.jl:9-13)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("bc,expected,tol", [
    (("rigid",), -49.088291, 1e-4),
    (("soft",), -44.997865, 1e-4),
    (("fluid",) + FLUID1, -94.278687, 1e-3),
])
def test_modal_goldens(bc, expected, tol):
    f = modal.form_function(bc, K1, A1)
    assert abs(modal.target_strength(f) - expected) < tol


def test_modal_derivative_rule_matches_scipy():
    x = 1.7
    for n in range(0, 12):
        assert abs(modal.jsd(n, x) - special.spherical_jn(n, x, derivative=True)) < 1e-13
        yd = (n / x) * modal.ys(n, x) - modal.ys(n + 1, x)
        ref = special.spherical_yn(n, x, derivative=True)
        assert abs(yd - ref) <= 1e-13 * max(1.0, abs(ref))


def test_legendre_recurrence():
    x = 0.3
    for n in range(0, 15):
        assert abs(modal.legendre_p(n, x) - special.eval_legendre(n, x)) < 1e-13


def test_mode_count_rule():
    assert modal.default_mode_count(0.1) == 20
    assert modal.default_mode_count(10.0) == math.ceil(25 + 4 * 11 ** (1 / 3) + 10)


# ---------------------------------------------------------------------------
# quadrature rules
# ---------------------------------------------------------------------------
def test_seven_point_rule_is_degree_five():
    nodes = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.0, 1.0, 0], [0.0, 0, 1.0]]).T
    tri = np.array([[1], [2], [3]])  # face (1,2,3) of the unit simplex, 0-based
    m = O.Mesh(nodes, tri)
    xyz, w = m.quadrature()
    # map to barycentric (u, v) on the face: integrate u^a v^b over the triangle
    P0, P1, P2 = nodes[:, 1], nodes[:, 2], nodes[:, 3]
    area = 0.5 * np.linalg.norm(np.cross(P1 - P0, P2 - P0))
    M = np.column_stack([P1 - P0, P2 - P0])
    uv = np.linalg.lstsq(M, (xyz - P0).T, rcond=None)[0]
    for a in range(6):
        for b in range(6 - a):
            exact = 2 * area * math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            got = float(np.sum(w * uv[0] ** a * uv[1] ** b))
            assert abs(got - exact) < 1e-14, (a, b)


def _duffy_self(V, x, f, n=48):
    """Integrate f(R) over a flat triangle with vertices V (3x3) about the
    in-plane point x, splitting at x and Duffy-transforming each part."""
    gx, gw = np.polynomial.legendre.leggauss(n)
    u = 0.5 * (gx + 1)
    wu = 0.5 * gw
    total = 0j
    for e in range(3):
        P, Q = V[e], V[(e + 1) % 3]
        J = np.linalg.norm(np.cross(P - x, Q - P))
        for ui, wi in zip(u, wu):
            for vj, wj in zip(u, wu):
                d = (P - x) + vj * (Q - P)
                R = ui * np.linalg.norm(d)
                total += wi * wj * ui * J * f(R)
    return total


@pytest.mark.parametrize("k", [1e-6, 3.0, 40.0])
def test_self_single_layer_matches_duffy(k):
    nodes, tri = meshes.sphere(0.05, 1)
    m = O.Mesh(nodes, tri)
    for t in (0, 7, 33):
        V = nodes[:, tri[:, t]].T
        x = m.centroids[t]
        S = m.pair_ops(t, t, k)[0]
        ref = _duffy_self(V, x, lambda R: np.exp(1j * k * R) / (4 * math.pi * R))
        assert abs(S - ref) / abs(ref) < 1e-11


@pytest.mark.parametrize("k", [3.0, 40.0])
def test_self_hypersingular_wavenumber_part_matches_duffy(k):
    # H(k) - H(0) is weakly singular: integrand [e^{ikR}(1 - ikR) - 1]/(4 pi R^3)
    nodes, tri = meshes.sphere(0.05, 1)
    m = O.Mesh(nodes, tri)
    for t in (0, 19):
        V = nodes[:, tri[:, t]].T
        x = m.centroids[t]
        dH = m.pair_ops(t, t, k)[3] - m.pair_ops(t, t, 1e-9)[3]

        def f(R):
            kr = k * R
            if kr < 1e-3:  # series: k^2/(2R) + i k^3/3 + ...
                return (k * k / (2 * R) + 1j * k ** 3 / 3 - k ** 4 * R / 8) / (4 * math.pi)
            return (np.exp(1j * kr) * (1 - 1j * kr) - 1) / (4 * math.pi * R ** 3)

        ref = _duffy_self(V, x, f)
        assert abs(dH - ref) / abs(ref) < 1e-9


def test_flat_panel_self_double_layers_vanish():
    nodes, tri = meshes.sphere(1.0, 2)
    m = O.Mesh(nodes, tri)
    e = m.pair_ops(5, 5, 2.0)
    assert e[1] == 0 and e[2] == 0


@pytest.mark.parametrize("level", [2, 3])
def test_polyhedral_laplace_identities(level):
    """On a closed polyhedron, for x at a panel centroid: sum_j D0_ij = -1/2
    (solid angle) and sum_j H0_ij = 0 (D0[1] is constant inside).  This pins
    the sign and size of the finite-part self term (|H0_ii| ~ 10 here)."""
    nodes, tri = meshes.sphere(1.0, level)
    m = O.Mesh(nodes, tri)
    k = 1e-9
    D = m.assemble(O.OP_D, k)
    H = m.assemble(O.OP_H, k)
    assert np.abs(D.sum(axis=1) + 0.5).max() < 1e-6
    assert np.abs(H.sum(axis=1)).max() < 2e-5
    assert np.abs(np.diag(H)).min() > 1.0


def test_near_and_regular_quadrature_accuracy():
    nodes, tri = meshes.sphere(1.0, 2)
    m = O.Mesh(nodes, tri)
    worst_near = worst_reg = 0.0
    for j in range(1, m.n):
        ref = m.pair_ops(0, j, 3.0, level=6)
        got = m.pair_ops(0, j, 3.0)
        e = float((np.abs(got - ref) / np.abs(ref)).max())
        if m.is_near(0, j):
            worst_near = max(worst_near, e)
        else:
            worst_reg = max(worst_reg, e)
    assert worst_near < 5e-6
    assert worst_reg < 5e-5


def test_kernel_symmetries():
    """S is symmetric for equal-area panels pairs; D(i,j) and K'(j,i) coincide
    for the centroid rule when the Gauss points sit at the centroid, so check
    the exact reciprocity S_ij(A_i) ~ S_ji(A_j) on far pairs instead."""
    nodes, tri = meshes.sphere(1.0, 2)
    m = O.Mesh(nodes, tri)
    i, j = 0, 200
    assert not m.is_near(i, j)
    sij = m.pair_ops(i, j, 2.0)[0] / m.area[j]
    sji = m.pair_ops(j, i, 2.0)[0] / m.area[i]
    assert abs(sij - sji) / abs(sij) < 1e-3


# ---------------------------------------------------------------------------
# whole solves
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("bcname,level", [("rigid", 3), ("fluid", 3), ("soft", 4)])
def test_oracle_against_modal(bcname, level):
    nodes, tri = meshes.sphere(A1, level)
    m = O.Mesh(nodes, tri)
    if bcname == "fluid":
        p, q, d = O.solve_fluid(m, *FLUID1, K1)
        fm = modal.form_function(("fluid",) + FLUID1, K1, A1)
    else:
        p, q, d = O.solve_boundary(m, bcname, K1)
        fm = modal.form_function((bcname,), K1, A1)
    f = m.far_field(K1, -O.direction(0.0), p, q)[0]
    assert abs(O.target_strength(f) - modal.target_strength(fm)) < 0.1
    assert abs(f - fm) / abs(fm) < 0.01
    assert d["relative_residual"] < 1e-12


def test_oracle_convergence_with_refinement():
    errs = []
    fm = modal.form_function(("rigid",), K1, A1)
    for level in (2, 3):
        nodes, tri = meshes.sphere(A1, level)
        m = O.Mesh(nodes, tri)
        p, q, _ = O.solve_boundary(m, "rigid", K1)
        errs.append(abs(m.far_field(K1, -O.direction(0.0), p, q)[0] - fm) / abs(fm))
    assert errs[1] < 0.5 * errs[0]


def test_matched_medium_invisible():
    nodes, tri = meshes.sphere(1.0, 2)
    m = O.Mesh(nodes, tri)
    p, q, _ = O.solve_fluid(m, 1.0, 1.0, 1.0)
    f = m.far_field(1.0, -O.direction(0.0), p, q)[0]
    assert abs(f) < 1e-12


def test_calderon_blocks_approach_the_direct_operators():
    """The Calderon blocks (full.jl:201-207) equal the directly assembled K' and
    H up to the discretisation error, measured on smooth traces of a
    non-spherical body (S and D commute on a sphere).  A transposed product
    S D gives twice the K' error here, and a sign error in D D - I/4 an H
    error above 1."""
    m = O.Mesh(*meshes.spheroid(1.5, 1.0, 3))
    k = 1.0
    kp, hh = O.calderon_blocks(m, k)
    kd, hd = m.assemble(O.OP_KP, k), m.assemble(O.OP_H, k)
    p, q = m.incident(k, O.direction(0.7, 0.3))
    assert np.linalg.norm((kp - kd) @ q) / np.linalg.norm(kd @ q) < 0.07
    assert np.linalg.norm((hh - hd) @ p) / np.linalg.norm(hd @ p) < 0.1


def test_calderon_rule_follows_the_collocation_radius():
    """Mueller only, max(k, k_int) * radius <= pi/2 (full.jl:319-320)."""
    m = O.Mesh(*meshes.sphere(1.0, 2))
    r = O.collocation_radius(m)
    assert 0.97 < r < 1.0
    assert O.calderon_applies(m, 1.5, 1.0, "muller") == (1.5 * r <= math.pi / 2)
    assert O.calderon_applies(m, 1.0, 0.5, "muller") is False  # k_int = 2 decides
    assert O.calderon_applies(m, 0.1, 1.0, "cbie") is False


SPHERE_FLUID_GOLDENS = [
    # test/extended/bem/sphere/full3d.jl:209-233: (g, h, k) and amplitudes at -x, +x, +y
    (0.0012, 0.23, 1.0, (0.08627220103266989 + 0.5718263756162451j, -1.169649453390533 + 0.8439868202224056j,
                         -0.4123286201575636 + 0.7056764613690221j)),
    (1000.0, 2.0, 1.0, (-0.4684232462649632 + 0.0118836351691168j, 0.1742206481610343 + 0.08026589273243463j,
                        -0.2384449751929127 + 0.04496309639283315j)),
]


@pytest.mark.parametrize("g,h,k,refs", SPHERE_FLUID_GOLDENS)
def test_reference_sphere_fluid_goldens(g, h, k, refs):
    nodes, tri = meshes.sphere(1.0, 3)
    m = O.Mesh(nodes, tri)
    p, q, _ = O.solve_fluid(m, g, h, k, beta=0.0)
    f = m.far_field(k, np.array([[-1.0, 0, 0], [1.0, 0, 0], [0, 1.0, 0]]), p, q)
    for fi, r in zip(f, refs):
        assert abs(20 * math.log10(abs(fi)) - 20 * math.log10(abs(r))) < 0.1
        assert abs(fi - r) / abs(r) < 0.01


@pytest.mark.xfail(reason="resonant gas bubble at ka=0.0138 (|f|~72): with the Calderon blocks the "
                          "flat-panel Mueller solution converges toward the golden (34 % at L3, 9.4 % "
                          "and -0.035 dB at L4) but needs finer meshes than the oracle runs for 1 %",
                   strict=True)
def test_reference_gas_bubble_resonance_golden():
    nodes, tri = meshes.sphere(1.0, 3)
    m = O.Mesh(nodes, tri)
    p, q, _ = O.solve_fluid(m, 0.0012, 0.23, 0.0138, beta=0.0)
    f = m.far_field(0.0138, np.array([-1.0, 0, 0]), p, q)[0]
    ref = -0.2603525438230707 + 72.4628313282518j
    assert abs(f - ref) / abs(ref) < 0.01


@pytest.mark.slow
def test_reference_spheroid_rigid_golden_L4():
    """test/extended/bem/spheroid/full3d.jl:10-24 (k=2, beta=pi/3).  At 5120
    flat panels the amplitude error is ~1.07 %, so this CPU check relaxes the
    reference's 1 % bar to 1.2 % (TS stays inside 0.1 dB); the GPU test
    test_reference_rigid_spheroid_golden_n20480 (tests/test_gpu_full_size.py)
    holds the full 1 % bar at 20480 panels (0.11-0.57 % measured)."""
    nodes, tri = meshes.spheroid(1.5, 1.0, 4)
    m = O.Mesh(nodes, tri)
    p, q, _ = O.solve_boundary(m, "rigid", 2.0, beta=math.pi / 3)
    refs = (([-0.5, -math.sqrt(3) / 2, 0.0], 0.277466754734215 + 0.478594831065819j),
            ([0.5, math.sqrt(3) / 2, 0.0], 0.544249737180197 + 0.633566137972720j),
            ([0.0, 0.0, 1.0], -0.604161429137314 + 0.394521828342496j))
    for d, r in refs:
        f = m.far_field(2.0, np.array(d), p, q)[0]
        assert abs(20 * math.log10(abs(f)) - 20 * math.log10(abs(r))) < 0.1
        assert abs(f - r) / abs(r) < 0.012


def test_tetrahedron_fixture_solves():
    """test/basic/bem/formulations.jl:228-251: the 4-node supplied mesh."""
    nodes, tri = meshes.tetrahedron()
    assert O.validate(nodes, tri)[0] == 0
    m = O.Mesh(nodes, tri)
    for bc in ("rigid", "soft"):
        p, q, _ = O.solve_boundary(m, bc, 0.3, formulation="cbie")
        assert np.isfinite(O.target_strength(m.far_field(0.3, np.array([-1.0, 0, 0]), p, q)[0]))
    for k, form in ((0.3, "muller"), (0.3, "cbie"), (4.0, "muller")):
        p, q, _ = O.solve_fluid(m, 1.2, 1.1, k, formulation=form)
        assert np.isfinite(O.target_strength(m.far_field(k, -O.direction(0.0), p, q)[0]))


@pytest.mark.parametrize("g,h", [(1.5, 1.3), (0.5, 0.8)])
def test_fluid_cbie_and_muller_against_modal(g, h):
    """Both fluid formulations at strong contrast.  (At the weak 38 kHz sphere
    contrast CBIE's right-hand side is a difference of nearly cancelling
    interior operators and carries ~0.6 dB of O(h) error; Mueller is used
    there, as the reference's default.)"""
    nodes, tri = meshes.sphere(A1, 3)
    m = O.Mesh(nodes, tri)
    fm = modal.form_function(("fluid", g, h), K1, A1)
    for form in ("muller", "cbie"):
        p, q, _ = O.solve_fluid(m, g, h, K1, formulation=form)
        f = m.far_field(K1, -O.direction(0.0), p, q)[0]
        assert abs(O.target_strength(f) - modal.target_strength(fm)) < 0.1


def test_rigid_cbie_and_burton_miller_agree_off_resonance():
    nodes, tri = meshes.sphere(A1, 3)
    m = O.Mesh(nodes, tri)
    f = []
    for form in ("burton_miller", "cbie"):
        p, q, _ = O.solve_boundary(m, "rigid", K1, formulation=form)
        f.append(m.far_field(K1, -O.direction(0.0), p, q)[0])
    assert abs(f[0] - f[1]) / abs(f[0]) < 0.02


# ---------------------------------------------------------------------------
# mixed refinement semantics (test/basic/bem/formulations.jl:411-467)
# ---------------------------------------------------------------------------
def _fluid_small():
    nodes, tri = meshes.sphere(A1, 2)
    return O.Mesh(nodes, tri)


def test_mixed_refinement_accepts_and_matches_double():
    m = _fluid_small()
    pd, qd, _ = O.solve_fluid(m, *FLUID1, K1)
    pm, qm, d = O.solve_fluid(m, *FLUID1, K1, precision="mixed")
    assert d["factor_precision"] == "single"
    assert d["normwise_backward_error"] <= 1e-13 and d["componentwise_backward_error"] <= 1e-13
    fd = m.far_field(K1, -O.direction(0.0), pd, qd)[0]
    fm = m.far_field(K1, -O.direction(0.0), pm, qm)[0]
    assert abs(fm - fd) / abs(fd) < 1e-8


def test_mixed_condition_guard_forces_promotion():
    m = _fluid_small()
    _, _, d = O.solve_fluid(m, *FLUID1, K1, precision="mixed", refinement={"condition_guard": 1.0})
    assert d["factor_precision"] == "double" and d["refinement_reason"] == "condition"


def test_mixed_maxiter_zero_forces_iteration_limit():
    m = _fluid_small()
    _, _, d = O.solve_fluid(m, *FLUID1, K1, precision="mixed", refinement={"maxiter": 0})
    assert d["factor_precision"] == "double" and d["refinement_reason"] == "iteration_limit"


def test_refinement_option_validation():
    with pytest.raises(ValueError):
        O.refinement_options({"bogus": 1})
    with pytest.raises(ValueError):
        O.refinement_options({"tolerance": 2.0})
    with pytest.raises(ValueError):
        O.refinement_options({"maxiter": -1})


def test_equilibration_rule():
    A = np.array([[2.0, 0.5], [0.0, 0.0]], dtype=np.complex128)
    rows, cols = O.equilibrate(A)
    assert rows.tolist() == [2.0, 1.0]
    assert cols.tolist() == [1.0, 0.25]


def test_argument_errors():
    m = _fluid_small()
    for k in (0.0, -1.0, math.inf, math.nan):
        with pytest.raises(ValueError):
            O.solve_boundary(m, "rigid", k)
    for g, h in ((0.0, 1.0), (1.0, -1.0), (math.inf, 1.0), (1.0, math.nan)):
        with pytest.raises(ValueError):
            O.solve_fluid(m, g, h, 1.0)


@pytest.mark.parametrize("kind", [O.RIGID_BM, O.SOFT_BM, O.RIGID_CBIE, O.SOFT_CBIE, O.FLUID_MULLER,
                                  O.FLUID_CBIE, O.OP_S, O.OP_D, O.OP_KP, O.OP_H])
def test_assemble_rows_equal_full_assembly(kind):
    """The row sampler used by the large-surface checks returns exactly the
    rows of the full assembly (self, near and regular pairs, both fluid
    block rows): a swapped block or an offset row index fails."""
    nodes, tri = meshes.spheroid(0.02, 0.01, 2)
    m = O.Mesh(nodes, tri)
    A = m.assemble(kind, 300.0, 1.2, 1.1)
    rows = np.array([0, 7, m.n - 1, A.shape[0] - 1, A.shape[0] // 2 + 3])
    R = m.assemble_rows(kind, 300.0, rows, 1.2, 1.1)
    assert np.array_equal(R, A[rows, :])
    with pytest.raises(ValueError):
        m.assemble_rows(kind, 300.0, np.array([A.shape[0]]))


# ---------------------------------------------------------------------------
# pressure at points (src/fields/pressure.jl:404-472)
# ---------------------------------------------------------------------------
def _dirs(n, seed):
    v = np.random.default_rng(seed).standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def test_modal_near_field_far_limit_and_plane_wave_expansion():
    """The modal near field tends to f(theta) e^{ikr}/r (the reference's form
    function) and the series behind the interior coefficients reproduces the
    plane wave: two closed forms the near-field restatement must meet."""
    k, a = K1, A1
    for bc in (("rigid",), ("soft",), ("fluid",) + FLUID1):
        for th in (0.3, 1.7, math.pi):
            r = 2.0e6  # the next term of h_m's expansion is m(m+1)/(2kr) ~ 1e-8 here
            ps = modal.near_field(bc, k, a, r, math.cos(th))
            f = modal.form_function(bc, k, a, angle=th)
            assert abs(ps * r * np.exp(-1j * k * r) - f) < 1e-6 * abs(f)
    for kr, c in ((0.7, 0.2), (3.1, -0.6), (9.0, 0.95)):
        s = sum((2 * m + 1) * (1j ** m) * modal.js(m, kr) * modal.legendre_p(m, c) for m in range(60))
        assert abs(s - np.exp(1j * kr * c)) < 1e-12


@pytest.mark.parametrize("bc", ["rigid", "soft", "fluid"])
def test_oracle_pressure_matches_modal_near_field(bc):
    """Sphere a = 1 cm at 38 kHz (ka = 1.62), 5120 panels: the oracle's
    field from its own BEM traces against the exact series outside (r =
    1.5a .. 3a) and, for the fluid body, inside (r = 0.3a .. 0.6a), at the
    discretisation's accuracy (the amplitude errs ~1 % here)."""
    nodes, tri = meshes.sphere(A1, 4)
    m = O.Mesh(nodes, tri)
    if bc == "fluid":
        p, q, _ = O.solve_fluid(m, *FLUID1, K1, beta=0.0)
        mbc, fluid = ("fluid",) + FLUID1, FLUID1
    else:
        p, q, _ = O.solve_boundary(m, bc, K1, beta=0.0)
        mbc, fluid = (bc,), None
    dirs = _dirs(12, 3)
    radii = np.array([1.5, 2.0, 3.0, 1.5, 2.0, 3.0, 1.5, 2.0, 3.0, 2.5, 1.8, 2.2]) * A1
    pts = dirs * radii[:, None]
    vals, reg, wind = m.pressure(K1, pts, p, q, field="scattered", fluid=fluid)
    assert np.all(reg == 0) and np.all(np.abs(wind) < 1e-9)
    ref = np.array([modal.near_field(mbc, K1, A1, r, d[0]) for r, d in zip(radii, dirs)])
    assert np.max(np.abs(vals - ref)) < 0.02 * np.max(np.abs(ref))
    tot, _, _ = m.pressure(K1, pts, p, q, field="total", fluid=fluid)
    assert np.allclose(tot - vals, np.exp(1j * K1 * pts[:, 0]), atol=1e-12)
    if bc == "fluid":
        ri = np.array([0.3, 0.45, 0.6, 0.5]) * A1
        pin = _dirs(4, 5) * ri[:, None]
        vi, regi, wi = m.pressure(K1, pin, p, q, field="interior", fluid=fluid)
        assert np.all(regi == 1) and np.all(np.abs(wi - 1) < 1e-9)
        refi = np.array([modal.near_field(mbc, K1, A1, r, d[0]) for r, d in zip(ri, _dirs(4, 5))])
        assert np.max(np.abs(vi - refi)) < 0.02 * np.max(np.abs(refi))
    else:
        _, regi, _ = m.pressure(K1, np.array([[0.0, 0.0, 0.0]]), p, q, field="total")
        assert regi[0] == 2


def test_oracle_pressure_far_limit_matches_far_field():
    """At 1000 m the field of the surface traces equals the oracle's far-field
    amplitude times e^{ikR}/R (the representation formula and far_field are
    one integral apart): a sign or factor error in either fails."""
    nodes, tri = meshes.spheroid(0.02, 0.01, 2)
    m = O.Mesh(nodes, tri)
    k = 300.0
    p, q, _ = O.solve_fluid(m, 1.2, 1.1, k, beta=0.7)
    dirs = _dirs(5, 9)
    R = 1000.0
    vals, reg, _ = m.pressure(k, dirs * R, p, q, beta=0.7, field="scattered", fluid=(1.2, 1.1))
    f = m.far_field(k, dirs, p, q)
    assert np.all(reg == 0)
    assert np.max(np.abs(vals * R * np.exp(-1j * k * R) - f)) < 1e-4 * np.max(np.abs(f))


def test_oracle_pressure_matched_medium_and_winding():
    """g = h = 1: no scattered field outside, and the interior field is the
    incident wave; the winding number is 1 inside and 0 outside."""
    nodes, tri = meshes.sphere(1.0, 3)
    m = O.Mesh(nodes, tri)
    k = 2.0
    p, q, _ = O.solve_fluid(m, 1.0, 1.0, k, beta=0.4)
    out_pts = _dirs(6, 1) * 1.7
    in_pts = _dirs(6, 2) * 0.5
    vs, rs, ws = m.pressure(k, out_pts, p, q, beta=0.4, field="scattered", fluid=(1.0, 1.0))
    assert np.all(rs == 0) and np.max(np.abs(vs)) < 1e-10
    vi, ri, wi = m.pressure(k, in_pts, p, q, beta=0.4, field="total", fluid=(1.0, 1.0))
    d = O.direction(0.4)
    assert np.all(ri == 1) and np.all(np.abs(wi - 1.0) < 1e-9)
    assert np.max(np.abs(vi - np.exp(1j * k * in_pts @ d))) < 5e-3


# ---------------------------------------------------------------------------
# reduced-basis frequency sweep (src/postprocessing/sweeps.jl:113-306)
# ---------------------------------------------------------------------------
def _direct_cbie(m, g, h, k, beta, alpha):
    from oracle import reduced as RB
    A, b = RB.equations(m, g, h, k, beta, alpha)
    x = RB.full_solve(A, b)
    return m.far_field(k, -O.direction(beta, alpha), x[:m.n], x[m.n:])[0]


def test_reduced_sweep_reference_scenario():
    """test/basic/sampling.jl:117-139 on the flat-panel unit sphere: k = 0.6,
    0.7, 1.2, 0.7 with training 0.6, 0.8, g = 1.05, h = 1.02, incidence
    (0.7, 0.3): basis rank 2, the training frequency reduces, at least one
    query falls back, every returned residual <= 1e-10 and the amplitudes
    equal direct CBIE solves to 1e-8."""
    from oracle import reduced as RB
    nodes, tri = meshes.sphere(1.0, 2)
    m = O.Mesh(nodes, tri)
    ks, kt = [0.6, 0.7, 1.2, 0.7], [0.6, 0.8]
    amps, rep = RB.reduced_sweep(m, 1.05, 1.02, ks, kt, 0.7, 0.3)
    direct = np.array([_direct_cbie(m, 1.05, 1.02, k, 0.7, 0.3) for k in ks])
    assert rep["basis_rank"] == 2
    assert rep["queries"][0]["reduced"]
    assert any(not q["reduced"] for q in rep["queries"])
    assert all(max(q["residual"].values()) <= 1e-10 for q in rep["queries"])
    assert np.max(np.abs(amps - direct) / np.abs(direct)) < 1e-8
    # empty storage: every query is an independently validated full solve
    amps0, rep0 = RB.reduced_sweep(m, 1.05, 1.02, ks, kt, 0.7, 0.3, maxbytes=0)
    assert rep0["basis_rank"] == 0
    assert all(not q["reduced"] and q["reason"] == "empty" for q in rep0["queries"])
    assert np.max(np.abs(amps0 - direct) / np.abs(direct)) < 1e-12


def test_reduced_snapshot_space_and_least_squares():
    """The snapshot space keeps orthonormal columns and rejects a dependent
    snapshot; a right-hand side whose solution lies in the basis is solved
    by the reduced path exactly (rank = basis size, residual ~ eps)."""
    from oracle import reduced as RB
    rng = np.random.default_rng(4)
    n = 40
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 8 * np.eye(n)
    V = []
    cap = RB.capacity(n, 16, 64 * 1024 ** 2)
    xs = [rng.standard_normal(n) + 1j * rng.standard_normal(n) for _ in range(3)]
    for x in xs:
        assert RB.add_snapshot(V, cap, x)
    assert not RB.add_snapshot(V, cap, 2 * xs[0] - 1j * xs[2])
    B = np.stack(V, axis=1)
    assert np.allclose(B.conj().T @ B, np.eye(3), atol=1e-13)
    xt = B @ np.array([1.0, -2j, 0.5])
    res = RB.solve_reduced(A, A @ xt, B)
    assert res["reduced"] and res["rank"] == 3
    assert np.linalg.norm(res["x"] - xt) < 1e-12 * np.linalg.norm(xt)
    assert RB.capacity(n, 16, 0) == 0 and RB.capacity(n, 2, 10 ** 9) == 2
