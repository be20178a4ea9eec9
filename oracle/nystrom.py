"""CPU FP64 restatement of the reference's full-3D Nystrom discretisation with
density-interpolation corrections (TEST INFRASTRUCTURE ONLY).

The checker for the GPU library's curved-element Nystrom path: tests/ call
it; the product path never imports it.

What the reference does (its arithmetic lives in Inti.jl, which is not in
the tree):
  * `mesh(nodes, triangles; qorder)` takes 3-, 6- or 10-node triangles in
    Gmsh order (src/geometry/surface.jl:271-298) and builds an
    `Inti.Quadrature` of order `qorder` on each (curved) element
    (surface.jl:331); the unknowns are the quadrature nodes.
  * the layer operators are Nystrom matrices, w_j K(x_i, y_j) with the
    coincident entry zero, corrected by density interpolation
    (`correction=(method=:dim,)`, src/solvers/bem/full.jl:89-97):
      - the reference's own regular-wave correction
        (src/solvers/bem/quadrature/fluid.jl:11-163): regular solid
        harmonics times the normalised spherical Bessel series as the
        interpolation basis, the Green-identity defect
        multiplier * trace(x) + S flux - D pressure, and per element a
        row-scaled minimum-norm SVD fit with a 1e-13 relative cut;
      - Inti's `bdim_correction` (called at quadrature/fluid.jl:107-109 and
        by `Inti.single_double_layer`) restated from its published method
        (Faria, Perez-Arancibia and Bonnet 2021, doi:10.1016/j.cma.2021.113703):
        point sources G(., z_p) on a sphere about the quadrature's bounding
        box as the basis, the same defect and the same per-element fit
        (parity with Inti itself is unpinned: Inti is absent);
  * the systems, equilibration, LU, forcing and far field are those of the
    panel path (src/solvers/bem/full.jl:103-121,289-347,509-519), which
    `oracle.py` restates; `NystromMesh` exposes the interface those drivers
    use.

Quadrature rules: fully symmetric rules on the reference triangle
(0,0),(1,0),(0,1) of Dunavant (1985, IJNME 21:1129), degree >= qorder; the
6-node degree-4 rule (qorder 3, 4) is the unique fully symmetric one with
positive weights.  Node order inside an element is the table order below;
elements are consecutive (node e*nq + q), as Inti numbers a quadrature.
"""
import math

import numpy as np

from . import oracle as O

# ---------------------------------------------------------------------------
# quadrature rules: orbits of barycentric points (S3 centroid, S21 (a,a,1-2a),
# S111 (a,b,1-a-b) and permutations) with weights summing to 1
# ---------------------------------------------------------------------------
_RULES = {
    1: [("S3", (), 1.0)],
    2: [("S21", (1.0 / 6.0,), 1.0 / 3.0)],
    4: [("S21", (0.445948490915965,), 0.223381589678011),
        ("S21", (0.091576213509771,), 0.109951743655322)],
    6: [("S21", (0.249286745170910,), 0.116786275726379),
        ("S21", (0.063089014491502,), 0.050844906370207),
        ("S111", (0.053145049844817, 0.310352451033784), 0.082851075618374)],
    8: [("S3", (), 0.144315607677787),
        ("S21", (0.459292588292723,), 0.095091634267285),
        ("S21", (0.170569307751760,), 0.103217370534718),
        ("S21", (0.050547228317031,), 0.032458497623198),
        ("S111", (0.008394777409958, 0.263112829634638), 0.027230314174435)],
}
# qorder -> tabulated degree
_DEGREE = {1: 1, 2: 2, 3: 4, 4: 4, 5: 5, 6: 6, 7: 8, 8: 8}


def _radon7():
    s15 = math.sqrt(15.0)
    a, b = (6.0 - s15) / 21.0, (6.0 + s15) / 21.0
    wa, wb = (155.0 - s15) / 1200.0, (155.0 + s15) / 1200.0
    return [("S3", (), 9.0 / 40.0), ("S21", (a,), wa), ("S21", (b,), wb)]


def triangle_rule(qorder):
    """(bary (nq, 3), weights (nq,)) exact to degree >= qorder on the reference
    triangle; the weights sum to its area 1/2."""
    if qorder not in _DEGREE:
        raise ValueError("qorder must be between 1 and 8")
    deg = _DEGREE[qorder]
    orbits = _radon7() if deg == 5 else _RULES[deg]
    pts, wts = [], []
    for kind, par, w in orbits:
        if kind == "S3":
            pts.append((1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0)); wts.append(w)
        elif kind == "S21":
            a = par[0]
            c = 1.0 - 2.0 * a
            for p in ((a, a, c), (a, c, a), (c, a, a)):
                pts.append(p); wts.append(w)
        else:
            a, b = par
            c = 1.0 - a - b
            for p in ((a, b, c), (a, c, b), (b, a, c), (b, c, a), (c, a, b), (c, b, a)):
                pts.append(p); wts.append(w)
    return np.array(pts), 0.5 * np.array(wts)


# ---------------------------------------------------------------------------
# Lagrange shape functions in Gmsh node order (vertices, edge nodes along
# 0->1, 1->2, 2->0, then the interior node)
# ---------------------------------------------------------------------------
def _shape_lam(order, l0, l1, l2):
    """Values and derivatives with respect to (l0, l1, l2) treated as
    independent; returns N (npt, m) and dN (3, npt, m)."""
    one = np.ones_like(l0)
    zero = np.zeros_like(l0)
    if order == 1:
        N = [l0, l1, l2]
        dN = [[one, zero, zero], [zero, one, zero], [zero, zero, one]]
    elif order == 2:
        N = [l0 * (2 * l0 - 1), l1 * (2 * l1 - 1), l2 * (2 * l2 - 1), 4 * l0 * l1, 4 * l1 * l2, 4 * l2 * l0]
        dN = [[4 * l0 - 1, zero, zero], [zero, 4 * l1 - 1, zero], [zero, zero, 4 * l2 - 1],
              [4 * l1, 4 * l0, zero], [zero, 4 * l2, 4 * l1], [4 * l2, zero, 4 * l0]]
    elif order == 3:
        def v(l):
            return 0.5 * l * (3 * l - 1) * (3 * l - 2), 0.5 * (27 * l * l - 18 * l + 2)

        def e(a, b):  # node nearer a on edge a-b: 9/2 a b (3a - 1)
            return 4.5 * a * b * (3 * a - 1), 4.5 * b * (6 * a - 1), 4.5 * a * (3 * a - 1)
        N, dN = [], []
        for idx, l in enumerate((l0, l1, l2)):
            val, der = v(l)
            N.append(val)
            d = [zero, zero, zero]
            d[idx] = der
            dN.append(d)
        for (ia, ib) in ((0, 1), (1, 0), (1, 2), (2, 1), (2, 0), (0, 2)):
            L = (l0, l1, l2)
            val, da, db = e(L[ia], L[ib])
            N.append(val)
            d = [zero, zero, zero]
            d[ia] = da
            d[ib] = db
            dN.append(d)
        N.append(27 * l0 * l1 * l2)
        dN.append([27 * l1 * l2, 27 * l0 * l2, 27 * l0 * l1])
    else:
        raise ValueError("mesh_order must be 1, 2 or 3")
    N = np.array(N)
    dN = np.array([[np.broadcast_to(dN[a][c], l0.shape) for a in range(len(dN))] for c in range(3)])
    return N, dN


def shape(order, bary):
    """N (m, npt), dN/du (m, npt), dN/dv (m, npt) at barycentric points,
    with u = l1, v = l2 and l0 = 1 - u - v."""
    l0, l1, l2 = bary[:, 0], bary[:, 1], bary[:, 2]
    N, dN = _shape_lam(order, l0, l1, l2)
    Nu = dN[1] - dN[0]
    Nv = dN[2] - dN[0]
    return N.T, Nu.T, Nv.T


ORDER_OF = {3: 1, 6: 2, 10: 3}


class Quadrature:
    """Nystrom quadrature of a curved triangle mesh: nodes 3 x Nv, elements
    npt x N (0-based, Gmsh order)."""

    def __init__(self, nodes, tri0, qorder=4):
        nodes = np.asarray(nodes, dtype=np.float64)
        tri0 = np.asarray(tri0, dtype=np.int64)
        npt = tri0.shape[0]
        if npt not in ORDER_OF:
            raise ValueError("triangles must have 3, 6 or 10 rows in Gmsh ordering")
        self.order = ORDER_OF[npt]
        self.qorder = qorder
        bary, w = triangle_rule(qorder)
        N, Nu, Nv = shape(self.order, bary)
        P = nodes.T[tri0.T]                              # (N_el, npt, 3)
        X = np.einsum("qa,ead->eqd", N, P)
        Xu = np.einsum("qa,ead->eqd", Nu, P)
        Xv = np.einsum("qa,ead->eqd", Nv, P)
        cr = np.cross(Xu, Xv)
        J = np.linalg.norm(cr, axis=2)
        self.nel = tri0.shape[1]
        self.nq = len(w)
        self.coords = X.reshape(-1, 3)
        self.normals = (cr / J[:, :, None]).reshape(-1, 3)
        self.weights = (w[None, :] * J).reshape(-1)
        self.n = self.coords.shape[0]
        # interface of the oracle drivers
        self.centroids = self.coords

    # -- operators ---------------------------------------------------------
    def layers(self, k, derivative=False):
        """Uncorrected Nystrom matrices (S, D), or (K', H) with derivative:
        w_j K(x_i, y_j), zero where x_i = y_j."""
        return layer_matrices(self.coords, self.normals, self.weights, k, derivative)

    def incident(self, k, d):
        """Plane-wave traces (src/solvers/bem/incident.jl:64-76)."""
        ph = np.exp(1j * k * (self.coords @ d))
        return ph, 1j * k * (self.normals @ d) * ph

    def far_field(self, k, dirs, p, q):
        """src/solvers/bem/full.jl:509-519 over the quadrature nodes."""
        dirs = np.atleast_2d(dirs)
        out = []
        for xh in dirs:
            e = np.exp(-1j * k * (self.coords @ xh))
            out.append(-np.sum(self.weights * (1j * k * (self.normals @ xh) * p + q) * e) / (4 * math.pi))
        return np.array(out)

    def assemble(self, kind, k, g=1.0, h=1.0):
        """Systems of the oracle drivers (oracle.solve_boundary) with the
        point-source density interpolation of every layer operator
        (src/solvers/bem/full.jl:68-128)."""
        n = self.n
        I = np.eye(n)
        if kind in (O.OP_S, O.OP_D):
            S, D = corrected_layers(self, k, derivative=False, regular=False)
            return S if kind == O.OP_S else D
        a = 1j / k
        if kind == O.RIGID_CBIE:
            _, D = corrected_layers(self, k, False, False)
            return 0.5 * I - D
        if kind == O.SOFT_CBIE:
            S, _ = corrected_layers(self, k, False, False)
            return S
        S, D = corrected_layers(self, k, False, False)
        K, H = corrected_layers(self, k, True, False)
        if kind == O.RIGID_BM:
            return 0.5 * I - D - a * H
        if kind == O.SOFT_BM:
            return S + a * (0.5 * I + K)
        raise ValueError("kind not supported on Nystrom quadratures")


def layer_matrices(X, N, W, k, derivative=False):
    r = X[:, None, :] - X[None, :, :]
    R = np.linalg.norm(r, axis=2)
    same = R == 0
    Rs = np.where(same, 1.0, R)
    G = np.exp(1j * k * Rs) / (4 * math.pi * Rs) * W[None, :]
    if not derivative:
        b = np.einsum("ijd,jd->ij", r, N) / Rs
        S = G
        D = G * (1.0 / Rs - 1j * k) * b
        S = np.where(same, 0, S)
        D = np.where(same, 0, D)
        return S, D
    a = np.einsum("ijd,id->ij", r, N) / Rs
    b = np.einsum("ijd,jd->ij", r, N) / Rs
    c = N @ N.T
    K = G * (1j * k - 1.0 / Rs) * a
    ab = a * b
    H = G * (k * k * ab + (c - 3 * ab) / Rs ** 2 + 1j * k * (3 * ab - c) / Rs)
    return np.where(same, 0, K), np.where(same, 0, H)


# ---------------------------------------------------------------------------
# interpolation bases
# ---------------------------------------------------------------------------
OVERSAMPLE, RADIUS_MULTIPLIER = 3.0, 1.5


def fibonacci_sphere(ns, radius, center):
    """Golden-angle spiral of ns points on a sphere (a Fibonacci lattice)."""
    golden = math.pi * (3.0 - math.sqrt(5.0))
    pts = np.empty((ns, 3))
    for i in range(ns):
        y = 1.0 - 2.0 * i / (ns - 1) if ns > 1 else 0.0
        rho = math.sqrt(max(0.0, 1.0 - y * y))
        th = golden * i
        pts[i] = (math.cos(th) * rho * radius + center[0], y * radius + center[1],
                  math.sin(th) * rho * radius + center[2])
    return pts


def source_points(quad):
    """Point sources of the general correction: ceil(3 nq) points on the
    sphere of radius 1.5 x half the diagonal of the quadrature's bounding box,
    about the box centre."""
    lo, hi = quad.coords.min(axis=0), quad.coords.max(axis=0)
    xc = 0.5 * (lo + hi)
    R = RADIUS_MULTIPLIER * 0.5 * float(np.linalg.norm(hi - lo))
    ns = int(math.ceil(OVERSAMPLE * quad.nq))
    return fibonacci_sphere(ns, R, xc)


def point_traces(X, N, zs, k):
    """G(x, z_p) and its derivative along the normal at x (n x P)."""
    r = X[:, None, :] - zs[None, :, :]
    R = np.linalg.norm(r, axis=2)
    G = np.exp(1j * k * R) / (4 * math.pi * R)
    a = np.einsum("ipd,id->ip", r, N) / R
    return G, G * (1j * k - 1.0 / R) * a


def quadrature_size(quad):
    """(center, radius, rms) of src/solvers/bem/quadrature/fluid.jl:1-7."""
    c = quad.coords.mean(axis=0)
    d2 = np.sum((quad.coords - c) ** 2, axis=1)
    radius = float(np.sqrt(d2.max()))
    rms = float(np.sqrt(np.sum(quad.weights * d2) / np.sum(quad.weights)))
    return c, radius, rms


def regular_range(k, quad):
    """src/solvers/bem/quadrature/fluid.jl:9."""
    _, radius, rms = quadrature_size(quad)
    return k * radius <= 1 and 2 * k * rms <= 1


def regular_degree(nq):
    """src/solvers/bem/quadrature/fluid.jl:112-113."""
    return int(math.floor(math.sqrt(2 * nq))) - 1


def regular_traces(X, N, center, radius, k, degree):
    """Pressure and normal traces of the regular solid harmonics times the
    normalised spherical Bessel series (src/solvers/bem/quadrature/fluid.jl:19-61),
    columns in its order: m = 0..degree, l = m..degree, then the conjugate
    pair for m > 0."""
    v = (X - center) / radius
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    t = np.sum(v * v, axis=1)
    xy = x + 1j * y
    n = X.shape[0]
    ez = np.array([0.0, 0.0, 1.0])
    diagonal = np.ones(n, dtype=complex)
    gradient = np.zeros((n, 3), dtype=complex)
    vals, flux = [], []
    for m in range(degree + 1):
        if m > 0:
            gradient = -(2 * m - 1) * (xy[:, None] * gradient + diagonal[:, None] * np.array([1.0, 1j, 0.0]))
            diagonal = diagonal * (-(2 * m - 1) * xy)
        prev = np.zeros(n, dtype=complex)
        prev_g = np.zeros((n, 3), dtype=complex)
        cur, cur_g = diagonal, gradient
        for l in range(m, degree + 1):
            if l > m:
                nxt = ((2 * l - 1) * z * cur - (l + m - 1) * t * prev) / (l - m)
                nxt_g = ((2 * l - 1) * (z[:, None] * cur_g + cur[:, None] * ez)
                         - (l + m - 1) * (t[:, None] * prev_g + 2 * v * prev[:, None])) / (l - m)
                prev, prev_g = cur, cur_g
                cur, cur_g = nxt, nxt_g
            coefficient = 1.0
            radial = np.ones(n)
            derivative = np.zeros(n)
            for p in range(1, 21):
                coefficient *= -(k * radius) ** 2 / (2 * p * (2 * l + 2 * p + 1))
                radial = radial + coefficient * t ** p
                derivative = derivative + p * coefficient * t ** (p - 1)
            norm = math.sqrt((2 * l + 1) / (4 * math.pi) / math.prod(range(l - m + 1, l + m + 1)))
            value = norm * cur * radial
            fl = norm * np.sum(N * (radial[:, None] * cur_g + 2 * v * (cur * derivative)[:, None]), axis=1) / radius
            vals.append(value)
            flux.append(fl)
            if m > 0:
                vals.append(np.conj(value))
                flux.append(np.conj(fl))
    return np.array(vals).T, np.array(flux).T


def basis_traces(quad, k, regular):
    """(pressure, flux) of the interpolation basis at the quadrature nodes."""
    if regular:
        c, radius, _ = quadrature_size(quad)
        return regular_traces(quad.coords, quad.normals, c, radius, k, regular_degree(quad.nq))
    return point_traces(quad.coords, quad.normals, source_points(quad), k)


# ---------------------------------------------------------------------------
# the correction (src/solvers/bem/quadrature/fluid.jl:114-150)
# ---------------------------------------------------------------------------
def dim_correct(quad, A, B, pressure, flux, target_trace, multiplier=-0.5):
    """Corrects (A, B) = (S, D), or (K', H) with target_trace the flux, in
    place so that multiplier * trace + A flux - B pressure vanishes on the
    basis; the correction of target i sits on its own element's nodes."""
    nq = quad.nq
    defect = multiplier * target_trace + A @ flux - B @ pressure
    eps = np.finfo(np.float64).eps
    for e in range(quad.nel):
        cols = slice(e * nq, (e + 1) * nq)
        traces = np.vstack([pressure[cols], flux[cols]])           # 2nq x P
        scales = np.maximum(np.sqrt(np.sum(np.abs(traces) ** 2, axis=1)), eps)
        U, s, Vh = np.linalg.svd((traces / scales[:, None]).T, full_matrices=False)
        keep = s > 1e-13 * s[0]
        Vk = Vh[keep].conj().T
        weights = (Vk @ ((U[:, keep].conj().T @ defect[cols].T) / s[keep][:, None])) / scales[:, None]
        A[cols, cols] -= weights[nq:].T
        B[cols, cols] += weights[:nq].T
    return A, B


def corrected_layers(quad, k, derivative, regular):
    """Density-interpolated (S, D) or (K', H) at wavenumber k on the surface
    (target = source, multiplier -1/2; src/solvers/bem/quadrature/fluid.jl:94-97)."""
    A, B = quad.layers(k, derivative)
    pressure, flux = basis_traces(quad, k, regular)
    return dim_correct(quad, A, B, pressure, flux, flux if derivative else pressure)


def assemble_fluid(quad, g, h, k, formulation="muller"):
    """2n x 2n fluid system on the quadrature (src/solvers/bem/full.jl:289-347):
    regular-wave interpolation in both media when Mueller and max(k, k_int) is
    in the regular range, Calderon derivative blocks when
    max(k, k_int) * radius <= pi/2 (full.jl:187-207)."""
    ki = k / h
    n = quad.n
    regular = formulation == "muller" and regular_range(max(k, ki), quad)
    _, radius, _ = quadrature_size(quad)
    reconstruct = max(k, ki) * radius <= math.pi / 2
    Se, De = corrected_layers(quad, k, False, regular)
    Si, Di = corrected_layers(quad, ki, False, regular)
    I = np.eye(n)
    if formulation != "muller":
        return np.block([[0.5 * I - De, Se], [0.5 * I + Di, -g * Si]])
    if reconstruct:
        def cal(S, D):
            import scipy.linalg as sla
            lu = sla.lu_factor(S)
            return sla.lu_solve(lu, D @ S), sla.lu_solve(lu, D @ D - 0.25 * I)
        Ke, He = cal(Se, De)
        Ki, Hi = cal(Si, Di)
    else:
        Ke, He = corrected_layers(quad, k, True, regular)
        Ki, Hi = corrected_layers(quad, ki, True, regular)
    return np.block([[I - De + Di, Se - g * Si], [-He + Hi / g, I + Ke - Ki]])


def solve_fluid(quad, g, h, k, beta=0.0, alpha=0.0, formulation="muller", **kw):
    A = assemble_fluid(quad, g, h, k, formulation)
    return O.solve_fluid(quad, g, h, k, beta, alpha, formulation, A=A, **kw)


def solve_boundary(quad, boundary, k, beta=0.0, alpha=0.0, formulation="burton_miller"):
    return O.solve_boundary(quad, boundary, k, beta, alpha, formulation)


# ---------------------------------------------------------------------------
# curved icosphere meshes (the rule the product's mesher follows)
# ---------------------------------------------------------------------------
def curved_icosphere(level, axes=(1.0, 1.0, 1.0), order=2):
    """Icosphere of `level` with order-2 or order-3 Lagrange nodes projected to
    the unit sphere, then scaled per axis: nodes 3 x Nv, elements npt x N
    (0-based, Gmsh order).  Edge nodes are created per face in face order at
    the first encounter of each edge ((a,b), (b,c), (c,a)), placed from the
    lower to the higher vertex index at m = (a+b)/2 (order 2) or
    (2 lo + hi)/3 and (lo + 2 hi)/3 (order 3); order 3 then adds the face's
    interior node ((a+b)+c)/3 after its edges."""
    from . import meshes

    nodes, tri = meshes.icosphere(level)
    if order == 1:
        return nodes * np.asarray(axes, dtype=np.float64)[:, None], tri
    if order not in (2, 3):
        raise ValueError("mesh_order must be 1, 2 or 3")
    verts = [tuple(nodes[:, i]) for i in range(nodes.shape[1])]

    def proj(p):
        nn = math.sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2])
        return (p[0] / nn, p[1] / nn, p[2] / nn)

    cache = {}

    def edge(a, b):
        lo, hi = (a, b) if a < b else (b, a)
        ids = cache.get((lo, hi))
        if ids is None:
            A, B = verts[lo], verts[hi]
            if order == 2:
                pts = [proj(tuple(0.5 * (A[c] + B[c]) for c in range(3)))]
            else:
                pts = [proj(tuple((2.0 * A[c] + B[c]) / 3.0 for c in range(3))),
                       proj(tuple((A[c] + 2.0 * B[c]) / 3.0 for c in range(3)))]
            ids = []
            for p in pts:
                verts.append(p)
                ids.append(len(verts) - 1)
            cache[(lo, hi)] = ids
        return list(ids) if a < b else list(reversed(ids))

    elems = []
    for f in range(tri.shape[1]):
        a, b, c = (int(x) for x in tri[:, f])
        el = [a, b, c] + edge(a, b) + edge(b, c) + edge(c, a)
        if order == 3:
            A, B, C = verts[a], verts[b], verts[c]
            verts.append(proj(tuple(((A[i] + B[i]) + C[i]) / 3.0 for i in range(3))))
            el.append(len(verts) - 1)
        elems.append(el)
    out = np.asarray(verts, dtype=np.float64).T.copy()
    for c in range(3):
        out[c] *= axes[c]
    return out, np.asarray(elems, dtype=np.int64).T.copy()


# ---------------------------------------------------------------------------
# near-field pressure at points (src/fields/pressure.jl:404-472)
# ---------------------------------------------------------------------------
def sub_triangles(nodes, tri0):
    """Flat triangles through every node of each curved element (1, 4 or 9
    per element), outward like the element: the surface the winding number
    is taken over.  Returns (n_sub, 3, 3) vertex coordinates."""
    nodes = np.asarray(nodes, dtype=np.float64).T
    npt = tri0.shape[0]
    if npt == 3:
        local = [(0, 1, 2)]
    elif npt == 6:
        local = [(0, 3, 5), (3, 1, 4), (5, 4, 2), (3, 4, 5)]
    else:
        # lattice (i, j) of the cubic element: vertices, edge nodes, interior
        L = {(0, 0): 0, (3, 0): 1, (0, 3): 2, (1, 0): 3, (2, 0): 4, (2, 1): 5, (1, 2): 6,
             (0, 2): 7, (0, 1): 8, (1, 1): 9}
        local = []
        for j in range(3):
            for i in range(3 - j):
                local.append((L[(i, j)], L[(i + 1, j)], L[(i, j + 1)]))
                if i + j < 2:
                    local.append((L[(i + 1, j)], L[(i + 1, j + 1)], L[(i, j + 1)]))
    out = []
    for e in range(tri0.shape[1]):
        for (a, b, c) in local:
            out.append((nodes[tri0[a, e]], nodes[tri0[b, e]], nodes[tri0[c, e]]))
    return np.array(out)


def winding(subs, pts):
    """(1/4 pi) sum of the flat triangles' solid angles about each point
    (Van Oosterom and Strackee 1983)."""
    w = np.zeros(len(pts))
    for i, x in enumerate(pts):
        a, b, c = subs[:, 0] - x, subs[:, 1] - x, subs[:, 2] - x
        la, lb, lc = (np.linalg.norm(v, axis=1) for v in (a, b, c))
        num = np.einsum("ij,ij->i", a, np.cross(b, c))
        den = la * lb * lc + np.einsum("ij,ij->i", a, b) * lc + np.einsum("ij,ij->i", a, c) * lb \
            + np.einsum("ij,ij->i", b, c) * la
        w[i] = np.sum(2 * np.arctan2(num, den)) / (4 * math.pi)
    return w


def layer_values(quad, k, pts, inside, u, v, regular_ok):
    """S u - D v (inside) or D v - S u (outside) at off-surface points with
    the nearest element excluded and recovered by density interpolation
    (exclude_nearest, target_location, maxdist = 2 radius;
    src/fields/pressure.jl:441-462, src/solvers/bem/quadrature/fluid.jl:78-163).
    Points with k |x - c| <= 2 use the regular basis when `regular_ok`."""
    nq = quad.nq
    center, radius, _ = quadrature_size(quad)
    maxdist = 2 * radius
    sigma = -1.0 if inside else 0.0
    zs = source_points(quad)
    out = np.zeros(len(pts), dtype=complex)
    for i, x in enumerate(pts):
        r = x[None, :] - quad.coords
        R = np.linalg.norm(r, axis=1)
        G = np.exp(1j * k * R) / (4 * math.pi * R) * quad.weights
        Dk = G * (1.0 / R - 1j * k) * np.einsum("jd,jd->j", r, quad.normals) / R
        near = int(np.argmin(R))
        e = near // nq
        cols = slice(e * nq, (e + 1) * nq)
        corrected = R[near] <= maxdist
        if corrected:
            G = G.copy(); Dk = Dk.copy()
            G[cols] = 0
            Dk[cols] = 0
        val = (G @ u - Dk @ v) if inside else (Dk @ v - G @ u)
        if corrected:
            use_regular = regular_ok and k * np.linalg.norm(x - center) <= 2
            if use_regular:
                pres, flux = regular_traces(quad.coords, quad.normals, center, radius, k, regular_degree(nq))
                at, _ = regular_traces(x[None, :], np.zeros((1, 3)), center, radius, k, regular_degree(nq))
            else:
                pres, flux = point_traces(quad.coords, quad.normals, zs, k)
                at, _ = point_traces(x[None, :], np.zeros((1, 3)), zs, k)
            defect = sigma * at[0] + G @ flux - Dk @ pres
            traces = np.vstack([pres[cols], flux[cols]])
            scales = np.maximum(np.sqrt(np.sum(np.abs(traces) ** 2, axis=1)), np.finfo(np.float64).eps)
            U, s, Vh = np.linalg.svd((traces / scales[:, None]).T, full_matrices=False)
            keep = s > 1e-13 * s[0]
            w = (Vh[keep].conj().T @ ((U[:, keep].conj().T @ defect) / s[keep])) / scales
            dS, dD = -w[nq:], w[:nq]
            corr = (dS @ u[cols] - dD @ v[cols]) if inside else (dD @ v[cols] - dS @ u[cols])
            val = val + corr
        out[i] = val
    return out


def pressure(quad, nodes, tri0, k, pts, p, q, beta=0.0, alpha=0.0, field="total", fluid=None):
    """pressure(solution, points; field) of a Nystrom solution: (values,
    regions 0 exterior / 1 interior / 2 none, winding numbers)."""
    pts = np.atleast_2d(np.asarray(pts, dtype=np.float64))
    d = O.direction(beta, alpha)
    inc = np.exp(1j * k * (pts @ d))
    if field == "incident":
        return inc, np.zeros(len(pts), dtype=int), np.zeros(len(pts))
    w = winding(sub_triangles(nodes, tri0), pts)
    reg = np.full(len(pts), 2)
    out = np.full(len(pts), np.nan + 0j)
    for i, wi in enumerate(w):
        if field in ("total", "scattered") and wi < 0.75:
            reg[i] = 0
        elif fluid is not None and field in ("total", "interior") and wi > 0.25:
            reg[i] = 1
    ext = np.where(reg == 0)[0]
    if len(ext):
        out[ext] = layer_values(quad, k, pts[ext], False, q, p, regular_range(k, quad))
        if field == "total":
            out[ext] += inc[ext]
    inn = np.where(reg == 1)[0]
    if len(inn):
        g, h = fluid
        ki = k / h
        pinc, qinc = quad.incident(k, d)
        out[inn] = layer_values(quad, ki, pts[inn], True, g * (q + qinc), p + pinc, regular_range(ki, quad))
    return out, reg, w
