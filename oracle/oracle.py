"""CPU FP64 oracle driver (TEST INFRASTRUCTURE ONLY).

The checker for the GPU library: tests/, the smoke check and the bench's
cpu_baseline leg call it; the product path never imports it.

Assembly, incident traces, far field and mesh checks run in the C
restatement (bem_oracle.c, built by oracle/Makefile).  Dense LU is LAPACK
zgetrf/cgetrf through scipy, which is the routine the reference itself calls
(LinearAlgebra.lu! at src/solvers/bem/full.jl:404,
src/solvers/bem/fluid_system.jl:217).  The solve drivers restate:
  rigid/soft Burton-Miller RHS and outputs     src/solvers/bem/full.jl:130-184
  fluid equilibration, LU, unscaling           src/solvers/bem/full.jl:364-440
  fluid forcing and scattered traces           src/solvers/bem/full.jl:442-495
  low-frequency Calderon derivative blocks     src/solvers/bem/full.jl:187-207,319-339
  MixedRefinement (cFP32 LU + FP64 refinement) src/solvers/bem/fluid_system.jl:172-309
  residual report                              src/numerics/diagnostics.jl:9-18
"""
import ctypes
import math
import os

import numpy as np
import scipy.linalg as sla
from scipy.linalg import lapack

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

RIGID_BM, SOFT_BM, RIGID_CBIE, SOFT_CBIE, FLUID_MULLER, FLUID_CBIE = 0, 1, 2, 3, 4, 5
OP_S, OP_D, OP_KP, OP_H = 6, 7, 8, 9

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int64)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libbem_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.orc_mesh_create.restype = ctypes.c_void_p
        L.orc_mesh_create.argtypes = [ctypes.c_int64, _dp, ctypes.c_int64, _ip]
        L.orc_mesh_free.argtypes = [ctypes.c_void_p]
        L.orc_mesh_geometry.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_mesh_quadrature.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_is_near.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64]
        L.orc_pair_ops.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                   ctypes.c_double, _dp]
        L.orc_pair_ops_level.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                         ctypes.c_double, ctypes.c_int, _dp]
        L.orc_assemble.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int64]
        L.orc_assemble_rows.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_int64, _ip, _dp]
        L.orc_incident.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp, _dp, _dp]
        L.orc_pressure.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int64, _dp, _dp, _dp, _dp,
                                   ctypes.POINTER(ctypes.c_int), _dp]
        L.orc_far_field.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int64, _dp,
                                    _dp, _dp, _dp]
        L.orc_validate.argtypes = [ctypes.c_int64, _dp, ctypes.c_int64, _ip, _dp,
                                   ctypes.c_char_p, ctypes.c_int]
        L.orc_num_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp)


def direction(beta, alpha=0.0):
    """Polar angle from +x; azimuth from +y toward +z (src/solvers/bem/full.jl:15-18)."""
    return np.array([math.cos(beta), math.sin(beta) * math.cos(alpha),
                     math.sin(beta) * math.sin(alpha)])


def validate(nodes, tri0):
    nodes = np.ascontiguousarray(nodes.T, dtype=np.float64)
    tri0 = np.ascontiguousarray(tri0.T, dtype=np.int64)
    res = ctypes.c_double(0)
    buf = ctypes.create_string_buffer(256)
    code = lib().orc_validate(nodes.shape[0], _d(nodes), tri0.shape[0],
                              tri0.ctypes.data_as(_ip), ctypes.byref(res), buf, 256)
    return code, buf.value.decode(), res.value


class Mesh:
    """Flat-panel geometry of a closed triangulated surface (0-based triangles)."""

    def __init__(self, nodes, tri0):
        self.nodes = np.ascontiguousarray(np.asarray(nodes, dtype=np.float64).T)
        self.tri = np.ascontiguousarray(np.asarray(tri0, dtype=np.int64).T)
        self.n = self.tri.shape[0]
        self._h = lib().orc_mesh_create(self.nodes.shape[0], _d(self.nodes), self.n,
                                        self.tri.ctypes.data_as(_ip))
        c = np.empty((self.n, 3)); nr = np.empty((self.n, 3))
        area = np.empty(self.n); hmax = np.empty(self.n)
        lib().orc_mesh_geometry(self._h, _d(c), _d(nr), _d(area), _d(hmax))
        self.centroids, self.normals, self.area, self.hmax = c, nr, area, hmax

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.orc_mesh_free(self._h)
            self._h = None

    def quadrature(self):
        xyz = np.empty((7 * self.n, 3)); w = np.empty(7 * self.n)
        lib().orc_mesh_quadrature(self._h, _d(xyz), _d(w))
        return xyz, w

    def is_near(self, i, j):
        return bool(lib().orc_is_near(self._h, i, j))

    def pair_ops(self, i, j, k, level=None):
        out = np.empty(8)
        if level is None:
            lib().orc_pair_ops(self._h, i, j, k, _d(out))
        else:
            lib().orc_pair_ops_level(self._h, i, j, k, level, _d(out))
        return out.view(np.complex128)

    def assemble(self, kind, k, g=1.0, h=1.0):
        n = self.n * (2 if kind in (FLUID_MULLER, FLUID_CBIE) else 1)
        A = np.empty((n, n), dtype=np.complex128, order="F")
        rc = lib().orc_assemble(self._h, kind, k, g, h, A.ctypes.data_as(_dp), n)
        if rc != 0:
            raise ValueError("k must be finite and positive")
        return A

    def assemble_rows(self, kind, k, rows, g=1.0, h=1.0):
        """Rows of the matrix assemble() would return (row-major, complex)."""
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        n = self.n * (2 if kind in (FLUID_MULLER, FLUID_CBIE) else 1)
        out = np.empty((rows.size, n), dtype=np.complex128)
        rc = lib().orc_assemble_rows(self._h, kind, k, g, h, rows.size, rows.ctypes.data_as(_ip),
                                     out.ctypes.data_as(_dp))
        if rc != 0:
            raise ValueError("k must be finite and positive and rows inside the matrix")
        return out

    def incident(self, k, d):
        d = np.ascontiguousarray(d, dtype=np.float64)
        p = np.empty(self.n, dtype=np.complex128); q = np.empty(self.n, dtype=np.complex128)
        lib().orc_incident(self._h, k, _d(d), p.ctypes.data_as(_dp), q.ctypes.data_as(_dp))
        return p, q

    def pressure(self, k, points, p, q, beta=0.0, alpha=0.0, field="total", fluid=None):
        """Pressure at points (N x 3) from scattered traces p, q
        (src/fields/pressure.jl:404-472); fluid = (g, h) for a fluid body.
        Returns (values, regions, winding numbers); regions 0 exterior,
        1 interior, 2 no pressure for this field."""
        codes = {"total": 0, "scattered": 1, "incident": 2, "interior": 3}
        pts = np.ascontiguousarray(np.atleast_2d(points), dtype=np.float64)
        n = pts.shape[0]
        out = np.empty(n, dtype=np.complex128)
        reg = np.empty(n, dtype=np.int32)
        wind = np.empty(n)
        g, h = fluid if fluid is not None else (1.0, 1.0)
        d = np.ascontiguousarray(direction(beta, alpha))
        lib().orc_pressure(self._h, k, k / h, g, 1 if fluid is not None else 0, codes[field], _d(d), n, _d(pts),
                           np.ascontiguousarray(p, dtype=np.complex128).ctypes.data_as(_dp),
                           np.ascontiguousarray(q, dtype=np.complex128).ctypes.data_as(_dp),
                           out.ctypes.data_as(_dp), reg.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _d(wind))
        return out, reg, wind

    def far_field(self, k, dirs, p, q):
        dirs = np.ascontiguousarray(np.atleast_2d(dirs), dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.complex128)
        q = np.ascontiguousarray(q, dtype=np.complex128)
        out = np.empty(dirs.shape[0], dtype=np.complex128)
        lib().orc_far_field(self._h, k, dirs.shape[0], _d(dirs), p.ctypes.data_as(_dp),
                            q.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
        return out


def residual_report(residual, b):
    """src/numerics/diagnostics.jl:13-18."""
    absolute = float(np.linalg.norm(residual))
    rhs = float(np.linalg.norm(b))
    rel = (0.0 if absolute == 0 else math.inf) if rhs == 0 else absolute / rhs
    return {"absolute_residual": absolute, "relative_residual": rel}


# --------------------------------------------------------------------------
# rigid / pressure-release (dense direct; the reference's GMRES is replaced by
# LU, which is its converged limit: compression=(method=:none,), reltol -> 0)
# --------------------------------------------------------------------------
def solve_boundary(mesh, boundary, k, beta=0.0, alpha=0.0, formulation="burton_miller"):
    """Returns (p_scat, dpdn_scat, diagnostics) following full.jl:130-184."""
    if not (math.isfinite(k) and k > 0):
        raise ValueError("k must be finite and positive")
    d = direction(beta, alpha)
    p_inc, q_inc = mesh.incident(k, d)
    bm = formulation == "burton_miller"
    if boundary == "rigid":
        A = mesh.assemble(RIGID_BM if bm else RIGID_CBIE, k)
        B = mesh.assemble(SOFT_BM if bm else SOFT_CBIE, k)
        rhs = B @ q_inc            # S q + alpha (1/2 q + K' q)
    else:
        A = mesh.assemble(SOFT_BM if bm else SOFT_CBIE, k)
        B = mesh.assemble(RIGID_BM if bm else RIGID_CBIE, k)
        rhs = B @ p_inc            # -D p + 1/2 p - alpha H p
    lu = sla.lu_factor(A, check_finite=False)
    x = sla.lu_solve(lu, rhs, check_finite=False)
    diag = residual_report(A @ x - rhs, rhs)
    diag.update(method="direct", formulation=formulation, unknown_count=mesh.n)
    if boundary == "rigid":
        return x, -q_inc, diag
    return -p_inc, x, diag


# --------------------------------------------------------------------------
# fluid (src/solvers/bem/full.jl:273-495)
# --------------------------------------------------------------------------
def equilibrate(A):
    """Row max, then column max of the row-scaled matrix; zeros -> 1 (full.jl:388-394)."""
    rows = np.abs(A).max(axis=1)
    rows[rows == 0] = 1.0
    cols = (np.abs(A) / rows[:, None]).max(axis=0)
    cols[cols == 0] = 1.0
    return rows, cols


def collocation_radius(mesh):
    """Largest distance of a collocation point from their mean
    (src/solvers/bem/quadrature/fluid.jl:1-7, with the panel centroids as the
    collocation points)."""
    c = mesh.centroids.mean(axis=0)
    return float(np.max(np.linalg.norm(mesh.centroids - c, axis=1)))


def calderon_applies(mesh, k, h, formulation):
    """Mueller at max(k, k_int) * radius <= pi/2 (src/solvers/bem/full.jl:319-320)."""
    return formulation == "muller" and max(k, k / h) * collocation_radius(mesh) <= math.pi / 2


def calderon_blocks(mesh, k):
    """K' = S^-1 (D S) and H = S^-1 (D D - I/4) at wavenumber k
    (src/solvers/bem/full.jl:201-207)."""
    S = mesh.assemble(OP_S, k)
    D = mesh.assemble(OP_D, k)
    lu = sla.lu_factor(S)
    return sla.lu_solve(lu, D @ S), sla.lu_solve(lu, D @ D - 0.25 * np.eye(mesh.n))


def assemble_fluid(mesh, g, h, k, formulation="muller"):
    """The 2N x 2N fluid system (src/solvers/bem/full.jl:289-347): Mueller
    takes its derivative blocks from the Calderon identities at low frequency."""
    A = mesh.assemble(FLUID_MULLER if formulation == "muller" else FLUID_CBIE, k, g, h)
    if calderon_applies(mesh, k, h, formulation):
        n = mesh.n
        ke, he = calderon_blocks(mesh, k)
        ki, hi = calderon_blocks(mesh, k / h)
        A[n:, :n] = -he + hi / g
        A[n:, n:] = np.eye(n) + ke - ki
    return A


def fluid_forcing(A, n, p_inc, q_inc, formulation):
    if formulation == "muller":
        return np.concatenate([p_inc, q_inc])
    rhs = np.zeros(2 * n, dtype=np.complex128)
    rhs[n:] = -(A[n:, :] @ np.concatenate([p_inc, q_inc]))
    return rhs


def solve_fluid(mesh, g, h, k, beta=0.0, alpha=0.0, formulation="muller",
                precision="double", refinement=None, A=None):
    """Returns (p_scat, dpdn_scat, diagnostics)."""
    if not (math.isfinite(k) and k > 0):
        raise ValueError("k must be finite and positive")
    if not (math.isfinite(g) and g > 0 and math.isfinite(h) and h > 0):
        raise ValueError("fluid density and soundspeed contrasts must be finite and positive")
    n = mesh.n
    if A is None:
        A = assemble_fluid(mesh, g, h, k, formulation)
    p_inc, q_inc = mesh.incident(k, direction(beta, alpha))
    b = fluid_forcing(A, n, p_inc, q_inc, formulation)
    diag = {"method": "direct", "formulation": formulation, "unknown_count": 2 * n}
    if precision == "mixed":
        opts = refinement_options(refinement or {})
        F = factorize_refined(A, condition_guard=opts["condition_guard"])
        sol = solve_refined(F, b, tolerance=opts["tolerance"], maxiter=opts["maxiter"])
        x = sol["x"]
        diag.update(precision="mixed", factor_precision=sol["precision"],
                    refinement_reason=sol["reason"], refinement_iterations=sol["corrections"],
                    normwise_backward_error=sol["normwise"],
                    componentwise_backward_error=sol["componentwise"],
                    refinement_history=sol["history"],
                    single_reciprocal_condition=F.reciprocal_condition,
                    method="mixed_refinement")
        rows = F.rows
    else:
        rows, cols = equilibrate(A)
        scaled = A / rows[:, None] / cols[None, :]
        lu = sla.lu_factor(scaled, overwrite_a=True, check_finite=False)
        x = sla.lu_solve(lu, b / rows, check_finite=False) / cols
    r = A @ x - b
    diag.update(residual_report(r, b))
    sc = residual_report(r / rows, b / rows)
    diag["scaled_relative_residual"] = sc["relative_residual"]
    diag["scaled_absolute_residual"] = sc["absolute_residual"]
    p, q = x[:n], x[n:]
    if formulation == "muller":
        p, q = p - p_inc, q - q_inc
    return p, q, diag


# --------------------------------------------------------------------------
# MixedRefinement (src/solvers/bem/fluid_system.jl:172-309)
# --------------------------------------------------------------------------
def refinement_options(options):
    for key in options:
        if key not in ("tolerance", "maxiter", "condition_guard"):
            raise ValueError("unknown refinement option")
    tol = options.get("tolerance", 1e-13)
    maxiter = options.get("maxiter", 8)
    guard = options.get("condition_guard", 10 * float(np.finfo(np.float32).eps))
    if not (math.isfinite(tol) and 0 < tol < 1):
        raise ValueError("invalid refinement tolerance")
    if not (isinstance(maxiter, int) and maxiter >= 0):
        raise ValueError("maxiter must be a nonnegative integer")
    if not (math.isfinite(guard) and guard >= 0):
        raise ValueError("invalid condition guard")
    return {"tolerance": tol, "maxiter": maxiter, "condition_guard": guard}


class RefinedLU:
    def __init__(self, A, rows, cols, matrix_norm):
        self.A, self.rows, self.columns, self.matrix_norm = A, rows, cols, matrix_norm
        self.factor = None
        self.precision = "single"
        self.reason = "none"
        self.reciprocal_condition = math.nan

    def promote(self, reason):
        scaled = self.A / self.rows[:, None] / self.columns[None, :]
        self.factor = sla.lu_factor(scaled, overwrite_a=True, check_finite=False)
        self.precision = "double"
        self.reason = reason


def factorize_refined(A, condition_guard=10 * float(np.finfo(np.float32).eps)):
    n = A.shape[0]
    if A.shape[0] != A.shape[1] or n == 0:
        raise ValueError("expected a nonempty square matrix")
    if not np.all(np.isfinite(A)):
        raise ValueError("matrix must be finite")
    absA = np.abs(A)
    rows = absA.max(axis=1)
    rows[rows == 0] = 1.0
    cols = (absA / rows[:, None]).max(axis=0)
    cols[cols == 0] = 1.0
    sums = absA.sum(axis=1)
    if not np.all(np.isfinite(sums)):
        raise ValueError("matrix norm overflow")
    F = RefinedLU(A, rows, cols, float(sums.max()))
    low = (A / rows[:, None] / cols[None, :]).astype(np.complex64)
    if not np.all(np.isfinite(low)):
        F.promote("conversion")
        return F
    anorm = float(np.abs(low).sum(axis=1).max())
    lu, piv, info = lapack.cgetrf(low)
    if info != 0:
        F.promote("single_singular")
        return F
    F.factor = (lu, piv)
    rcond, _ = lapack.cgecon(lu, anorm, norm="I")
    F.reciprocal_condition = float(rcond)
    if not math.isfinite(F.reciprocal_condition) or F.reciprocal_condition <= condition_guard:
        F.promote("condition")
    return F


def correction(F, rhs):
    scaled = rhs / F.rows
    mag = float(np.abs(scaled).max())
    if not math.isfinite(mag):
        return np.full(rhs.shape, complex(math.nan, math.nan))
    if mag == 0:
        return np.zeros_like(rhs)
    if F.precision == "single":
        lu, piv = F.factor
        work = (scaled / mag).astype(np.complex64)
        val, _ = lapack.cgetrs(lu, piv, work)
        val = val.astype(np.complex128)
    else:
        val = sla.lu_solve(F.factor, scaled / mag, check_finite=False)
    return val * mag / F.columns


def backward_errors(F, x, b):
    residual = b - F.A @ x
    den = np.abs(b) + np.abs(F.A) @ np.abs(x)
    if not (np.all(np.isfinite(residual)) and np.all(np.isfinite(den))):
        return residual, math.inf, math.inf
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = np.where(den == 0, np.where(residual == 0, 0.0, math.inf),
                         np.abs(residual) / np.where(den == 0, 1.0, den))
    cw = float(ratio.max()) if ratio.size else 0.0
    scale = F.matrix_norm * float(np.abs(x).max()) + float(np.abs(b).max())
    rn = float(np.abs(residual).max())
    if not math.isfinite(scale):
        nw = math.inf
    elif scale == 0:
        nw = 0.0 if rn == 0 else math.inf
    else:
        nw = rn / scale
    return residual, nw, cw


def solve_refined(F, b, tolerance=1e-13, maxiter=8):
    if b.shape[0] != F.A.shape[0]:
        raise ValueError("right-hand side length")
    if not np.all(np.isfinite(b)):
        raise ValueError("right-hand side must be finite")
    x = correction(F, b)
    history = []
    corrections = attempts = stalls = 0
    previous = math.inf
    while True:
        residual, nw, cw = backward_errors(F, x, b)
        metric = max(nw, cw)
        history.append((F.precision, nw, cw))
        if math.isfinite(metric) and metric <= tolerance:
            return {"x": x, "accepted": True, "precision": F.precision, "reason": F.reason,
                    "corrections": corrections, "history": history,
                    "normwise": nw, "componentwise": cw}
        stalls = stalls + 1 if metric >= previous * 0.9 else 0
        if not math.isfinite(metric):
            reason = "nonfinite"
        elif attempts >= maxiter:
            reason = "iteration_limit"
        elif stalls >= 2:
            reason = "stagnation"
        else:
            reason = None
        if reason is not None:
            if F.precision == "double":
                raise RuntimeError(
                    "double-precision fallback failed backward-error acceptance: " + reason)
            F.promote(reason)
            x = correction(F, b)
            previous, attempts, stalls = math.inf, 0, 0
            continue
        delta = correction(F, residual)
        updated = x + delta
        if np.array_equal(updated, x):
            if F.precision == "double":
                raise RuntimeError("double-precision correction stagnated")
            F.promote("stagnation")
            x = correction(F, b)
            previous, attempts, stalls = math.inf, 0, 0
            continue
        x = updated
        previous = metric
        corrections += 1
        attempts += 1


def target_strength(f):
    """src/postprocessing/target_strength.jl:8."""
    return 20.0 * math.log10(abs(f))
