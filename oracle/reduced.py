"""Reduced-basis frequency sweep oracle (TEST INFRASTRUCTURE ONLY).

Restates the reference's FrequencyReduction module and the reduced
frequency_sweep on the oracle's dense FP64 CBIE fluid equations:
  snapshot space, bounded by capacity and bytes   src/postprocessing/sweeps.jl:115-147
  current-equation reduced solve with fallback    src/postprocessing/sweeps.jl:149-200
  CBIE equations and the full solve               src/postprocessing/sweeps.jl:209-220
  training, frozen basis, queries                 src/postprocessing/sweeps.jl:247-306
Pivoted QR is LAPACK zgeqp3 through scipy (the reference's qr(.., ColumnNorm())).
"""
import math

import numpy as np
import scipy.linalg as sla

from oracle import oracle as O

REASONS = ("accepted", "empty", "residual", "rank", "nonfinite")


def capacity(n, cap, maxbytes):
    if n <= 0:
        raise ValueError("positive state dimension required")
    if cap < 0 or maxbytes < 0:
        raise ValueError("negative storage bound")
    return min(cap, n, maxbytes // 16 // n)


def add_snapshot(V, cap, x, tolerance=1e-12):
    """V: list of orthonormal columns; returns True when x enlarged it."""
    if len(V) >= cap:
        return False
    mag = np.linalg.norm(x)
    if not np.isfinite(mag):
        raise ValueError("snapshot norm overflow")
    if mag == 0:
        return False
    v = x / mag
    for _ in range(2):
        for q in V:
            v = v - np.vdot(q, v) * q
    rem = np.linalg.norm(v)
    if not rem > tolerance:
        return False
    V.append(v / rem)
    return True


def relative_residual(A, x, b, rows):
    r = A @ x - b

    def rel(rr, bb):
        nb, nr = np.linalg.norm(bb), np.linalg.norm(rr)
        return (0.0 if nr == 0 else math.inf) if nb == 0 else nr / nb

    return {"original": rel(r, b), "scaled": rel(r / rows, b / rows)}


def full_solve(A, b):
    """_factor_fluid_system(A; condition_limit=0) + _solve_fluid_system."""
    rows, cols = O.equilibrate(A)
    lu = sla.lu_factor(A / rows[:, None] / cols[None, :], check_finite=False)
    return sla.lu_solve(lu, b / rows, check_finite=False) / cols


def solve_reduced(A, b, V, tolerance=1e-10, rank_tolerance=1e-12):
    n = A.shape[0]
    m = 0 if V is None else V.shape[1]
    rows = np.abs(A).max(axis=1)
    rows[rows == 0] = 1.0
    if not np.all(np.isfinite(rows)):
        raise ValueError("row norm overflow")
    reason = "empty" if m == 0 else "residual"
    x = np.zeros(n, dtype=np.complex128)
    rank = 0
    if np.linalg.norm(b) != 0 and m > 0:
        images = (A @ V) / rows[:, None]
        if np.all(np.isfinite(images)):
            Q, R, P = sla.qr(images, mode="economic", pivoting=True)
            d = np.abs(np.diag(R))
            rank = int(np.count_nonzero(d > rank_tolerance * d.max()))
            if rank == m:
                z = sla.solve_triangular(R, Q.conj().T @ (b / rows))
                y = np.empty(m, dtype=np.complex128)
                y[P] = z
                x = V @ y
            else:
                reason = "rank"
        else:
            reason = "nonfinite"
    projected = relative_residual(A, x, b, rows)
    if np.all(np.isfinite(x)) and max(projected.values()) <= tolerance:
        return {"x": x, "reduced": True, "reason": "accepted", "rank": rank, "projected": projected,
                "residual": projected}
    if not np.all(np.isfinite(x)):
        reason = "nonfinite"
    x = full_solve(A, b)
    residual = relative_residual(A, x, b, rows)
    if not (np.all(np.isfinite(x)) and max(residual.values()) <= tolerance):
        raise RuntimeError("full solve failed current-equation residual acceptance")
    return {"x": x, "reduced": False, "reason": reason, "rank": rank, "projected": projected, "residual": residual}


def equations(mesh, g, h, k, beta, alpha):
    A = O.assemble_fluid(mesh, g, h, k, formulation="cbie")
    p, q = mesh.incident(k, O.direction(beta, alpha))
    return A, O.fluid_forcing(A, mesh.n, p, q, "cbie")


def reduced_sweep(mesh, g, h, ks, train_ks, beta=math.pi / 2, alpha=0.0, tolerance=1e-10, rank_tolerance=1e-12,
                  cap=16, maxbytes=64 * 1024 ** 2):
    n = mesh.n
    c = capacity(2 * n, cap, maxbytes)
    V = []
    training = []
    for k in train_ks:
        A, b = equations(mesh, g, h, k, beta, alpha)
        res = solve_reduced(A, b, None, tolerance, rank_tolerance)
        kept = add_snapshot(V, c, res["x"], rank_tolerance)
        training.append({"k": k, "kept": kept, "rank": len(V), "residual": res["residual"]})
    basis = np.stack(V, axis=1) if V else None
    back = -O.direction(beta, alpha)
    amps, queries = [], []
    for k in ks:
        A, b = equations(mesh, g, h, k, beta, alpha)
        res = solve_reduced(A, b, basis, tolerance, rank_tolerance)
        x = res["x"]
        amps.append(mesh.far_field(k, back, x[:n], x[n:])[0])
        queries.append({key: res[key] for key in ("reduced", "reason", "rank", "projected", "residual")} |
                       {"k": k})
    return np.array(amps), {"training": training, "queries": queries, "basis_rank": len(V)}
