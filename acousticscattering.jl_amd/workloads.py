"""Synthetic workloads of the benchmark and the tests.

Served stream: Poisson arrivals of random bodies of revolution.  The meridian
is r(s) = a sin(pi s) (1 + eps cos(3 pi s)), x(s) = -aspect a cos(pi s),
s in [0, 1] (aspect = 1, eps = 0 is a sphere of radius a), with a ~ U[5, 50]
mm, aspect ~ U[1, 8], eps ~ U[0, 0.2]; the boundary is rigid, pressure-release
or fluid-filled (g, h ~ U[1.00, 1.06]) with equal odds; ka ~ U[0.5, 15] with a
the characteristic radius max_s hypot(x, r) (the reference's
`_characteristic_radius`, src/geometry/mesh.jl:6).  Each body is meshed by
rings of revolution at edge <= lambda / 6, clamped to N in [2000, 40000]
triangles.

Frequency sweep: 256 frequencies log-spaced over 1-200 kHz at c = 1500 m/s on
a fluid-filled prolate spheroid (semi-axes 0.02 x 0.01 m along x, g = 1.0357,
h = 1.0279); each frequency takes the smallest mapped icosphere (levels 3-5,
N = 1280, 5120, 20480 triangles) whose longest edge is <= lambda / 6.
"""
import math

import numpy as np

import bemhip as bh

N_MIN, N_MAX = 2000, 40000


def revolved_meridian(a, aspect, eps, ns):
    """ns meridian samples equally spaced in arc length (poles first and last);
    returns xs, rs, arc length, max radius and characteristic radius."""
    s = np.linspace(0.0, 1.0, 8001)
    x = -aspect * a * np.cos(np.pi * s)
    r = a * np.sin(np.pi * s) * (1.0 + eps * np.cos(3.0 * np.pi * s))
    arc = np.concatenate([[0.0], np.cumsum(np.hypot(np.diff(x), np.diff(r)))])
    si = np.interp(np.linspace(0.0, arc[-1], ns), arc, s)
    xs = -aspect * a * np.cos(np.pi * si)
    rs = a * np.sin(np.pi * si) * (1.0 + eps * np.cos(3.0 * np.pi * si))
    rs[0] = rs[-1] = 0.0
    return xs, rs, arc[-1], r.max(), float(np.hypot(x, r).max())


def ring_counts(arc, rmax, edge):
    """Meridian samples and azimuthal segments for target edge length `edge`."""
    ns = max(3, int(math.ceil(arc / edge)) + 1)
    nphi = max(3, int(math.ceil(2.0 * math.pi * rmax / edge)))
    return ns, nphi


def revolved_body(a, aspect, eps, k, n_min=N_MIN, n_max=N_MAX):
    """Mesh of the body at edge <= 2 pi / (6 k), with the triangle count
    clamped to [n_min, n_max]; returns (nodes 3xNv, triangles 3xN 1-based)."""
    _, _, arc, rmax, _ = revolved_meridian(a, aspect, eps, 3)
    edge = 2.0 * math.pi / k / 6.0 / math.sqrt(2.0)  # grid spacing: a cell's diagonal is the longest edge
    for _ in range(60):
        ns, nphi = ring_counts(arc, rmax, edge)
        nt = 2 * (ns - 2) * nphi
        if nt < n_min:
            edge *= math.sqrt(nt / n_min) * 0.98
        elif nt > n_max:
            edge *= math.sqrt(nt / n_max) * 1.02
        else:
            break
    xs, rs, _, _, _ = revolved_meridian(a, aspect, eps, ns)
    return bh.revolution_arrays(xs, rs, nphi)


def characteristic_radius(a, aspect, eps):
    return revolved_meridian(a, aspect, eps, 3)[4]


def serving_stream(rate, count, seed):
    """`count` requests with Poisson arrivals at `rate` per second: dicts with
    arrival time t [s], body parameters, boundary, k, and the mesh arrays."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    out = []
    for i in range(count):
        a = rng.uniform(5e-3, 50e-3)
        aspect = rng.uniform(1.0, 8.0)
        eps = rng.uniform(0.0, 0.2)
        kind = int(rng.integers(3))
        if kind == 2:
            g, h = rng.uniform(1.0, 1.06, 2)
            bc = bh.FluidFilled(float(g), float(h))
        else:
            bc = bh.Rigid() if kind == 0 else bh.PressureRelease()
        ka = rng.uniform(0.5, 15.0)
        k = ka / characteristic_radius(a, aspect, eps)
        nodes, tri = revolved_body(a, aspect, eps, k)
        out.append({"t": float(t[i]), "a": a, "aspect": aspect, "eps": eps, "boundary": bc, "k": float(k),
                     "ka": float(ka), "nodes": nodes, "triangles": tri})
    return out


SWEEP_C = 1500.0
SWEEP_AXES = (0.02, 0.01)
SWEEP_FLUID = (1.0357, 1.0279)


def sweep_plan(ctx, n_points=256, max_level=5, min_level=3):
    """The frequency sweep: returns plan [(f, k, level)] and the device meshes
    {level: Mesh}.  k = 2 pi f / c (src/postprocessing/sweeps.jl:107)."""
    a, b = SWEEP_AXES
    freqs = np.logspace(math.log10(1e3), math.log10(200e3), n_points)
    ks = 2 * math.pi * freqs / SWEEP_C
    levels = list(range(min_level, max_level + 1))
    meshes = {}
    for L in levels:
        nodes, tri = bh.icosphere_arrays(L, a, b, b)
        meshes[L] = bh.Mesh(nodes, tri, ctx=ctx)
    plan = []
    for f, k in zip(freqs, ks):
        lam = SWEEP_C / f
        lev = levels[-1]
        for L in levels:
            if meshes[L].resolution <= lam / 6.0:
                lev = L
                break
        plan.append((float(f), float(k), lev))
    return plan, meshes
