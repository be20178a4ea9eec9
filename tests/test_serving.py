"""Serving: the continuous request loop (bem_server_*).

Each served result must equal the solo solve of the same request (same
kernels, deterministic reductions: the check is 1e-12 relative, expected
exact), resident factors must serve repeated requests, and admission must
respect the HBM budget.  The generator of the served stream is checked on
the CPU."""
import math
import sys

import numpy as np
import pytest

import bemhip as bh
import workloads as W


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_stream_is_seeded_and_in_range():
    a = W.serving_stream(200, 12, 11)
    b = W.serving_stream(200, 12, 11)
    assert [r["k"] for r in a] == [r["k"] for r in b]
    assert all(np.array_equal(x["triangles"], y["triangles"]) for x, y in zip(a, b))
    ts = [r["t"] for r in a]
    assert ts == sorted(ts) and ts[0] > 0
    for r in a:
        nt = r["triangles"].shape[1]
        assert W.N_MIN <= nt <= W.N_MAX
        assert 0.5 <= r["ka"] <= 15
        res = bh.validate_surface(r["nodes"], r["triangles"])
        # meshed at lambda/6 unless the N floor forced a finer mesh
        assert res <= 2 * math.pi / r["k"] / 6 * 1.05 or nt <= W.N_MIN * 1.25


def test_unit_aspect_without_ripple_is_a_sphere():
    a = 0.02
    nodes, tri = W.revolved_body(a, 1.0, 0.0, 100.0)
    rad = np.linalg.norm(nodes, axis=0)
    assert np.abs(rad - a).max() < 1e-12 * a
    assert W.characteristic_radius(a, 1.0, 0.0) == pytest.approx(a, rel=1e-9)


@pytest.fixture(scope="module")
def ctx():
    say("[gpu] creating context on device 0")
    c = bh.Context(0)
    yield c
    c.close()


def _solo(ctx, mesh, bc, k, beta=math.pi / 2, alpha=0.0):
    sol = bh.bem(mesh, bc, k, incidence_angle=beta, incidence_azimuth=alpha, precision="mixed")
    return bh.scattering_amplitude(sol)


@pytest.mark.gpu
def test_results_are_the_same_alone_and_in_the_mix(ctx):
    say("[gpu] serving: 8 mixed requests on 3 lanes vs solo solves")
    reqs = W.serving_stream(800, 8, 3)
    meshes = [bh.Mesh(r["nodes"], r["triangles"], ctx=ctx) for r in reqs]
    srv = bh.Server(ctx, lanes=3, keep_factors=False)
    try:
        tickets = [srv.submit(m, r["boundary"], r["k"]) for m, r in zip(meshes, reqs)]
        served = [srv.wait(t) for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    assert st["completed"] == 8 and st["failed"] == 0
    for m, r, s in zip(meshes, reqs, served):
        f0 = _solo(ctx, m, r["boundary"], r["k"])
        assert abs(s.amplitude - f0) <= 1e-12 * abs(f0), (type(r["boundary"]).__name__, m.n)
        assert s.diagnostics["converged"]
        assert 0 <= s.queue_ms <= s.total_ms and s.run_ms > 0


@pytest.mark.gpu
def test_resident_factors_serve_repeated_requests(ctx):
    say("[gpu] serving: factor reuse across incidences")
    mesh = bh.mesh(bh.Sphere(0.01), level=3, ctx=ctx)
    bc = bh.FluidFilled(1.0357, 1.0279)
    k = 161.6
    angles = [(0.0, 0.0), (0.7, 0.2), (math.pi / 2, 1.0)]
    with bh.Server(ctx, lanes=2, keep_factors=True) as srv:
        out = []
        for beta, alpha in angles:  # one at a time: each later one finds the factors resident
            out.append(srv.wait(srv.submit(mesh, bc, k, incidence_angle=beta, incidence_azimuth=alpha)))
        st = srv.stats()
    assert st["resident_operators"] == 1 and st["factor_reuses"] == 2
    for (beta, alpha), s in zip(angles, out):
        f0 = _solo(ctx, mesh, bc, k, beta, alpha)
        assert abs(s.amplitude - f0) <= 1e-12 * abs(f0)
    # reuse skips assembly and factorization
    assert out[1].run_ms < out[0].run_ms


@pytest.mark.gpu
def test_admission_respects_the_budget():
    say("[gpu] serving: admission under a small HBM budget")
    probe = bh.Context(0)
    mesh = bh.mesh(bh.Sphere(0.01), level=3, ctx=probe)
    used0, _ = probe.memory()
    probe.close()
    # room for the meshes, lanes and about 1.5 solves of this size
    n = 2 * 1280
    need = 16 * n * n + (64 << 20)
    small = bh.Context(0, hbm_budget_bytes=used0 + int(1.5 * need) + (512 << 20))
    try:
        mesh = bh.mesh(bh.Sphere(0.01), level=3, ctx=small)
        bc = bh.FluidFilled(1.0357, 1.0279)
        with bh.Server(small, lanes=3, keep_factors=False) as srv:
            ts = [srv.submit(mesh, bc, 150.0 + 5 * i) for i in range(3)]
            res = [srv.wait(t) for t in ts]
            assert srv.stats()["completed"] == 3
            assert all(r.diagnostics["converged"] for r in res)
            big = bh.mesh(bh.Sphere(0.01), level=5, ctx=small)
            with pytest.raises(bh.AdmissionError):
                srv.submit(big, bc, 150.0)
    finally:
        small.close()


@pytest.mark.gpu
def test_server_argument_errors(ctx):
    say("[gpu] serving: argument errors")
    mesh = bh.mesh(bh.Sphere(0.01), level=2, ctx=ctx)
    with bh.Server(ctx, lanes=1) as srv:
        with pytest.raises(ValueError):
            srv.submit(mesh, bh.Rigid(), -1.0)
        with pytest.raises(ValueError):
            srv.submit(mesh, bh.Rigid(), 10.0, incidence_angle=float("nan"))
        with pytest.raises(RuntimeError):
            srv.wait(123456)
    with pytest.raises(ValueError):
        bh.Server(ctx, lanes=17)
