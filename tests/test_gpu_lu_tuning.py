"""GPU: factorization shapes and the tuner.

  * The bf16-split outer trailing updates (three bf16 pieces per FP32 value,
    six piece products per real product on the bf16 matrix cores) give a
    complex-FP32 factorization as accurate as the FP32 3M updates:
    ||P A - L U||_F / ||A||_F within twice the FP32 form's (about 2e-5 on
    these random Gaussian matrices, pivot growth included), at sizes with and
    without 128-aligned trailing matrices.
  * A factorization replayed from a captured HIP graph equals the directly
    launched one bit for bit (the launch sequence does not depend on data).
  * The tuner's report, and sweeps under every tuning mode give the same
    amplitudes (each solve is refined to the same FP64 backward error).
"""
import sys

import numpy as np
import pytest

import bemhip as bh
from oracle import meshes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    say("[gpu] creating context on device 0")
    c = bh.Context(0)
    yield c
    c.close()


def _matrix(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))


def _factor_residual(A, LU, piv):
    n = A.shape[0]
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    PA = A.copy()
    for i in range(n):
        p = piv[i] - 1
        if p != i:
            PA[[i, p]] = PA[[p, i]]
    return np.linalg.norm(PA - L @ U) / np.linalg.norm(A)


@pytest.mark.parametrize("n", [1000, 1536, 2600])
def test_split_updates_as_accurate_as_fp32(ctx, n):
    say(f"[gpu] complex FP32 LU, n = {n}: bf16-split vs FP32 outer updates")
    A = _matrix(n, n)
    LUs, ps = bh.debug_lu(A, single=True, ctx=ctx, updates="split")
    LUf, pf = bh.debug_lu(A, single=True, ctx=ctx, updates="fp32")
    rs, rf = _factor_residual(A, LUs, ps), _factor_residual(A, LUf, pf)
    same = np.mean(ps == pf)
    say(f"[gpu]   ||PA - LU|| / ||A||: split {rs:.2e}, fp32 {rf:.2e}; pivots equal {same:.3f}")
    assert rf < 1e-4
    assert rs <= 2.0 * rf


def test_split_solve_matches_lapack(ctx):
    """One solve with the complex-FP32 factors against LAPACK's complex64
    solve, both measured from the FP64 solution (the bar of the FP32-update
    factorization's own test)."""
    import scipy.linalg as sla
    n = 1536
    A = _matrix(n, 7)
    b = np.random.default_rng(8).standard_normal(n) + 0j
    x0 = np.linalg.solve(A, b)
    ref32 = np.abs(sla.lu_solve(sla.lu_factor(A.astype(np.complex64)), b.astype(np.complex64)) - x0).max()
    _, _, xs = bh.debug_lu(A, single=True, b=b, ctx=ctx, updates="split")
    _, _, xf = bh.debug_lu(A, single=True, b=b, ctx=ctx, updates="fp32")
    es, ef = np.abs(xs - x0).max(), np.abs(xf - x0).max()
    say(f"[gpu] FP32 factor solves vs FP64: split {es:.2e}, fp32 updates {ef:.2e}, LAPACK complex64 {ref32:.2e}")
    assert es < 20 * ref32 + 1e-6 * np.abs(x0).max()


@pytest.mark.parametrize("single,updates", [(True, "split"), (True, "fp32"), (False, None)])
def test_graph_replay_equals_direct_launch(ctx, single, updates):
    n = 1300
    say(f"[gpu] LU n = {n}, single={single}, updates={updates}: captured graph vs direct")
    A = _matrix(n, 3)
    LUd, pd = bh.debug_lu(A, single=single, ctx=ctx, updates=updates)
    LUg, pg = bh.debug_lu(A, single=single, ctx=ctx, updates=updates, graph=True)
    assert np.array_equal(pd, pg)
    assert np.array_equal(LUd, LUg)


def test_tuner_report(ctx):
    for n, prec in ((2560, "mixed"), (2560, "double")):
        r = ctx.autotune(n, prec)
        say(f"[gpu] tuner n = {n} {prec}: {r}")
        assert r["tuned"] and r["unknowns"] == n and r["candidates"] >= 4
        assert 0 < r["best_ms"] <= r["default_ms"]
        assert r["outer"] in (256, 512) and r["group"] in (4, 8, 16)
        assert not (r["split"] and prec == "double")
        assert ctx.lu_tune(n, prec) == r


def test_sweep_same_under_every_tuning_mode(ctx):
    nodes, tri = meshes.spheroid(0.02, 0.01, 3)
    gm = bh.Mesh(nodes, tri + 1, ctx=ctx)
    nodes4, tri4 = meshes.spheroid(0.02, 0.01, 4)
    gm4 = bh.Mesh(nodes4, tri4 + 1, ctx=ctx)
    fl = bh.FluidFilled(1.0357, 1.0279)
    ks = [60.0, 300.0, 520.0, 700.0]
    ms = [gm, gm, gm4, gm4]
    probs = [bh.make_problem(fl, k, precision="mixed") for k in ks]
    out = {}
    for mode in ("default", "split", "tuned"):
        ctx.set_lu_tuning(mode)
        a, _, d = bh.batched_sweep(ms, probs, [(0.0, 0.0)] * len(ks), ctx=ctx)
        assert all(x["converged"] and x["factor_precision"] == "single" for x in d)
        out[mode] = a
        say(f"[gpu] sweep under {mode}: {np.round(a, 8)}")
    ctx.set_lu_tuning("tuned")
    for mode in ("split", "tuned"):
        rel = np.abs(out[mode] - out["default"]) / np.abs(out["default"])
        assert rel.max() < 1e-9, (mode, rel)
