"""GPU: the reduced-basis frequency sweep (the reference's
frequency_sweep(surface, material::FluidFilled, frequencies, c;
training_frequencies), src/postprocessing/sweeps.jl:208-306) through
bem_reduced_sweep, against the oracle's restatement (oracle/reduced.py):
the same accept/fall-back decisions, reasons and ranks, amplitudes to 1e-9,
residuals within the tolerance; the reference's own scenario
(test/basic/sampling.jl:117-163) and a band sweep on the spheroid."""
import math
import sys

import numpy as np
import pytest

import bemhip as bh
from oracle import meshes
from oracle import oracle as O
from oracle import reduced as RB

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


def _compare(gpu, ref, ref_amps, tol=1e-10):
    g = gpu["diagnostics"]
    assert g["basis_rank"] == ref["basis_rank"]
    for q, r in zip(g["queries"], ref["queries"]):
        assert q["reduced"] == r["reduced"] and q["reason"] == r["reason"] and q["rank"] == r["rank"], (q, r)
        assert max(q["residual"]["original"], q["residual"]["scaled"]) <= tol
    err = np.max(np.abs(gpu["sweep"]["amplitudes"] - ref_amps) / np.abs(ref_amps))
    say(f"[gpu]   amplitudes vs oracle: {err:.2e}; reduced {[q['reduced'] for q in g['queries']]}")
    assert err < 1e-9


def test_reduced_sweep_reference_scenario(ctx):
    say("[gpu] reduced sweep, unit sphere L2, reference scenario")
    nodes, tri = meshes.sphere(1.0, 2)
    gm, om = bh.Mesh(nodes, tri + 1, ctx=ctx), O.Mesh(nodes, tri)
    c = 1500.0
    ks, kt = np.array([0.6, 0.7, 1.2, 0.7]), np.array([0.6, 0.8])
    f, ft = ks * c / (2 * math.pi), kt * c / (2 * math.pi)
    mat = bh.FluidFilled(1.05, 1.02)
    out = bh.frequency_sweep(gm, mat, f, c, training_frequencies=ft, incidence_angle=0.7, incidence_azimuth=0.3,
                             return_diagnostics=True)
    ref_amps, ref = RB.reduced_sweep(om, 1.05, 1.02, ks, kt, 0.7, 0.3)
    _compare(out, ref, ref_amps)
    assert out["diagnostics"]["basis_rank"] == 2 and out["diagnostics"]["queries"][0]["reduced"]
    assert out["diagnostics"]["basis_bytes"] == 16 * 2 * gm.n * 2
    empty = bh.frequency_sweep(gm, mat, f, c, training_frequencies=ft, incidence_angle=0.7, incidence_azimuth=0.3,
                               maxbytes=0, return_diagnostics=True)
    assert empty["diagnostics"]["basis_rank"] == 0
    assert all(not q["reduced"] and q["reason"] == "empty" for q in empty["diagnostics"]["queries"])
    assert np.max(np.abs(empty["sweep"]["amplitudes"] - ref_amps) / np.abs(ref_amps)) < 1e-8
    plain = bh.frequency_sweep(gm, mat, f[:1], c, training_frequencies=ft, capacity=0, incidence_angle=0.7,
                               incidence_azimuth=0.3)
    assert abs(plain["amplitudes"][0] - ref_amps[0]) < 1e-9 * abs(ref_amps[0])
    for bad in ([], [math.nan], [-1.0]):
        with pytest.raises(ValueError):
            bh.frequency_sweep(gm, mat, bad, c, training_frequencies=ft)
        with pytest.raises(ValueError):
            bh.frequency_sweep(gm, mat, f, c, training_frequencies=bad)
    for opt in ({"capacity": -1}, {"maxbytes": -1}, {"tolerance": math.nan}, {"rank_tolerance": 0.0},
                {"incidence_angle": math.inf}):
        with pytest.raises(ValueError):
            bh.frequency_sweep(gm, mat, f, c, training_frequencies=ft, **opt)
    with pytest.raises(ValueError):
        bh.frequency_sweep(gm, mat, f, 0.0, training_frequencies=ft)


def test_reduced_sweep_band_on_spheroid(ctx):
    """A band of 12 queries between 6 training frequencies on the
    0.02 x 0.01 m fluid spheroid (1280 panels, 2560 unknowns): at the
    default 1e-10 every query falls back to the full solve (six snapshots do
    not resolve the band that finely); at 1e-4 the basis solutions are
    accepted, and both runs decide exactly as the oracle does."""
    say("[gpu] reduced sweep, spheroid L3, 6 training + 12 queries")
    nodes, tri = meshes.spheroid(0.02, 0.01, 3)
    gm, om = bh.Mesh(nodes, tri + 1, ctx=ctx), O.Mesh(nodes, tri)
    c, g, h = 1500.0, 1.0357, 1.0279
    ft = np.linspace(5e3, 15e3, 6)
    f = np.linspace(5.5e3, 14.5e3, 12)
    out = bh.frequency_sweep(gm, bh.FluidFilled(g, h), f, c, training_frequencies=ft, incidence_angle=0.0,
                             return_diagnostics=True)
    ks, kt = 2 * math.pi * f / c, 2 * math.pi * ft / c
    ref_amps, ref = RB.reduced_sweep(om, g, h, ks, kt, 0.0, 0.0)
    _compare(out, ref, ref_amps)
    loose = bh.frequency_sweep(gm, bh.FluidFilled(g, h), f, c, training_frequencies=ft, incidence_angle=0.0,
                               tolerance=1e-4, return_diagnostics=True)
    ref_amps4, ref4 = RB.reduced_sweep(om, g, h, ks, kt, 0.0, 0.0, tolerance=1e-4)
    _compare(loose, ref4, ref_amps4, tol=1e-4)
    assert sum(q["reduced"] for q in loose["diagnostics"]["queries"]) >= 6
