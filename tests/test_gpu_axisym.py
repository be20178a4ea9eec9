"""GPU: the axisymmetric (meridian x Fourier-mode) BEM through the C ABI
(bem_axi_operators, bem_axi_solve, bem_axi_far_field).

  * The device's mode operators K_m, V_m equal the oracle's restatement
    (oracle/axisym.py, the same quadrature) to 1e-11 relative (summation order
    and FMA contraction differ; the quadrature itself is good to 1e-8), over two mode
    chunks; the static Gauss identity sum_j K0_ij = -1/2 holds to 1e-7.
  * Traces of every mode and far fields equal the oracle's solves to 1e-10
    for rigid, pressure-release and fluid bodies, axial and oblique.
  * Physics at the reference's own bars (test/extended/bem/sphere/axisymmetric.jl):
    rigid and soft spheres at 38 kHz within 0.06 / 0.02 dB of the modal
    series at 16 / 32 panels; the fluid sphere (1.05, 1.02) within 0.15 dB
    at 48 panels; the gas sphere at k = 0.014 within 0.1 dB and 1 %; an
    oblique rigid sphere (beta = pi/3) within 1 %.
  * Each mode's system is solved to a relative residual below 1e-10; bad
    meridians are refused with the reference's messages.
"""
import math
import sys

import numpy as np
import pytest

import bemhip as bh
from oracle import axisym as X
from oracle import modal

pytestmark = pytest.mark.gpu

K38 = 2 * math.pi * 38000.0 / 1477.4


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    say("[gpu] creating context on device 0")
    c = bh.Context(0)
    yield c
    c.close()


def _ts(f):
    return 20 * math.log10(abs(f))


def test_operators_match_oracle(ctx):
    say("[gpu] axisymmetric operators, spheroid 16 panels, k = 300, modes 0..9")
    rho, z = X.spheroid_mesh(0.02, 0.01, 16)
    K, V = bh.axisymmetric_operators(bh.MeridianMesh(rho, z), 300.0, range(10), ctx=ctx)
    K0, V0 = X.operators(X.Panels(rho, z), 300.0, list(range(10)))
    ek = np.abs(K - K0).max() / np.abs(K0).max()
    ev = np.abs(V - V0).max() / np.abs(V0).max()
    say(f"[gpu]   max |K - K_oracle| / max |K| {ek:.2e}, V {ev:.2e}")
    assert ek < 1e-11 and ev < 1e-11


def test_static_gauss_identity(ctx):
    rho, z = X.sphere_mesh(0.01, 48)
    K, _ = bh.axisymmetric_operators(bh.MeridianMesh(rho, z), 1e-6, [0], ctx=ctx)
    err = np.abs(K[0].sum(axis=1) + 0.5).max()
    say(f"[gpu] static identity max |sum_j K0_ij + 1/2| {err:.2e}")
    assert err < 1e-7


CASES = [("rigid", bh.Rigid(), ("rigid",)), ("soft", bh.PressureRelease(), ("soft",)),
         ("fluid", bh.FluidFilled(1.2, 1.1), ("fluid", 1.2, 1.1))]


@pytest.mark.parametrize("name,bc,obc", CASES)
@pytest.mark.parametrize("beta", [0.0, math.pi / 3])
def test_solves_match_oracle(ctx, name, bc, obc, beta):
    say(f"[gpu] axisymmetric {name} spheroid, 24 panels, k = 300, beta = {beta:.4f}")
    rho, z = X.spheroid_mesh(0.02, 0.01, 24)
    k, m_max = 300.0, 6
    sol = bh.bem(bh.MeridianMesh(rho, z), bc, k, incidence_angle=beta, m_max=m_max, ctx=ctx)
    P = X.Panels(rho, z)
    modes, ps, ds = X.solve(P, k, obc, beta, m_max)
    assert sol.n_modes == len(modes)
    assert sol.diagnostics["relative_residual"] < 1e-10
    for c in range(len(modes)):
        p, d = sol.traces(c)
        scale = max(np.abs(ps[0]).max(), np.abs(ds[0]).max() / k)
        assert np.abs(p - ps[c]).max() <= 1e-10 * scale, (c, np.abs(p - ps[c]).max())
        assert np.abs(d - ds[c]).max() <= 1e-10 * scale * k, (c, np.abs(d - ds[c]).max())
    dirs = [(math.pi - beta, math.pi), (0.3, 0.0), (2.0, 1.1)]
    f = sol.far_field(dirs)
    f0 = np.array([X.far_field(P, k, modes, ps, ds, th, ph) for th, ph in dirs])
    err = np.abs(f - f0).max() / np.abs(f0).max()
    say(f"[gpu]   far field vs oracle {err:.2e}, residual {sol.diagnostics['relative_residual']:.2e}")
    assert err < 1e-10
    assert bh.scattering_amplitude(sol) == pytest.approx(complex(f[0]), rel=1e-14)


@pytest.mark.parametrize("bc,obc", [(bh.Rigid(), ("rigid",)), (bh.PressureRelease(), ("soft",))])
def test_sphere_against_modal(ctx, bc, obc):
    ts_m = _ts(modal.form_function(obc, K38, 0.01))
    ts_f = _ts(modal.form_function(obc, K38, 0.01, angle=0.0))
    d = {}
    for n in (16, 32):
        sol = bh.bem(bh.Sphere(0.01), bc, K38, incidence_angle=0.0, n=n, ctx=ctx)
        d[n] = abs(bh.target_strength(sol) - ts_m)
        if n == 16:
            fwd = abs(bh.target_strength(sol, angle=0.0) - ts_f)
    say(f"[gpu] sphere {obc[0]}: |dTS| {d[16]:.4f} (16), {d[32]:.4f} (32) dB, forward {fwd:.4f} dB")
    assert d[16] < 0.06 and d[32] < 0.02 and d[32] < d[16]
    assert fwd < 0.06


def test_fluid_and_gas_spheres(ctx):
    a = 0.05
    sol = bh.bem(bh.Sphere(a), bh.FluidFilled(1.05, 1.02), K38, incidence_angle=0.0, n=48, ctx=ctx)
    dts = abs(bh.target_strength(sol) - _ts(modal.form_function(("fluid", 1.05, 1.02), K38, a)))
    gas = bh.bem(bh.Sphere(1.0), bh.FluidFilled(0.0012, 0.23), 0.014, incidence_angle=0.0, n=96, ctx=ctx)
    fm = modal.form_function(("fluid", 0.0012, 0.23), 0.014, 1.0)
    f = bh.scattering_amplitude(gas)
    say(f"[gpu] fluid sphere |dTS| {dts:.4f} dB; gas sphere |dTS| {abs(_ts(f) - _ts(fm)):.4f} dB, "
        f"rel {abs(f - fm) / abs(fm):.4f}")
    assert dts < 0.15
    assert abs(_ts(f) - _ts(fm)) < 0.1 and abs(f - fm) / abs(fm) < 0.01


def test_oblique_sphere(ctx):
    beta = math.pi / 3
    sol = bh.bem(bh.Sphere(0.01), bh.Rigid(), K38, incidence_angle=beta, n=32, m_max=20, ctx=ctx)
    f = bh.scattering_amplitude(sol)
    fm = modal.form_function(("rigid",), K38, 0.01)
    say(f"[gpu] oblique rigid sphere rel {abs(f - fm) / abs(fm):.4f}, residual "
        f"{sol.diagnostics['relative_residual']:.1e}")
    assert abs(f - fm) / abs(fm) < 0.01
    assert sol.n_modes == 21


def test_meridian_errors(ctx):
    rho, z = X.sphere_mesh(0.01, 12)
    with pytest.raises(ValueError, match="north pole to the south pole"):
        bh.bem(bh.MeridianMesh(rho[::-1], z[::-1]), bh.Rigid(), 100.0, incidence_angle=0.0, ctx=ctx)
    bad = rho.copy()
    bad[0] = 1e-3
    with pytest.raises(ValueError, match="meet the axis"):
        bh.bem(bh.MeridianMesh(bad, z), bh.Rigid(), 100.0, incidence_angle=0.0, ctx=ctx)
    with pytest.raises(ValueError):
        bh.bem(bh.Sphere(0.01), bh.Rigid(), 100.0, incidence_angle=0.3, m_max=-1, ctx=ctx)
