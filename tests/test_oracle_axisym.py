"""Pinning the axisymmetric oracle (oracle/axisym.py), CPU only.

  * Quadrature: single K_m / V_m entries (the panel itself, a neighbour, a
    far panel, the pole panel; modes 0, 1 and 5) against nested adaptive
    integration (scipy.integrate.quad) of the same ring kernels, to 1e-8 of
    the pair's largest entry.
  * The Gauss identity sum_j K0_ij = -1/2 at every collocation point of a
    closed meridian at k -> 0, to the reference's 1e-7
    (test/extended/bem/sphere/axisymmetric.jl:42-46).
  * The sphere's modal series at the reference's own bars
    (test/extended/bem/sphere/axisymmetric.jl:48-111): rigid and soft at 38
    kHz within 0.06 dB (16 panels) and 0.02 dB (32 panels), improving with
    refinement, forward scatter within 0.06 dB; the fluid sphere (g, h) =
    (1.05, 1.02) within 0.15 dB at 48 panels; the stiff and soft fluid
    limits against rigid and pressure-release (0.25 and 0.1 dB); the gas
    sphere at k = 0.014 within 0.1 dB and 1 % (axisymmetric.jl:141-149).
  * Oblique incidence: a sphere's backscatter does not depend on the
    incidence angle, so the mode sum at beta = pi/3 meets the modal series
    (1 %); the matched medium's amplitude is discretisation error, below
    5e-3 of a rigid body's and shrinking as O(h^2).
"""
import math

import numpy as np
import pytest
from scipy import integrate

from oracle import axisym as X
from oracle import modal

K38 = 2 * math.pi * 38000.0 / 1477.4


def _ts(f):
    return 20 * math.log10(abs(f))


def _brute(P, i, j, k, m):
    xr, xz = P.rm[i], P.zm[i]

    def inner(s, which, part):
        yr, yz = P.point(j, s)
        d2 = (xr - yr) ** 2 + (xz - yz) ** 2

        def f(phi):
            sh2 = math.sin(0.5 * phi) ** 2
            r = math.sqrt(d2 + 4 * xr * yr * sh2)
            G = np.exp(1j * k * r) / (4 * math.pi * r)
            if which == "V":
                v = G
            else:
                proj = P.nr[j] * (xr - P.rm[j]) + P.nz[j] * (xz - P.zm[j]) - 2 * P.nr[j] * xr * sh2
                v = -(1j * k - 1 / r) * G * proj / r
            v *= math.cos(m * phi)
            return v.real if part == 0 else v.imag

        w = math.sqrt(d2 / (xr * yr))
        pts = [p for p in (w, 2 * w, 4 * w, 8 * w) if p < math.pi]
        val = integrate.quad(f, 0, math.pi, points=pts or None, epsabs=1e-14, epsrel=1e-12, limit=400)[0]
        return 2 * val * P.L[j] * yr

    out = {}
    for which in ("K", "V"):
        parts = []
        for part in (0, 1):
            pts = [0.5] if i == j else None
            parts.append(integrate.quad(lambda s: inner(s, which, part), 0, 1, points=pts, epsabs=1e-14,
                                        epsrel=1e-11, limit=400)[0])
        out[which] = parts[0] + 1j * parts[1]
    return out


@pytest.mark.parametrize("i,j", [(5, 5), (5, 6), (5, 11), (0, 0), (3, 14)])
def test_entries_against_adaptive_integration(i, j):
    P = X.Panels(*X.spheroid_mesh(0.02, 0.01, 16))
    k = 300.0
    modes = [0, 1, 5]
    K, V = X.pair_KV(P, i, j, k, modes)
    refs = [_brute(P, i, j, k, m) for m in modes]
    # relative to the pair's largest entry over the modes (mode 0): higher
    # modes of a panel at the axis are orders of magnitude smaller
    sk = max(abs(r["K"]) for r in refs)
    sv = max(abs(r["V"]) for r in refs)
    for c, m in enumerate(modes):
        assert abs(K[c] - refs[c]["K"]) <= 1e-8 * sk, (i, j, m, K[c], refs[c]["K"])
        assert abs(V[c] - refs[c]["V"]) <= 1e-8 * sv, (i, j, m, V[c], refs[c]["V"])


def test_gauss_identity_static_double_layer():
    P = X.Panels(*X.sphere_mesh(0.01, 48))
    K, _ = X.operators(P, 1e-6, [0])
    assert np.abs(K[0].sum(axis=1) + 0.5).max() < 1e-7


def _axial(bc, k, a, n):
    P = X.Panels(*X.sphere_mesh(a, n))
    modes, ps, ds = X.solve(P, k, bc, 0.0, 0)
    return P, modes, ps, ds


@pytest.mark.parametrize("bc", [("rigid",), ("soft",)])
def test_sphere_axial_against_modal(bc):
    ts_m = _ts(modal.form_function(bc, K38, 0.01))
    ts_f = _ts(modal.form_function(bc, K38, 0.01, angle=0.0))
    d = {}
    for n in (16, 32):
        P, modes, ps, ds = _axial(bc, K38, 0.01, n)
        d[n] = abs(_ts(X.far_field(P, K38, modes, ps, ds, math.pi, math.pi)) - ts_m)
        if n == 16:
            fwd = abs(_ts(X.far_field(P, K38, modes, ps, ds, 0.0, math.pi)) - ts_f)
    assert d[16] < 0.06 and d[32] < 0.02 and d[32] < d[16]
    assert fwd < 0.06


def test_fluid_sphere_and_limits():
    a = 0.05
    P, modes, ps, ds = _axial(("fluid", 1.05, 1.02), K38, a, 48)
    f = X.far_field(P, K38, modes, ps, ds, math.pi, math.pi)
    assert abs(_ts(f) - _ts(modal.form_function(("fluid", 1.05, 1.02), K38, a))) < 0.15
    ts = {}
    for name, bc in (("stiff", ("fluid", 1e8, 1e8)), ("rigid", ("rigid",)), ("softfl", ("fluid", 1e-8, 1.0)),
                     ("soft", ("soft",))):
        P, modes, ps, ds = _axial(bc, K38, a, 24)
        ts[name] = _ts(X.far_field(P, K38, modes, ps, ds, math.pi, math.pi))
    assert abs(ts["stiff"] - ts["rigid"]) < 0.25
    assert abs(ts["softfl"] - ts["soft"]) < 0.1


def test_gas_sphere_low_frequency():
    bc, k = ("fluid", 0.0012, 0.23), 0.014
    P, modes, ps, ds = _axial(bc, k, 1.0, 96)
    f = X.far_field(P, k, modes, ps, ds, math.pi, math.pi)
    fm = modal.form_function(bc, k, 1.0)
    assert abs(_ts(f) - _ts(fm)) < 0.1
    assert abs(f - fm) / abs(fm) < 0.01


def test_oblique_sphere_and_matched_medium():
    beta = math.pi / 3
    P = X.Panels(*X.sphere_mesh(0.01, 32))
    modes, ps, ds = X.solve(P, K38, ("rigid",), beta, 20)
    f = X.far_field(P, K38, modes, ps, ds, math.pi - beta, math.pi)
    fm = modal.form_function(("rigid",), K38, 0.01)
    assert abs(f - fm) / abs(fm) < 0.01
    # matched medium: the discrete scattered field is the discretisation error
    # of the incident wave's interior representation, O(h^2) against a rigid
    # body's amplitude
    ratio = []
    for n in (24, 48):
        P = X.Panels(*X.spheroid_mesh(0.02, 0.01, n))
        modes, ps, ds = X.solve(P, 300.0, ("fluid", 1.0, 1.0), beta, 6)
        f = X.far_field(P, 300.0, modes, ps, ds, math.pi - beta, math.pi)
        modes, ps, ds = X.solve(P, 300.0, ("rigid",), beta, 6)
        ratio.append(abs(f) / abs(X.far_field(P, 300.0, modes, ps, ds, math.pi - beta, math.pi)))
    assert ratio[0] < 5e-3 and ratio[1] < ratio[0] / 3


def test_spheroid_mesh_equal_arc_length():
    rho, z = X.spheroid_mesh(0.03, 0.01, 40)
    L = np.hypot(np.diff(rho), np.diff(z))
    assert np.max(L) / np.min(L) < 1.01
    assert rho[0] == 0 and abs(rho[-1]) < 1e-15 and z[0] == 0.03 and abs(z[-1] + 0.03) < 1e-15
