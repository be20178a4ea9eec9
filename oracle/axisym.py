"""Axisymmetric (meridian x Fourier-mode) BEM oracle (TEST INFRASTRUCTURE ONLY).

The checker for the GPU library's axisymmetric path: tests/ call it; the
product path never imports it.  A numpy restatement of the same
discretisation the device kernels use, so their matrices, traces and
amplitudes can be compared entry for entry; the quadrature itself is pinned
in tests/test_oracle_axisym.py by brute-force adaptive integration
(scipy.integrate.quad) of single entries, by the Gauss identity
sum_j K0_ij = -1/2 on a closed meridian, and by the sphere's modal series.

Model (the reference's, src/solvers/bem/axisymmetric.jl):
  meridian panels, outward normal (-dz, drho)/L         src/geometry/meridian.jl:136-162
  sphere mesh (uniform polar angle), spheroid mesh
  (equal arc length along the ellipse)                  src/geometry/meridian.jl:201-259
  ring kernels G and dG/dn_y with
  proj = n.(x - y_mid) - 2 n_rho rho_x sin^2(dphi/2)   axisymmetric.jl:109-145,174-177
  mode-m operators K_m, V_m = int int kernel cos(m dphi) axisymmetric.jl:325-465
  collocation at panel midpoints, piecewise-constant traces
  rigid   (1/2 I - K_m) p = V_m dpdn_inc_m              axisymmetric.jl:995-1019
  soft    V_m dpdn = K_m p - p/2, p = -p_inc_m          axisymmetric.jl:1035-1057
  fluid   the reduced 2n system in (P, D):
          (1/2 I - Ke) P + Ve D = 0
          (1/2 I + Ki) P - g Vi D = -[(1/2 I + Ki) p_inc - g Vi dpdn_inc]
                                                        axisymmetric.jl:1185-1300
  incident modes eps_m i^m J_m(k rho sin b) e^{ik z cos b} (and the normal
  derivative with J_m')                                 axisymmetric.jl:967-979,1021-1025
  far field f(theta, phi) = -1/2 sum_j int sum_m (-i)^m cos(m phi)
          [(dp + ik cos(theta) n_z p) J_m(u) - k sin(theta) n_rho p J_m'(u)]
          rho e^{-ik z cos(theta)} L ds,  u = k rho sin(theta)
                                                        src/postprocessing/farfield.jl:55-100
  default direction: backscatter (theta = pi - beta, phi = pi)
                                                        src/solvers/shared/postprocessing.jl:10-14
Oblique incidence samples the incident modes at the midpoints, as the
reference's oblique solvers do; axial incidence of rigid and soft bodies
integrates the incident wave along the panels in the right-hand side, as its
axial solvers do (axisymmetric.jl:683-729); fluid bodies sample it at the
midpoints in both (axisymmetric.jl:806-857).

Quadrature (the device kernels' rule):
  meridian  far pairs (midpoint distance >= 2 L): 8-point Gauss on [0, 1];
            otherwise split at the point of the panel closest to the
            collocation point (the midpoint for the panel itself) and, on
            each side of length a, s = s_c -+ a t^5 with 16-point Gauss in t
            (the grading makes the logarithmic singularity of the self term
            smooth enough for 1e-10: t^3 left 1e-7).
  azimuth   dphi in [0, pi], doubled (the kernels are even in dphi);
            intervals [0, b1], [b1, b2], ... with b1 = min(w, h) and
            b_{i+1} = min(2 b_i, b_i + h, pi), where w = d / sqrt(rho_x rho_y)
            (d the meridian distance; w >= 1e-12) is the width of the ring
            kernel's peak at dphi = 0 and h = pi / (ceil(m_hi + 2 k sqrt(rho_x rho_y)) + 1)
            resolves cos(m dphi) and the phase of e^{ikr}; 8-point Gauss on
            each interval.
"""
import math

import numpy as np
from scipy import special

GL8 = np.polynomial.legendre.leggauss(8)
GL16 = np.polynomial.legendre.leggauss(16)
NEAR_FACTOR = 2.0
W_FLOOR = 1e-12


def _unit(rule):
    x, w = rule
    return 0.5 * (x + 1.0), 0.5 * w


U8 = _unit(GL8)
U16 = _unit(GL16)


# ---------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------
def sphere_mesh(a, n):
    psi = np.linspace(0.0, math.pi, n + 1)
    return a * np.sin(psi), a * np.cos(psi)


def spheroid_mesh(a, b, n):
    """Semi-axis a along the axis of symmetry, b equatorial; nodes at equal
    arc length along the meridian ellipse, north pole to south pole."""
    psi = np.linspace(0.0, math.pi, max(2000, 40 * n))
    speed = np.hypot(b * np.cos(psi), a * np.sin(psi))
    s = np.concatenate([[0.0], np.cumsum(0.5 * (speed[1:] + speed[:-1]) * np.diff(psi))])
    targets = np.linspace(0.0, s[-1], n + 1)
    nodes = np.interp(targets, s, psi)
    nodes[-1] = math.pi
    return b * np.sin(nodes), a * np.cos(nodes)


class Panels:
    def __init__(self, rho, z):
        rho, z = np.asarray(rho, float), np.asarray(z, float)
        self.r1, self.z1, self.r2, self.z2 = rho[:-1], z[:-1], rho[1:], z[1:]
        dr, dz = self.r2 - self.r1, self.z2 - self.z1
        self.L = np.hypot(dr, dz)
        self.nr, self.nz = -dz / self.L, dr / self.L
        self.rm, self.zm = 0.5 * (self.r1 + self.r2), 0.5 * (self.z1 + self.z2)
        self.n = len(self.L)

    def point(self, j, s):
        return self.r1[j] + s * (self.r2[j] - self.r1[j]), self.z1[j] + s * (self.z2[j] - self.z1[j])


# ---------------------------------------------------------------------------
# quadrature rules
# ---------------------------------------------------------------------------
def meridian_rule(P, i_or_point, j):
    """(s, w) on panel j for the collocation point of panel i (int) or an
    arbitrary point (rho, z)."""
    if isinstance(i_or_point, (int, np.integer)):
        xr, xz, self_pair = P.rm[i_or_point], P.zm[i_or_point], i_or_point == j
    else:
        (xr, xz), self_pair = i_or_point, False
    L = P.L[j]
    if not self_pair and math.hypot(xr - P.rm[j], xz - P.zm[j]) >= NEAR_FACTOR * L:
        return U8
    if self_pair:
        sc = 0.5
    else:
        dr, dz = P.r2[j] - P.r1[j], P.z2[j] - P.z1[j]
        sc = ((xr - P.r1[j]) * dr + (xz - P.z1[j]) * dz) / (L * L)
        sc = min(1.0, max(0.0, sc))
    t, wt = U16
    s, w = [], []
    for a, sign in ((sc, -1.0), (1.0 - sc, 1.0)):
        if a <= 0.0:
            continue
        s.append(sc + sign * a * t ** 5)
        w.append(5.0 * a * t ** 4 * wt)
    return np.concatenate(s), np.concatenate(w)


def azimuth_breaks(w, h):
    w = max(w, W_FLOOR)
    b = [0.0, min(w, h, math.pi)]
    while b[-1] < math.pi:
        b.append(min(2.0 * b[-1], b[-1] + h, math.pi))
    return b


def azimuth_rule(xr, yr, d, k, m_hi):
    rr = xr * yr
    w = d / math.sqrt(rr) if rr > 0.0 else math.inf
    h = math.pi / (math.ceil(m_hi + 2.0 * k * math.sqrt(max(rr, 0.0))) + 1)
    b = azimuth_breaks(w, h)
    t, wt = U8
    phi, wp = [], []
    for lo, hi in zip(b[:-1], b[1:]):
        phi.append(lo + (hi - lo) * t)
        wp.append((hi - lo) * wt)
    return np.concatenate(phi), np.concatenate(wp)


# ---------------------------------------------------------------------------
# operators
# ---------------------------------------------------------------------------
def pair_KV(P, x, j, k, modes, axial=False):
    """K_m and V_m entries of panel j at the collocation point of panel x
    (int) or at a point x = (rho, z), for every m in `modes`.  axial: also
    the panel's integrals of the mode-0 kernels against the axial incident
    wave e^{ikz} and its normal derivative ik n_z e^{ikz} along the panel
    (returned as two more values), as the reference's axial solvers
    integrate them (axisymmetric.jl:218-237,259-283)."""
    if isinstance(x, (int, np.integer)):
        xr, xz = P.rm[x], P.zm[x]
    else:
        xr, xz = x
    s, ws = meridian_rule(P, x, j)
    modes = np.asarray(modes)
    m_hi = int(modes.max())
    K = np.zeros(len(modes), complex)
    V = np.zeros(len(modes), complex)
    RK = RV = 0j
    nr, nz, L = P.nr[j], P.nz[j], P.L[j]
    # n . (x - y) is constant along a straight panel: taken at its midpoint,
    # so it is exactly zero for the panel's own collocation point
    # (axisymmetric.jl:174-177)
    mproj = nr * (xr - P.rm[j]) + nz * (xz - P.zm[j])
    for sq, wq in zip(s, ws):
        yr, yz = P.point(j, sq)
        d = math.hypot(xr - yr, xz - yz)
        phi, wp = azimuth_rule(xr, yr, d, k, m_hi)
        sh2 = np.sin(0.5 * phi) ** 2
        r = np.sqrt(d * d + 4.0 * xr * yr * sh2)
        G = np.exp(1j * k * r) / (4.0 * math.pi * r)
        proj = mproj - 2.0 * nr * xr * sh2
        dG = -(1j * k - 1.0 / r) * G * proj / r
        c = np.cos(np.outer(modes, phi))
        W = 2.0 * wp * wq * L * yr
        K += c @ (dG * W)
        V += c @ (G * W)
        if axial:
            e = np.exp(1j * k * yz)
            RK += np.sum(dG * W) * e
            RV += np.sum(G * W) * (1j * k * nz * e)
    if axial:
        return K, V, RK, RV
    return K, V


def operators(P, k, modes, axial=False):
    """K[m], V[m] (n x n complex each) for every m in `modes`; axial: also
    the row sums rK_i = sum_j int dG e^{ikz} and rV_i = sum_j int G ik n_z e^{ikz}
    of the mode-0 kernels against the axial incident wave."""
    modes = list(modes)
    n = P.n
    K = np.zeros((len(modes), n, n), complex)
    V = np.zeros((len(modes), n, n), complex)
    rK = np.zeros(n, complex)
    rV = np.zeros(n, complex)
    for i in range(n):
        for j in range(n):
            out = pair_KV(P, i, j, k, modes, axial=axial)
            K[:, i, j] = out[0]
            V[:, i, j] = out[1]
            if axial:
                rK[i] += out[2]
                rV[i] += out[3]
    if axial:
        return K, V, rK, rV
    return K, V


# ---------------------------------------------------------------------------
# incidence and solves
# ---------------------------------------------------------------------------
def default_mode_count(ka):
    return max(20, math.ceil(2.5 * ka + 4.0 * np.cbrt(ka + 1.0) + 10.0))


def panel_count(k, radius, per_wavelength=30):
    return max(20, math.ceil(per_wavelength * k * radius / 2))


def incident_modes(P, k, beta, m):
    eps = 1.0 if m == 0 else 2.0
    u = k * P.rm * math.sin(beta)
    jm = special.jv(m, u)
    jd = special.jvp(m, u)
    ph = eps * (1j ** m) * np.exp(1j * k * P.zm * math.cos(beta))
    p = ph * jm
    dp = ph * (1j * k * math.cos(beta) * P.nz * jm + k * math.sin(beta) * P.nr * jd)
    return p, dp


def solve(P, k, boundary, beta, m_max):
    """boundary: ("rigid",), ("soft",) or ("fluid", g, h).  Returns the
    per-mode scattered traces (lists of arrays), modes 0..m_max (only 0 at
    beta = 0)."""
    axial = beta == 0.0
    modes = [0] if axial else list(range(m_max + 1))
    n = P.n
    I = np.eye(n)
    if axial:
        K, V, rK, rV = operators(P, k, modes, axial=True)
    else:
        K, V = operators(P, k, modes)
    if boundary[0] == "fluid":
        g, h = boundary[1], boundary[2]
        Ki, Vi = operators(P, k / h, modes)
    ps, ds = [], []
    for c, m in enumerate(modes):
        pinc, dinc = incident_modes(P, k, beta, m)
        if boundary[0] == "rigid":
            # axial: G applied to the incident normal derivative along the
            # panels (axisymmetric.jl:683-703); oblique: to its midpoint
            # samples per mode (axisymmetric.jl:995-1019)
            p = np.linalg.solve(0.5 * I - K[c], rV if axial else V[c] @ dinc)
            d = -dinc
        elif boundary[0] == "soft":
            p = -pinc
            # axial: dG applied to p_scat = -e^{ikz} along the panels
            # (axisymmetric.jl:705-729)
            d = np.linalg.solve(V[c], (-rK if axial else K[c] @ p) - 0.5 * p)
        else:
            A = np.block([[0.5 * I - K[c], V[c]], [0.5 * I + Ki[c], -g * Vi[c]]])
            rhs = np.concatenate([np.zeros(n, complex), -((0.5 * I + Ki[c]) @ pinc - g * (Vi[c] @ dinc))])
            x = np.linalg.solve(A, rhs)
            p, d = x[:n], x[n:]
        ps.append(p)
        ds.append(d)
    return modes, ps, ds


def far_field(P, k, modes, ps, ds, theta, phi):
    st, ct = math.sin(theta), math.cos(theta)
    t, wt = U8
    total = 0j
    for j in range(P.n):
        yr, yz = P.point(j, t)
        u = k * yr * st
        val = np.zeros(len(t), complex)
        for c, m in enumerate(modes):
            jm = special.jv(m, u)
            jd = special.jvp(m, u)
            wm = ((-1j) ** m) * math.cos(m * phi)
            val += wm * ((ds[c][j] + 1j * k * ct * P.nz[j] * ps[c][j]) * jm - k * st * P.nr[j] * ps[c][j] * jd)
        total += np.sum(val * yr * np.exp(-1j * k * yz * ct) * P.L[j] * wt)
    return -0.5 * total
