"""Sphere modal-series oracle (TEST INFRASTRUCTURE ONLY).

Restates the reference's exact sphere solution:
  mode count   max(20, ceil(2.5 ka + 4 cbrt(ka+1) + 10))      src/solvers/modal/sphere.jl:1-3
  rigid        A_m = -j_m'(ka) / h_m'(ka)                      src/solvers/modal/sphere.jl:5-8
  soft         A_m = -j_m(ka) / h_m(ka)                        src/solvers/modal/sphere.jl:10-13
  fluid        closed form with gh = g*h, ka_int = ka/h        src/solvers/modal/sphere.jl:25-40
  amplitude    f = (-i/k) sum_m (2m+1) P_m(cos angle) A_m      src/solvers/modal/sphere.jl:279-288
  near field   p_s = sum_m (2m+1) i^m A_m h_m(kr) P_m(cos theta) outside; inside a fluid
               sphere p = sum_m (2m+1) i^m B_m j_m(k_int r) P_m(cos theta) with
               B_m = (j_m(ka) + A_m h_m(ka)) / j_m(k_int a) (pressure continuity at r = a);
               the plane-wave expansion e^{ikr cos theta} = sum (2m+1) i^m j_m(kr) P_m
               (textbook; checked against the far-field limit in the tests)
  derivatives  f_n' = (n/x) f_n - f_{n+1}                      src/numerics/special_functions.jl:27-35
  Legendre     three-term recurrence                           src/numerics/special_functions.jl:56-64
Spherical Bessel functions come from scipy.special (the reference calls
SpecialFunctions.besselj/bessely of half-integer order).
"""
import math

import numpy as np
from scipy import special


def default_mode_count(ka):
    return max(20, math.ceil(2.5 * ka + 4.0 * np.cbrt(ka + 1.0) + 10.0))


def js(n, x):
    return special.spherical_jn(n, x)


def ys(n, x):
    return special.spherical_yn(n, x)


def hs(n, x):
    return js(n, x) + 1j * ys(n, x)


def jsd(n, x):
    return (n / x) * js(n, x) - js(n + 1, x)


def hsd(n, x):
    return (n / x) * hs(n, x) - hs(n + 1, x)


def legendre_p(n, x):
    if n == 0:
        return 1.0
    p0, p1 = 1.0, x
    for k in range(2, n + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    return p1


def modal_coefficient(boundary, m, k, a):
    ka = k * a
    kind = boundary[0]
    if kind == "rigid":
        return -jsd(m, ka) / hsd(m, ka)
    if kind == "soft":
        return -js(m, ka) / hs(m, ka)
    if kind == "fluid":
        g, h = boundary[1], boundary[2]
        kai = ka / h
        gh = g * h
        jd_int, jd_ext = jsd(m, kai), jsd(m, ka)
        j_int, j_ext = js(m, kai), js(m, ka)
        h_ext, hd_ext = hs(m, ka), hsd(m, ka)
        den = gh * j_int * hd_ext - h_ext * jd_int
        return (j_ext * jd_int - gh * j_int * jd_ext) / den
    raise ValueError(kind)


def form_function(boundary, k, a, angle=math.pi, m_max=None):
    """Sphere far-field amplitude f(angle) [m]; angle = pi is backscatter."""
    if m_max is None:
        m_max = default_mode_count(k * a)
    x = math.cos(angle)
    total = 0j
    for m in range(m_max + 1):
        total += (2 * m + 1) * legendre_p(m, x) * modal_coefficient(boundary, m, k, a)
    return -1j / k * total


def target_strength(f):
    return 20.0 * math.log10(abs(f))


def near_field(boundary, k, a, r, cos_theta, m_max=None):
    """Scattered pressure at radius r >= a, or the interior pressure of a
    fluid sphere at r < a, for a unit plane wave; theta is the angle between
    the point and the incidence direction."""
    if m_max is None:
        m_max = default_mode_count(k * a) + 10
    total = 0j
    for m in range(m_max + 1):
        A = modal_coefficient(boundary, m, k, a)
        pm = legendre_p(m, cos_theta)
        if r >= a:
            total += (2 * m + 1) * (1j ** m) * A * hs(m, k * r) * pm
        else:
            g, h = boundary[1], boundary[2]
            ki = k / h
            B = (js(m, k * a) + A * hs(m, k * a)) / js(m, ki * a)
            total += (2 * m + 1) * (1j ** m) * B * js(m, ki * r) * pm
    return total
