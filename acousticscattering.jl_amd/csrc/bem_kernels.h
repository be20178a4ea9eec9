// Point kernels of the Helmholtz layer operators and the system-entry
// combinations, shared by the panel and the Nystrom assembly kernels.
//
// Kernel convention G = e^{ikR}/(4 pi R) (reference docs/src/models/
// boundary_methods.md:9-13).  With r = x - y, R = |r|, a = r.n_x/R,
// b = r.n_y/R, c = n_x.n_y:
//   S: G          D: G (1/R - ik) b          K': G (ik - 1/R) a
//   H: G [ k^2 ab + (c - 3ab)/R^2 + i k (3ab - c)/R ]
#pragma once
#include "bem_device.h"

namespace bem {

__device__ inline void fsincos(double x, double* s, double* c) { sincos(x, s, c); }
__device__ inline void fsincos(float x, float* s, float* c) { sincosf(x, s, c); }
__device__ inline double frsqrt(double x) { return rsqrt(x); }
__device__ inline float frsqrt(float x) { return rsqrtf(x); }

template <typename R>
struct Acc4 {
    C2<R> S, D, K, H;
};

// Add one quadrature point's contribution to the accumulators selected by NEED
// (bit 0: S, 1: D, 2: K', 3: H).
template <typename R, int NEED>
__device__ __forceinline__ void kpoint(R k, R rx, R ry, R rz, R nx0, R nx1, R nx2, R ny0, R ny1,
                                       R ny2, R c, R w, Acc4<R>& acc) {
    const R R2 = rx * rx + ry * ry + rz * rz;
    const R iR = frsqrt(R2);
    const R Rr = R2 * iR;
    R sn, cs;
    fsincos(k * Rr, &sn, &cs);
    const R gr = w * iR * R(0.07957747154594767);  // 1/(4 pi)
    const R Gx = gr * cs, Gy = gr * sn;
    if (NEED & 1) { acc.S.x += Gx; acc.S.y += Gy; }
    if (NEED & 2) {
        const R b = (rx * ny0 + ry * ny1 + rz * ny2) * iR;
        const R tx = b * iR, ty = -k * b;
        acc.D.x += Gx * tx - Gy * ty;
        acc.D.y += Gx * ty + Gy * tx;
    }
    if (NEED & 4) {
        const R a = (rx * nx0 + ry * nx1 + rz * nx2) * iR;
        const R tx = -a * iR, ty = k * a;
        acc.K.x += Gx * tx - Gy * ty;
        acc.K.y += Gx * ty + Gy * tx;
    }
    if (NEED & 8) {
        const R a = (rx * nx0 + ry * nx1 + rz * nx2) * iR;
        const R b = (rx * ny0 + ry * ny1 + rz * ny2) * iR;
        const R ab = a * b;
        const R tx = k * k * ab + (c - R(3) * ab) * iR * iR;
        const R ty = k * iR * (R(3) * ab - c);
        acc.H.x += Gx * tx - Gy * ty;
        acc.H.y += Gx * ty + Gy * tx;
    }
}

template <int KIND>
struct Need {
    static constexpr int v = (KIND == ASM_BM || KIND == ASM_MULLER) ? 15 : 3;  // operator bits S, D, K', H
    static constexpr bool two_media = (KIND == ASM_MULLER || KIND == ASM_FLUID_CBIE);
};

// Combine operator values into the off-diagonal system entries.
template <typename R, int KIND>
__device__ __forceinline__ void store_entries(C2<R>* __restrict__ out0, C2<R>* __restrict__ out1,
                                              int64_t ld, int nt, int i, int j, double k, double g,
                                              const Acc4<double>& e, const Acc4<double>& f,
                                              bool diag) {
    const double ik = 1.0 / k;
    if (KIND == ASM_BM) {
        // 1/2 d - D - (i/k) H  ;  S + (i/k)(1/2 d + K')
        double h = diag ? 0.5 : 0.0;
        out0[(int64_t)i * ld + j] = C2<R>{R(h - e.D.x + e.H.y * ik), R(-e.D.y - e.H.x * ik)};
        out1[(int64_t)i * ld + j] = C2<R>{R(e.S.x - e.K.y * ik), R(e.S.y + (e.K.x + h) * ik)};
    } else if (KIND == ASM_CBIE_RS) {
        double h = diag ? 0.5 : 0.0;
        out0[(int64_t)i * ld + j] = C2<R>{R(h - e.D.x), R(-e.D.y)};
        out1[(int64_t)i * ld + j] = C2<R>{R(e.S.x), R(e.S.y)};
    } else if (KIND == ASM_SD) {
        out0[(int64_t)i * ld + j] = C2<R>{R(e.S.x), R(e.S.y)};
        out1[(int64_t)i * ld + j] = C2<R>{R(e.D.x), R(e.D.y)};
    } else if (KIND == ASM_MULLER) {
        // [d - De + Di, Se - g Si; -He + Hi/g, d + K'e - K'i]
        double d1 = diag ? 1.0 : 0.0, ig = 1.0 / g;
        out0[(int64_t)i * ld + j] = C2<R>{R(d1 - e.D.x + f.D.x), R(-e.D.y + f.D.y)};
        out0[(int64_t)i * ld + j + nt] = C2<R>{R(e.S.x - g * f.S.x), R(e.S.y - g * f.S.y)};
        out0[(int64_t)(i + nt) * ld + j] = C2<R>{R(-e.H.x + f.H.x * ig), R(-e.H.y + f.H.y * ig)};
        out0[(int64_t)(i + nt) * ld + j + nt] = C2<R>{R(d1 + e.K.x - f.K.x), R(e.K.y - f.K.y)};
    } else {
        // [1/2 d - De, Se; 1/2 d + Di, -g Si]
        double h = diag ? 0.5 : 0.0;
        out0[(int64_t)i * ld + j] = C2<R>{R(h - e.D.x), R(-e.D.y)};
        out0[(int64_t)i * ld + j + nt] = C2<R>{R(e.S.x), R(e.S.y)};
        out0[(int64_t)(i + nt) * ld + j] = C2<R>{R(h + f.D.x), R(f.D.y)};
        out0[(int64_t)(i + nt) * ld + j + nt] = C2<R>{R(-g * f.S.x), R(-g * f.S.y)};
    }
}

template <typename R>
__device__ inline Acc4<double> widen(const Acc4<R>& a) {
    return Acc4<double>{{a.S.x, a.S.y}, {a.D.x, a.D.y}, {a.K.x, a.K.y}, {a.H.x, a.H.y}};
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ inline void wave_sum_acc(Acc4<double>& a) {
    a.S.x = wave_sum(a.S.x); a.S.y = wave_sum(a.S.y);
    a.D.x = wave_sum(a.D.x); a.D.y = wave_sum(a.D.y);
    a.K.x = wave_sum(a.K.x); a.K.y = wave_sum(a.K.y);
    a.H.x = wave_sum(a.H.x); a.H.y = wave_sum(a.H.y);
}


// Solid angle of the flat triangle V (9 doubles) about x, from Van Oosterom
// and Strackee (IEEE Trans. Biomed. Eng. 30 (1983) 125-126):
// tan(Omega/2) = a.(b x c) / (|a||b||c| + (a.b)|c| + (a.c)|b| + (b.c)|a|).
__device__ inline double tri_solid_angle(const double* V, double x0, double x1, double x2) {
    const double a0 = V[0] - x0, a1 = V[1] - x1, a2 = V[2] - x2;
    const double b0 = V[3] - x0, b1 = V[4] - x1, b2 = V[5] - x2;
    const double c0 = V[6] - x0, c1 = V[7] - x1, c2 = V[8] - x2;
    const double la = sqrt(a0 * a0 + a1 * a1 + a2 * a2), lb = sqrt(b0 * b0 + b1 * b1 + b2 * b2),
                 lc = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    const double num = a0 * (b1 * c2 - b2 * c1) + a1 * (b2 * c0 - b0 * c2) + a2 * (b0 * c1 - b1 * c0);
    const double den = la * lb * lc + (a0 * b0 + a1 * b1 + a2 * b2) * lc + (a0 * c0 + a1 * c1 + a2 * c2) * lb +
                       (b0 * c0 + b1 * c1 + b2 * c2) * la;
    return 2.0 * atan2(num, den);
}

}  // namespace bem
