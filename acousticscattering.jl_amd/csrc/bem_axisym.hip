// Axisymmetric (meridian x Fourier-mode) BEM on the device (gfx950).
//
// A body of revolution is a meridian of straight panels in the half-plane
// (rho, z), z along the axis of symmetry; the unknowns are piecewise-constant
// Fourier-mode traces at the panel midpoints (reference
// src/solvers/bem/axisymmetric.jl).  The mode-m operators
//   K_m(i, j) = int_panel_j int_0^{2 pi} dG/dn_y(x_i, y) cos(m dphi) dphi rho_y ds
//   V_m(i, j) = int_panel_j int_0^{2 pi} G(x_i, y)       cos(m dphi) dphi rho_y ds
// with r^2 = d^2 + 4 rho_x rho_y sin^2(dphi / 2) (d the meridian distance) are
// ring integrals whose integrand peaks at dphi = 0 with width
// w = d / sqrt(rho_x rho_y), and whose meridian integral is logarithmically
// singular at the panel's own midpoint.  Rules (identical in the oracle):
//   meridian  8-point Gauss for far pairs (midpoint distance >= 2 L); else a
//             split at the closest point s_c, s = s_c -+ a t^5 on each side
//             with 16-point Gauss in t;
//   azimuth   [0, pi] doubled; intervals b1 = min(w, h), b_{i+1} =
//             min(2 b_i, b_i + h, pi), h = pi / (ceil(m_hi + 2 k sqrt(rho_x rho_y)) + 1),
//             8-point Gauss on each.
// One wave per panel pair: lanes take the azimuthal nodes of a meridian node
// in turn, modes m0 .. m0 + 7 ride along through the Chebyshev recurrence of
// cos(m dphi), and one wave reduction per value ends the pair.  FP64
// throughout.
#include <cmath>

#include "bem_device.h"
#include "bem_internal.h"

namespace bem {

constexpr double kAxiPi = 3.14159265358979323846;
constexpr double kAxiNear = 2.0;
constexpr double kAxiWFloor = 1e-12;

__device__ inline double axi_wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// The azimuthal intervals of one meridian node, walked in order.
struct AxiBreaks {
    double w, h, lo, hi;
    int idx;  // interval index of [lo, hi]
    __device__ void start(double w_, double h_) {
        w = fmax(w_, kAxiWFloor);
        h = h_;
        idx = 0;
        lo = 0.0;
        hi = fmin(fmin(w, h), kAxiPi);
    }
    __device__ bool advance() {  // false past pi
        if (hi >= kAxiPi) return false;
        const double nhi = fmin(fmin(2.0 * hi, hi + h), kAxiPi);
        lo = hi;
        hi = nhi;
        ++idx;
        return true;
    }
};

__device__ inline int axi_interval_count(double w, double h) {
    AxiBreaks b;
    b.start(w, h);
    int n = 1;
    while (b.advance()) ++n;
    return n;
}

__global__ __launch_bounds__(256) void k_axi_ops(AxiPanels P, AxiRules R, double k, int m0, int nm, int axial,
                                                 z64* __restrict__ K, z64* __restrict__ V, z64* __restrict__ RK,
                                                 z64* __restrict__ RV) {
    const int lane = threadIdx.x & 63;
    const int64_t pair = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int n = P.n;
    if (pair >= (int64_t)n * n) return;
    const int i = (int)(pair / n), j = (int)(pair % n);
    const double xr = P.rm[i], xz = P.zm[i];
    const double r1 = P.r1[j], z1 = P.z1[j], dr = P.r2[j] - r1, dz = P.z2[j] - z1;
    const double L = P.L[j], nr = P.nr[j], nz = P.nz[j];
    // n . (x - y) is constant along the straight panel: at its midpoint, so
    // exactly zero for the panel's own collocation point
    const double mproj = nr * (xr - P.rm[j]) + nz * (xz - P.zm[j]);
    const bool self = i == j;
    const bool far = !self && hypot(xr - P.rm[j], xz - P.zm[j]) >= kAxiNear * L;
    double sc = 0.5;
    if (!far && !self) sc = fmin(1.0, fmax(0.0, ((xr - r1) * dr + (xz - z1) * dz) / (L * L)));
    const int m_hi = m0 + nm - 1;
    double aK[2 * kAxiModes], aV[2 * kAxiModes];
#pragma unroll
    for (int c = 0; c < 2 * kAxiModes; ++c) aK[c] = aV[c] = 0.0;
    double rk_re = 0, rk_im = 0, rv_re = 0, rv_im = 0;
    const int nq = far ? 8 : 32;
    for (int q = 0; q < nq; ++q) {
        double s, ws;
        if (far) {
            s = R.x8[q];
            ws = R.w8[q];
        } else {
            const int side = q >> 4, t = q & 15;
            const double a = side == 0 ? sc : 1.0 - sc;
            if (a <= 0.0) continue;  // wave-uniform
            const double tt = R.x16[t], t2 = tt * tt, t4 = t2 * t2;
            s = side == 0 ? sc - a * t4 * tt : sc + a * t4 * tt;
            ws = 5.0 * a * t4 * R.w16[t];
        }
        const double yr = r1 + s * dr, yz = z1 + s * dz;
        const double ddr = xr - yr, ddz = xz - yz;
        const double d2 = ddr * ddr + ddz * ddz, d = sqrt(d2);
        const double rr = xr * yr;
        const double w = rr > 0.0 ? d / sqrt(rr) : INFINITY;
        const double h = kAxiPi / (ceil((double)m_hi + 2.0 * k * sqrt(fmax(rr, 0.0))) + 1.0);
        const int ni = axi_interval_count(w, h);
        const double wq = 2.0 * ws * L * yr;
        // the lane's nodes, in increasing order
        AxiBreaks b;
        b.start(w, h);
        double lk_re = 0, lk_im = 0, lv_re = 0, lv_im = 0;  // mode-0 sums of this node (axial rhs)
        for (int e = lane; e < 8 * ni; e += 64) {
            const int iv = e >> 3, g = e & 7;
            while (b.idx < iv) b.advance();
            const double phi = b.lo + (b.hi - b.lo) * R.x8[g];
            const double wp = (b.hi - b.lo) * R.w8[g] * wq;
            double sh, ch;
            sincos(0.5 * phi, &sh, &ch);
            const double sh2 = sh * sh;
            const double r = sqrt(d2 + 4.0 * rr * sh2);
            const double ir = 1.0 / r;
            double sn, cs;
            sincos(k * r, &sn, &cs);
            const double g0 = ir * 0.07957747154594767;  // 1 / (4 pi r)
            const double G_re = cs * g0, G_im = sn * g0;
            const double proj = mproj - 2.0 * nr * xr * sh2;
            // dG/dn_y = -(ik - 1/r) G proj / r = (1/r - ik) G proj / r
            const double f = proj * ir;
            const double dG_re = (ir * G_re + k * G_im) * f, dG_im = (ir * G_im - k * G_re) * f;
            // cos(m phi), m = m0 .., by the Chebyshev recurrence
            const double c1 = cos(phi);
            double cm = cos((double)m0 * phi), cn = cos((double)(m0 + 1) * phi);
#pragma unroll
            for (int c = 0; c < kAxiModes; ++c) {
                if (c < nm) {
                    const double wc = wp * cm;
                    aK[2 * c] += dG_re * wc;
                    aK[2 * c + 1] += dG_im * wc;
                    aV[2 * c] += G_re * wc;
                    aV[2 * c + 1] += G_im * wc;
                }
                const double nx = 2.0 * c1 * cn - cm;
                cm = cn;
                cn = nx;
            }
            if (axial) {
                lk_re += dG_re * wp;
                lk_im += dG_im * wp;
                lv_re += G_re * wp;
                lv_im += G_im * wp;
            }
        }
        if (axial) {
            // against e^{ik z_y} (K) and ik n_z e^{ik z_y} (V) at this node
            double es, ec;
            sincos(k * yz, &es, &ec);
            rk_re += lk_re * ec - lk_im * es;
            rk_im += lk_re * es + lk_im * ec;
            const double tv_re = lv_re * ec - lv_im * es, tv_im = lv_re * es + lv_im * ec;
            rv_re += -k * nz * tv_im;
            rv_im += k * nz * tv_re;
        }
    }
#pragma unroll
    for (int c = 0; c < kAxiModes; ++c) {
        if (c >= nm) break;
        const double kr = axi_wave_sum(aK[2 * c]), ki = axi_wave_sum(aK[2 * c + 1]);
        const double vr = axi_wave_sum(aV[2 * c]), vi = axi_wave_sum(aV[2 * c + 1]);
        if (lane == 0) {
            const int64_t o = ((int64_t)c * n + i) * n + j;
            K[o] = z64{kr, ki};
            V[o] = z64{vr, vi};
        }
    }
    if (axial) {
        rk_re = axi_wave_sum(rk_re);
        rk_im = axi_wave_sum(rk_im);
        rv_re = axi_wave_sum(rv_re);
        rv_im = axi_wave_sum(rv_im);
        if (lane == 0) {
            RK[(int64_t)i * n + j] = z64{rk_re, rk_im};
            RV[(int64_t)i * n + j] = z64{rv_re, rv_im};
        }
    }
}

void launch_axi_ops(hipStream_t s, const AxiPanels& P, const AxiRules& R, double k, int m0, int nm, int axial, z64* K,
                    z64* V, z64* RK, z64* RV) {
    const int64_t pairs = (int64_t)P.n * P.n;
    hipLaunchKernelGGL(k_axi_ops, dim3((unsigned)((pairs + 3) / 4)), dim3(256), 0, s, P, R, k, m0, nm, axial, K, V, RK,
                       RV);
}

// out (row-major, leading dimension ldo, at row/column offsets) = a I + b X + c Y
// over n x n blocks (X, Y row-major n x n, either may be null).
__global__ void k_axi_block(int n, z64 a, z64 b, const z64* __restrict__ X, z64 c, const z64* __restrict__ Y,
                            z64* __restrict__ out, int64_t ldo) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    const int i = (int)(e / n), j = (int)(e % n);
    z64 v = i == j ? a : z64{0.0, 0.0};
    if (X) v = cadd(v, cmul(b, X[e]));
    if (Y) v = cadd(v, cmul(c, Y[e]));
    out[(int64_t)i * ldo + j] = v;
}
void launch_axi_block(hipStream_t s, int n, z64 a, z64 b, const z64* X, z64 c, const z64* Y, z64* out, int64_t ldo) {
    const int64_t tot = (int64_t)n * n;
    hipLaunchKernelGGL(k_axi_block, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, a, b, X, c, Y, out, ldo);
}

// y_i = sum_j M[i][j] (n x n row-major), fixed order per row.
__global__ void k_axi_rowsum(int n, const z64* __restrict__ M, z64* __restrict__ y) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    if (i >= n) return;
    double re = 0, im = 0;
    for (int j = 0; j < n; ++j) {
        re += M[(int64_t)i * n + j].x;
        im += M[(int64_t)i * n + j].y;
    }
    y[i] = z64{re, im};
}
void launch_axi_rowsum(hipStream_t s, int n, const z64* M, z64* y) {
    hipLaunchKernelGGL(k_axi_rowsum, dim3((n + 63) / 64), dim3(64), 0, s, n, M, y);
}

// J_0 .. J_mx of u >= 0 into J (mx + 1 values): Miller's backward recurrence
// from an order well above max(mx, u), rescaled against overflow and
// normalised by J_0 + 2 (J_2 + J_4 + ...) = 1.
__device__ void axi_bessel(int mx, double u, double* J) {
    if (u < 1e-300) {
        J[0] = 1.0;
        for (int m = 1; m <= mx; ++m) J[m] = 0.0;
        return;
    }
    const int top = 2 * ((max(mx, (int)u) + 16 + (int)sqrt(40.0 * fmax((double)mx, u))) / 2);
    double jp = 0.0, jc = 1e-280, sum = 0.0;
    for (int m = 0; m <= mx; ++m) J[m] = 0.0;
    for (int m = top; m >= 1; --m) {
        const double jm = 2.0 * m / u * jc - jp;  // J_{m-1}
        jp = jc;
        jc = jm;
        if (m - 1 <= mx) J[m - 1] = jc;
        if ((m - 1) % 2 == 0 && m - 1 > 0) sum += jc;
        if (fabs(jc) > 1e250) {
            jc *= 1e-250;
            jp *= 1e-250;
            sum *= 1e-250;
            for (int q = m - 1; q <= mx; ++q) J[q] *= 1e-250;
        }
    }
    const double norm = 1.0 / (J[0] + 2.0 * sum);
    for (int m = 0; m <= mx; ++m) J[m] *= norm;
}

// Far field f(theta, phi) of the mode traces P[c*n + j], D[c*n + j] (modes
// m0 .. m0 + nm - 1), one workgroup per direction, 8 Gauss nodes per panel:
//   f = -1/2 sum_j int sum_m (-i)^m cos(m phi)
//       [(D + ik cos(theta) n_z P) J_m(u) - k sin(theta) n_rho P J_m'(u)] rho e^{-ik z cos(theta)} L ds.
constexpr int kAxiMaxModes = 256;  // far field: modes 0 .. 255
__global__ __launch_bounds__(256) void k_axi_far(AxiPanels P, AxiRules R, double k, int nm,
                                                 const z64* __restrict__ Pm, const z64* __restrict__ Dm,
                                                 const double* __restrict__ dirs, z64* __restrict__ out) {
    __shared__ double red[2][256];
    const double th = dirs[2 * blockIdx.x], ph = dirs[2 * blockIdx.x + 1];
    double st, ct;
    sincos(th, &st, &ct);
    const int n = P.n;
    double J[kAxiMaxModes + 2];
    double acc_re = 0, acc_im = 0;
    for (int e = threadIdx.x; e < 8 * n; e += 256) {
        const int j = e >> 3, q = e & 7;
        const double s = R.x8[q];
        const double yr = P.r1[j] + s * (P.r2[j] - P.r1[j]), yz = P.z1[j] + s * (P.z2[j] - P.z1[j]);
        const double u = k * yr * st;
        axi_bessel(nm, u, J);
        double v_re = 0, v_im = 0;
        for (int m = 0; m < nm; ++m) {
            const double jd = m == 0 ? -J[1] : 0.5 * (J[m - 1] - J[m + 1]);
            const z64 p = Pm[(int64_t)m * n + j], d = Dm[(int64_t)m * n + j];
            // (d + ik ct nz p) J - k st nr p jd
            const double a_re = d.x - k * ct * P.nz[j] * p.y, a_im = d.y + k * ct * P.nz[j] * p.x;
            const double t_re = a_re * J[m] - k * st * P.nr[j] * p.x * jd;
            const double t_im = a_im * J[m] - k * st * P.nr[j] * p.y * jd;
            // (-i)^m cos(m phi)
            const double cw = cos(m * ph);
            const int r4 = m & 3;
            const double w_re = (r4 == 0 ? 1.0 : r4 == 2 ? -1.0 : 0.0) * cw;
            const double w_im = (r4 == 1 ? -1.0 : r4 == 3 ? 1.0 : 0.0) * cw;
            v_re += w_re * t_re - w_im * t_im;
            v_im += w_re * t_im + w_im * t_re;
        }
        double es, ec;
        sincos(-k * yz * ct, &es, &ec);
        const double wgt = yr * P.L[j] * R.w8[q];
        acc_re += (v_re * ec - v_im * es) * wgt;
        acc_im += (v_re * es + v_im * ec) * wgt;
    }
    red[0][threadIdx.x] = acc_re;
    red[1][threadIdx.x] = acc_im;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = z64{-0.5 * red[0][0], -0.5 * red[1][0]};
}
void launch_axi_far(hipStream_t s, const AxiPanels& P, const AxiRules& R, double k, int nm, const z64* Pm,
                    const z64* Dm, int ndir, const double* dirs, z64* out) {
    hipLaunchKernelGGL(k_axi_far, dim3(ndir), dim3(256), 0, s, P, R, k, nm, Pm, Dm, dirs, out);
}

}  // namespace bem
