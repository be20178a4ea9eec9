// Helmholtz operator assembly on flat triangle panels (CDNA4 / gfx950).
//
// Kernel convention G = e^{ikR}/(4 pi R) (reference docs/src/models/
// boundary_methods.md:9-13).  With r = x - y, R = |r|, a = r.n_x/R,
// b = r.n_y/R, c = n_x.n_y:
//   S: G          D: G (1/R - ik) b          K': G (ik - 1/R) a
//   H: G [ k^2 ab + (c - 3ab)/R^2 + i k (3ab - c)/R ]
// Regular pairs use the 7-point rule (one unknown per panel, collocated at
// the centroid).  Near pairs (|c_i - c_j| < 2 hmax_j) use the same rule on
// 64 sub-triangles in FP64; self panels use closed-form radial integration
// with Gauss-Legendre in the polar angle (single layer, and the Hadamard
// finite part of the hypersingular operator).  Systems assembled:
//   Burton-Miller rigid/soft  (reference src/solvers/bem/full.jl:103-121)
//   Mueller and CBIE fluid    (reference src/solvers/bem/full.jl:341,344)
// Output matrices are row-major (row = target panel) so that the LU's row
// interchanges and panel reads are contiguous.
#include <cmath>
#include <vector>

#include "bem_device.h"
#include "bem_internal.h"
#include "bem_kernels.h"

namespace bem {

__constant__ double c_q7[7][4];          // l0, l1, l2, w
__constant__ double c_near[kNearPts][4]; // l0, l1, l2, w
__constant__ double c_gl[kSelfGL][2];    // node, weight on [-1, 1]

int upload_rules() {
    double l[7][3], w[7];
    q7_rule(l, w);
    double q7[7][4];
    for (int q = 0; q < 7; ++q) {
        for (int c = 0; c < 3; ++c) q7[q][c] = l[q][c];
        q7[q][3] = w[q];
    }
    std::vector<double> l3, nw;
    near_rule(l3, nw);
    std::vector<double> nr(4 * kNearPts);
    for (int p = 0; p < kNearPts; ++p) {
        for (int c = 0; c < 3; ++c) nr[4 * p + c] = l3[3 * p + c];
        nr[4 * p + 3] = nw[p];
    }
    double gx[kSelfGL], gw[kSelfGL], gl[kSelfGL][2];
    gauss_legendre(kSelfGL, gx, gw);
    for (int g = 0; g < kSelfGL; ++g) { gl[g][0] = gx[g]; gl[g][1] = gw[g]; }
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_q7), q7, sizeof q7) != hipSuccess) return 1;
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_near), nr.data(), sizeof(double) * nr.size()) != hipSuccess) return 1;
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_gl), gl, sizeof gl) != hipSuccess) return 1;
    return 0;
}


// ---------------------------------------------------------------------------
// Regular pairs.  One lane per source panel j (so the row-major stores of a
// row are coalesced); each wave walks ROWS target panels, whose centroid and
// normal are wave-uniform.
// ---------------------------------------------------------------------------
template <typename R>
struct MeshView {
    const R *cen, *nrm, *area, *gp;
};
__device__ inline MeshView<double> view_of(const DevMesh& m, double) { return {m.cen, m.nrm, m.area, m.gp}; }
__device__ inline MeshView<float> view_of(const DevMesh& m, float) { return {m.cen32, m.nrm32, m.area32, m.gp32}; }

constexpr int kAsmRows = 16;

template <typename R, int KIND>
__global__ __launch_bounds__(256) void k_assemble(DevMesh m, AsmParams p, C2<R>* __restrict__ out0,
                                                  C2<R>* __restrict__ out1) {
    const MeshView<R> v = view_of(m, R(0));
    const int nt = m.nt;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + lane;
    const bool vj = j < nt;
    const int jj = vj ? j : nt - 1;
    R gx[7], gy[7], gz[7], gw[7];
    const R aj = v.area[jj];
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        gx[q] = v.gp[(size_t)(3 * q) * nt + jj];
        gy[q] = v.gp[(size_t)(3 * q + 1) * nt + jj];
        gz[q] = v.gp[(size_t)(3 * q + 2) * nt + jj];
        gw[q] = R(c_q7[q][3]) * aj;
    }
    const R ny0 = v.nrm[jj], ny1 = v.nrm[nt + jj], ny2 = v.nrm[2 * nt + jj];
    const R k = R(p.k), ki = R(p.kint);
    const int i0 = (blockIdx.y * 4 + wv) * kAsmRows;
    for (int r = 0; r < kAsmRows; ++r) {
        const int i = i0 + r;
        if (i >= nt) break;
        const R x0 = v.cen[i], x1 = v.cen[nt + i], x2 = v.cen[2 * nt + i];
        const R nx0 = v.nrm[i], nx1 = v.nrm[nt + i], nx2 = v.nrm[2 * nt + i];
        const R c = nx0 * ny0 + nx1 * ny1 + nx2 * ny2;
        Acc4<R> e{}, f{};
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            const R rx = x0 - gx[q], ry = x1 - gy[q], rz = x2 - gz[q];
            kpoint<R, Need<KIND>::v>(k, rx, ry, rz, nx0, nx1, nx2, ny0, ny1, ny2, c, gw[q], e);
            if (Need<KIND>::two_media)
                kpoint<R, Need<KIND>::v>(ki, rx, ry, rz, nx0, nx1, nx2, ny0, ny1, ny2, c, gw[q], f);
        }
        if (vj && i != j)
            store_entries<R, KIND>(out0, out1, p.ld, nt, i, j, p.k, p.g, widen(e), widen(f), false);
    }
}

// ---------------------------------------------------------------------------
// Near pairs: one wave per pair, one lane per sub-triangle (7 points each),
// FP64, fixed-order wave reduction.
// ---------------------------------------------------------------------------
template <typename R, int KIND>
__global__ __launch_bounds__(256) void k_near(DevMesh m, AsmParams p, C2<R>* __restrict__ out0,
                                              C2<R>* __restrict__ out1) {
    const int lane = threadIdx.x & 63;
    const int64_t e = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (e >= m.nnear) return;
    const int nt = m.nt;
    const int i = m.near_row[e], j = m.near_col[e];
    const double x0 = m.cen[i], x1 = m.cen[nt + i], x2 = m.cen[2 * nt + i];
    const double nx0 = m.nrm[i], nx1 = m.nrm[nt + i], nx2 = m.nrm[2 * nt + i];
    const double ny0 = m.nrm[j], ny1 = m.nrm[nt + j], ny2 = m.nrm[2 * nt + j];
    const double c = nx0 * ny0 + nx1 * ny1 + nx2 * ny2;
    const double* V = m.vert + 9 * (size_t)j;
    const double aj = m.area[j];
    Acc4<double> a{}, b{};
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        const int pt = lane * 7 + q;
        const double l0 = c_near[pt][0], l1 = c_near[pt][1], l2 = c_near[pt][2];
        const double w = c_near[pt][3] * aj;
        const double y0 = l0 * V[0] + l1 * V[3] + l2 * V[6];
        const double y1 = l0 * V[1] + l1 * V[4] + l2 * V[7];
        const double y2 = l0 * V[2] + l1 * V[5] + l2 * V[8];
        const double rx = x0 - y0, ry = x1 - y1, rz = x2 - y2;
        kpoint<double, Need<KIND>::v>(p.k, rx, ry, rz, nx0, nx1, nx2, ny0, ny1, ny2, c, w, a);
        if (Need<KIND>::two_media)
            kpoint<double, Need<KIND>::v>(p.kint, rx, ry, rz, nx0, nx1, nx2, ny0, ny1, ny2, c, w, b);
    }
    wave_sum_acc(a);
    if (Need<KIND>::two_media) wave_sum_acc(b);
    if (lane == 0) store_entries<R, KIND>(out0, out1, p.ld, nt, i, j, p.k, p.g, a, b, false);
}

// Self panel (x at the centroid, in the panel's plane).  Per edge, with d the
// distance to the edge line and theta the polar angle from its foot:
//   S = (1/4pi) sum_e int (e^{ikR}-1)/(ik) dtheta,  R = d/cos(theta)
//   H = (1/4pi) [2 pi i k - sum_e int e^{ikR}/R dtheta]   (Hadamard f.p.)
__device__ void self_terms(const double* V, const double* x, double k, C2<double>& S, C2<double>& H) {
    double sSx = 0, sSy = 0, sHx = 0, sHy = 0;
    for (int e = 0; e < 3; ++e) {
        const double* P = V + 3 * e;
        const double* Q = V + 3 * ((e + 1) % 3);
        double t0 = Q[0] - P[0], t1 = Q[1] - P[1], t2 = Q[2] - P[2];
        const double L = sqrt(t0 * t0 + t1 * t1 + t2 * t2);
        t0 /= L; t1 /= L; t2 /= L;
        const double p0 = P[0] - x[0], p1 = P[1] - x[1], p2 = P[2] - x[2];
        const double sP = p0 * t0 + p1 * t1 + p2 * t2;
        const double f0 = p0 - sP * t0, f1 = p1 - sP * t1, f2 = p2 - sP * t2;
        const double dist = sqrt(f0 * f0 + f1 * f1 + f2 * f2);
        const double th0 = atan2(sP, dist), th1 = atan2(sP + L, dist);
        const double half = 0.5 * (th1 - th0), mid = 0.5 * (th1 + th0);
        for (int gq = 0; gq < kSelfGL; ++gq) {
            const double th = mid + half * c_gl[gq][0];
            const double wg = half * c_gl[gq][1];
            const double R = dist / cos(th);
            const double xr = k * R;
            double sn, cs;
            sincos(xr, &sn, &cs);
            const double sh = sin(0.5 * xr);
            if (k > 0) {
                sSx += wg * sn / k;
                sSy += wg * 2.0 * sh * sh / k;
            } else {
                sSx += wg * R;
            }
            sHx += wg * cs / R;
            sHy += wg * sn / R;
        }
    }
    const double i4p = 0.07957747154594767;
    S = C2<double>{sSx * i4p, sSy * i4p};
    H = C2<double>{-sHx * i4p, (2.0 * M_PI * k - sHy) * i4p};
}

template <typename R, int KIND>
__global__ __launch_bounds__(128) void k_self(DevMesh m, AsmParams p, C2<R>* __restrict__ out0,
                                              C2<R>* __restrict__ out1) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nt = m.nt;
    if (i >= nt) return;
    double V[9], x[3];
    for (int c = 0; c < 9; ++c) V[c] = m.vert[9 * (size_t)i + c];
    x[0] = m.cen[i]; x[1] = m.cen[nt + i]; x[2] = m.cen[2 * nt + i];
    Acc4<double> a{}, b{};
    self_terms(V, x, p.k, a.S, a.H);
    if (Need<KIND>::two_media) self_terms(V, x, p.kint, b.S, b.H);
    store_entries<R, KIND>(out0, out1, p.ld, nt, i, i, p.k, p.g, a, b, true);
}

template <typename R, int KIND>
static void launch_assemble_k(hipStream_t s, const DevMesh& m, const AsmParams& p, C2<R>* o0, C2<R>* o1) {
    dim3 grid((m.nt + 63) / 64, (m.nt + 4 * kAsmRows - 1) / (4 * kAsmRows));
    hipLaunchKernelGGL((k_assemble<R, KIND>), grid, dim3(256), 0, s, m, p, o0, o1);
}

template <typename R>
void launch_assemble(hipStream_t s, const DevMesh& m, const AsmParams& p, C2<R>* o0, C2<R>* o1) {
    switch (p.kind) {
    case ASM_BM: launch_assemble_k<R, ASM_BM>(s, m, p, o0, o1); break;
    case ASM_CBIE_RS: launch_assemble_k<R, ASM_CBIE_RS>(s, m, p, o0, o1); break;
    case ASM_MULLER: launch_assemble_k<R, ASM_MULLER>(s, m, p, o0, o1); break;
    case ASM_SD: launch_assemble_k<R, ASM_SD>(s, m, p, o0, o1); break;
    default: launch_assemble_k<R, ASM_FLUID_CBIE>(s, m, p, o0, o1); break;
    }
}

template <typename R, int KIND>
static void launch_near_self_k(hipStream_t s, const DevMesh& m, const AsmParams& p, C2<R>* o0, C2<R>* o1) {
    if (m.nnear > 0) {
        dim3 grid((unsigned)((m.nnear + 3) / 4));
        hipLaunchKernelGGL((k_near<R, KIND>), grid, dim3(256), 0, s, m, p, o0, o1);
    }
    hipLaunchKernelGGL((k_self<R, KIND>), dim3((m.nt + 127) / 128), dim3(128), 0, s, m, p, o0, o1);
}

template <typename R>
void launch_near_self(hipStream_t s, const DevMesh& m, const AsmParams& p, C2<R>* o0, C2<R>* o1) {
    switch (p.kind) {
    case ASM_BM: launch_near_self_k<R, ASM_BM>(s, m, p, o0, o1); break;
    case ASM_CBIE_RS: launch_near_self_k<R, ASM_CBIE_RS>(s, m, p, o0, o1); break;
    case ASM_MULLER: launch_near_self_k<R, ASM_MULLER>(s, m, p, o0, o1); break;
    case ASM_SD: launch_near_self_k<R, ASM_SD>(s, m, p, o0, o1); break;
    default: launch_near_self_k<R, ASM_FLUID_CBIE>(s, m, p, o0, o1); break;
    }
}

template void launch_assemble<double>(hipStream_t, const DevMesh&, const AsmParams&, z64*, z64*);
template void launch_assemble<float>(hipStream_t, const DevMesh&, const AsmParams&, z32*, z32*);
template void launch_near_self<double>(hipStream_t, const DevMesh&, const AsmParams&, z64*, z64*);
template void launch_near_self<float>(hipStream_t, const DevMesh&, const AsmParams&, z32*, z32*);

// ---------------------------------------------------------------------------
// Plane-wave traces (reference src/core/incident.jl:64-76) at the centroids.
// ---------------------------------------------------------------------------
__global__ void k_incident(DevMesh m, double k, int nrhs, const double* __restrict__ dirs,
                           z64* __restrict__ p, z64* __restrict__ q) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    const int nt = m.nt;
    if (i >= nt || r >= nrhs) return;
    const double d0 = dirs[3 * r], d1 = dirs[3 * r + 1], d2 = dirs[3 * r + 2];
    const double ph = k * (d0 * m.cen[i] + d1 * m.cen[nt + i] + d2 * m.cen[2 * nt + i]);
    double sn, cs;
    sincos(ph, &sn, &cs);
    const double dn = k * (d0 * m.nrm[i] + d1 * m.nrm[nt + i] + d2 * m.nrm[2 * nt + i]);
    p[(size_t)r * nt + i] = z64{cs, sn};
    q[(size_t)r * nt + i] = z64{-dn * sn, dn * cs};  // i k (d.n) e^{i k d.c}
}

void launch_incident(hipStream_t s, const DevMesh& m, double k, int nrhs, const double* dirs, z64* p, z64* q) {
    hipLaunchKernelGGL(k_incident, dim3((m.nt + 255) / 256, nrhs), dim3(256), 0, s, m, k, nrhs, dirs, p, q);
}

// ---------------------------------------------------------------------------
// Far field f(xhat) = -(1/4pi) sum_j sum_g w_jg (i k (xhat.n_j) p_j + q_j)
// e^{-ik xhat.y_jg}  (reference src/solvers/bem/full.jl:509-519), FP64,
// deterministic two-pass reduction.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_far_partial(DevMesh m, double k, int ndir,
                                                     const double* __restrict__ dirs,
                                                     const z64* __restrict__ p,
                                                     const z64* __restrict__ q, z64* partial) {
    const int rd = blockIdx.y;  // r * ndir + d
    const int r = rd / ndir, d = rd % ndir;
    const int nt = m.nt;
    const double x0 = dirs[3 * d], x1 = dirs[3 * d + 1], x2 = dirs[3 * d + 2];
    double sx = 0, sy = 0;
    for (int j = blockIdx.x * 256 + threadIdx.x; j < nt; j += gridDim.x * 256) {
        const double xn = x0 * m.nrm[j] + x1 * m.nrm[nt + j] + x2 * m.nrm[2 * nt + j];
        const z64 pj = p[(size_t)r * nt + j], qj = q[(size_t)r * nt + j];
        // coef = i k xn p + q
        const double cx = -k * xn * pj.y + qj.x, cy = k * xn * pj.x + qj.y;
        double ex = 0, ey = 0;
        if (m.nq) {
            // Nystrom meshes: the unknown's own quadrature node and weight
            const double ph = -k * (x0 * m.cen[j] + x1 * m.cen[nt + j] + x2 * m.cen[2 * nt + j]);
            sincos(ph, &ey, &ex);
        } else {
#pragma unroll
            for (int g = 0; g < 7; ++g) {
                const double ph = -k * (x0 * m.gp[(size_t)(3 * g) * nt + j] +
                                        x1 * m.gp[(size_t)(3 * g + 1) * nt + j] +
                                        x2 * m.gp[(size_t)(3 * g + 2) * nt + j]);
                double sn, cs;
                sincos(ph, &sn, &cs);
                ex += c_q7[g][3] * cs;
                ey += c_q7[g][3] * sn;
            }
        }
        const double a = m.area[j];
        sx += a * (cx * ex - cy * ey);
        sy += a * (cx * ey + cy * ex);
    }
    __shared__ double bx[256], by[256];
    bx[threadIdx.x] = sx;
    by[threadIdx.x] = sy;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if (threadIdx.x < w) {
            bx[threadIdx.x] += bx[threadIdx.x + w];
            by[threadIdx.x] += by[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)rd * gridDim.x + blockIdx.x] = z64{bx[0], by[0]};
}

__global__ void k_far_final(int count, int nblk, const z64* partial, z64* out) {
    const int rd = blockIdx.x * blockDim.x + threadIdx.x;
    if (rd >= count) return;
    double sx = 0, sy = 0;
    for (int b = 0; b < nblk; ++b) {
        sx += partial[(size_t)rd * nblk + b].x;
        sy += partial[(size_t)rd * nblk + b].y;
    }
    const double i4p = 0.07957747154594767;
    out[rd] = z64{-sx * i4p, -sy * i4p};
}

void launch_far_field(hipStream_t s, const DevMesh& m, double k, int nrhs, int ndir, const double* dirs,
                      const z64* p, const z64* q, z64* partial, z64* out) {
    int nblk = (m.nt + 255) / 256;
    if (nblk > kFarBlocks) nblk = kFarBlocks;
    hipLaunchKernelGGL(k_far_partial, dim3(nblk, nrhs * ndir), dim3(256), 0, s, m, k, ndir, dirs, p, q, partial);
    hipLaunchKernelGGL(k_far_final, dim3((nrhs * ndir + 63) / 64), dim3(64), 0, s, nrhs * ndir, nblk, partial, out);
}

// ---------------------------------------------------------------------------
// Pressure at points (reference src/fields/pressure.jl:404-472), FP64.
// One wave per point, lanes striding over the panels:
//  1. the winding number w = (1/4pi) sum_j Omega_j(x) of the closed surface
//     about x, with the solid angle of each flat panel from Van Oosterom and
//     Strackee (IEEE Trans. Biomed. Eng. 30 (1983) 125-126):
//     tan(Omega/2) = a.(b x c) / (|a||b||c| + (a.b)|c| + (a.c)|b| + (b.c)|a|),
//     a, b, c the vertices relative to x; w ~ 1 inside, 0 outside, 1/2 on a
//     face.  Outside/on (w < 3/4) and inside/on (w > 1/4) follow the
//     reference's region rule (pressure.jl:496-501);
//  2. the layer potentials of the surface traces at x:
//     exterior  p_s(x) = D p - S q                      (wavenumber k)
//     interior  p_i(x) = S (g (q + q_inc)) - D (p + p_inc)  (wavenumber k_int)
//     with S u = sum_j u_j int_Tj G, D u = sum_j u_j int_Tj dG/dn_y, the
//     7-point rule, or 64 sub-triangles x 7 points when |x - c_j| < 2 hmax_j
//     (the near rule of the assembly).  The total exterior field adds the
//     incident plane wave.  Points whose field has no pressure in their
//     region get REGION_NONE and NaN.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_pressure(DevMesh m, FieldParams fp, int npts,
                                                  const double* __restrict__ pts, const z64* __restrict__ p,
                                                  const z64* __restrict__ q, z64* __restrict__ out,
                                                  int* __restrict__ region, double* __restrict__ winding) {
    const int lane = threadIdx.x & 63;
    const int pt = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pt >= npts) return;
    const int nt = m.nt;
    const double x0 = pts[3 * pt], x1 = pts[3 * pt + 1], x2 = pts[3 * pt + 2];
    // incident plane wave at x
    const double phx = fp.k * (fp.d[0] * x0 + fp.d[1] * x1 + fp.d[2] * x2);
    double isn, ics;
    sincos(phx, &isn, &ics);
    if (fp.field == FIELD_INCIDENT) {
        if (lane == 0) {
            out[pt] = z64{ics, isn};
            region[pt] = REGION_EXTERIOR;
            winding[pt] = 0.0;
        }
        return;
    }
    double om = 0.0;
    for (int j = lane; j < nt; j += 64) om += tri_solid_angle(m.vert + 9 * (size_t)j, x0, x1, x2);
    // lane 0's sum for every lane, so the region test is wave-uniform
    om = __shfl(wave_sum(om), 0, 64);
    const double w = om * 0.07957747154594767;  // / (4 pi)
    const bool outside_or_on = w < 0.75, inside_or_on = w > 0.25;
    int reg = REGION_NONE;
    if ((fp.field == FIELD_TOTAL || fp.field == FIELD_SCATTERED) && outside_or_on) reg = REGION_EXTERIOR;
    else if (fp.fluid && (fp.field == FIELD_TOTAL || fp.field == FIELD_INTERIOR) && inside_or_on)
        reg = REGION_INTERIOR;
    if (reg == REGION_NONE) {
        if (lane == 0) {
            out[pt] = z64{NAN, NAN};
            region[pt] = reg;
            winding[pt] = w;
        }
        return;
    }
    const bool inner = reg == REGION_INTERIOR;
    const double kk = inner ? fp.kint : fp.k;
    double sx = 0.0, sy = 0.0;  // S u
    double dx = 0.0, dy = 0.0;  // D v
    for (int j = lane; j < nt; j += 64) {
        const double* V = m.vert + 9 * (size_t)j;
        const double cj0 = m.cen[j], cj1 = m.cen[nt + j], cj2 = m.cen[2 * nt + j];
        const double n0 = m.nrm[j], n1 = m.nrm[nt + j], n2 = m.nrm[2 * nt + j];
        // densities: u multiplies S, v multiplies D
        z64 u = q[j], v = p[j];
        if (inner) {
            const double ph = fp.k * (fp.d[0] * cj0 + fp.d[1] * cj1 + fp.d[2] * cj2);
            double sn, cs;
            sincos(ph, &sn, &cs);
            const double dn = fp.k * (fp.d[0] * n0 + fp.d[1] * n1 + fp.d[2] * n2);
            v = z64{v.x + cs, v.y + sn};                                // p + p_inc
            u = z64{fp.g * (u.x - dn * sn), fp.g * (u.y + dn * cs)};    // g (q + q_inc)
        }
        double e0 = V[3] - V[0], e1 = V[4] - V[1], e2 = V[5] - V[2];
        double hm = e0 * e0 + e1 * e1 + e2 * e2;
        e0 = V[6] - V[3]; e1 = V[7] - V[4]; e2 = V[8] - V[5];
        hm = fmax(hm, e0 * e0 + e1 * e1 + e2 * e2);
        e0 = V[6] - V[0]; e1 = V[7] - V[1]; e2 = V[8] - V[2];
        hm = fmax(hm, e0 * e0 + e1 * e1 + e2 * e2);
        const double r0 = x0 - cj0, r1 = x1 - cj1, r2 = x2 - cj2;
        const bool near = r0 * r0 + r1 * r1 + r2 * r2 < kNearEta * kNearEta * hm;
        const int npq = near ? kNearPts : 7;
        const double aj = m.area[j];
        Acc4<double> acc{};
        for (int g = 0; g < npq; ++g) {
            double l0, l1, l2, wq;
            if (near) { l0 = c_near[g][0]; l1 = c_near[g][1]; l2 = c_near[g][2]; wq = c_near[g][3]; }
            else { l0 = c_q7[g][0]; l1 = c_q7[g][1]; l2 = c_q7[g][2]; wq = c_q7[g][3]; }
            const double y0 = l0 * V[0] + l1 * V[3] + l2 * V[6];
            const double y1 = l0 * V[1] + l1 * V[4] + l2 * V[7];
            const double y2 = l0 * V[2] + l1 * V[5] + l2 * V[8];
            kpoint<double, 3>(kk, x0 - y0, x1 - y1, x2 - y2, 0.0, 0.0, 0.0, n0, n1, n2, 0.0, wq * aj, acc);
        }
        sx += acc.S.x * u.x - acc.S.y * u.y;
        sy += acc.S.x * u.y + acc.S.y * u.x;
        dx += acc.D.x * v.x - acc.D.y * v.y;
        dy += acc.D.x * v.y + acc.D.y * v.x;
    }
    sx = wave_sum(sx); sy = wave_sum(sy);
    dx = wave_sum(dx); dy = wave_sum(dy);
    if (lane == 0) {
        z64 val = inner ? z64{sx - dx, sy - dy} : z64{dx - sx, dy - sy};
        if (!inner && fp.field == FIELD_TOTAL) { val.x += ics; val.y += isn; }
        out[pt] = val;
        region[pt] = reg;
        winding[pt] = w;
    }
}

void launch_pressure(hipStream_t s, const DevMesh& m, const FieldParams& fp, int npts, const double* pts,
                     const z64* p, const z64* q, z64* out, int* region, double* winding) {
    if (npts <= 0) return;
    hipLaunchKernelGGL(k_pressure, dim3((npts + 3) / 4), dim3(256), 0, s, m, fp, npts, pts, p, q, out, region,
                       winding);
}

// ---------------------------------------------------------------------------
// Hierarchical path: entries by unknown index (the host maps tree positions
// to unknowns).  Rigid kinds have one unknown per panel; the Mueller kind has
// two (pressure I = i, normal derivative I = N + i).  Regular pairs in FP32
// (7-point rule); near pairs and same-panel pairs in FP64 into the
// dense-block slots the host resolved.
// ---------------------------------------------------------------------------
template <int KIND>
struct HNeed {
    static constexpr int v = KIND == HK_BM       ? 10
                             : KIND == HK_CBIE   ? 2
                             : KIND == HK_BM_RHS ? 5
                             : KIND == HK_MULLER ? 15
                                                 : 1;
};
struct HParF {
    float k, ki, g;
};
// single-medium entry from the operator values (diag adds the 1/2 I terms)
template <int KIND, typename R>
__device__ __forceinline__ z32 hcombine(const Acc4<R>& e, R ik, bool diag) {
    const R h = diag ? R(0.5) : R(0);
    if (KIND == HK_BM) return z32{float(h - e.D.x + e.H.y * ik), float(-e.D.y - e.H.x * ik)};
    if (KIND == HK_CBIE) return z32{float(h - e.D.x), float(-e.D.y)};
    if (KIND == HK_BM_RHS) return z32{float(e.S.x - e.K.y * ik), float(e.S.y + (e.K.x + h) * ik)};
    return z32{float(e.S.x), float(e.S.y)};
}
// Mueller entry of quadrant (a, b) from both media's operator values
// (full.jl:341); diag adds the identity of the diagonal quadrants.
template <typename R>
__device__ __forceinline__ z32 mcombine(const Acc4<R>& e, const Acc4<R>& f, int a, int b, R g, bool diag) {
    const R d1 = diag ? R(1) : R(0);
    if (a == 0 && b == 0) return z32{float(d1 - e.D.x + f.D.x), float(-e.D.y + f.D.y)};
    if (a == 0) return z32{float(e.S.x - g * f.S.x), float(e.S.y - g * f.S.y)};
    if (b == 0) return z32{float(-e.H.x + f.H.x / g), float(-e.H.y + f.H.y / g)};
    return z32{float(d1 + e.K.x - f.K.x), float(e.K.y - f.K.y)};
}

// 7-point regular sums of the V operators for target panel i, source panel j
template <int V>
__device__ __forceinline__ Acc4<float> hsum(const MeshView<float>& v, int nt, int i, int j, float k) {
    const float x0 = v.cen[i], x1 = v.cen[nt + i], x2 = v.cen[2 * nt + i];
    const float nx0 = v.nrm[i], nx1 = v.nrm[nt + i], nx2 = v.nrm[2 * nt + i];
    const float ny0 = v.nrm[j], ny1 = v.nrm[nt + j], ny2 = v.nrm[2 * nt + j];
    const float c = nx0 * ny0 + nx1 * ny1 + nx2 * ny2;
    const float aj = v.area[j];
    Acc4<float> e{};
    if (!v.gp) {
        // Nystrom meshes: the source node itself with its weight
        kpoint<float, V>(k, x0 - v.cen[j], x1 - v.cen[nt + j], x2 - v.cen[2 * nt + j], nx0, nx1, nx2, ny0, ny1,
                         ny2, c, aj, e);
        return e;
    }
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        const float rx = x0 - v.gp[(size_t)(3 * q) * nt + j];
        const float ry = x1 - v.gp[(size_t)(3 * q + 1) * nt + j];
        const float rz = x2 - v.gp[(size_t)(3 * q + 2) * nt + j];
        kpoint<float, V>(k, rx, ry, rz, nx0, nx1, nx2, ny0, ny1, ny2, c, float(c_q7[q][3]) * aj, e);
    }
    return e;
}

template <int KIND>
__device__ __forceinline__ int hpanel(int I, int nt) {
    return (KIND == HK_MULLER && I >= nt) ? I - nt : I;
}

// regular entry (I, J) of unknowns on different panels
template <int KIND>
__device__ __forceinline__ z32 hentry(const MeshView<float>& v, int nt, int I, int J, const HParF& p) {
    if constexpr (KIND == HK_MULLER) {
        const int a = I >= nt, b = J >= nt;
        const int i = I - a * nt, j = J - b * nt;
        const Acc4<float> e = hsum<15>(v, nt, i, j, p.k);
        const Acc4<float> f = hsum<15>(v, nt, i, j, p.ki);
        return mcombine(e, f, a, b, p.g, false);
    } else {
        return hcombine<KIND>(hsum<HNeed<KIND>::v>(v, nt, I, J, p.k), 1.f / p.k, false);
    }
}

// regular entry of the block scaled for ACA (Mueller: rows N + i by 1/sig,
// columns N + j by sig)
template <int KIND>
__device__ __forceinline__ z32 hentry_aca(const MeshView<float>& v, int nt, int I, int J, const HParF& p, float sig) {
    z32 a = hentry<KIND>(v, nt, I, J, p);
    if (KIND == HK_MULLER) {
        const float f = (I >= nt ? 1.f / sig : 1.f) * (J >= nt ? sig : 1.f);
        a.x *= f;
        a.y *= f;
    }
    return a;
}

template <int KIND>
__global__ __launch_bounds__(256) void k_hdense(DevMesh mesh, HParF p, const int* __restrict__ perm,
                                                const HBlk* __restrict__ blk, z32* __restrict__ val) {
    const HBlk b = blk[blockIdx.x];
    const MeshView<float> v = view_of(mesh, 0.f);
    const int nt = mesh.nt;
    for (int e = threadIdx.x; e < b.m * b.n; e += 256) {
        const int r = e / b.n, c = e - r * b.n;
        const int I = perm[b.row_lo + r], J = perm[b.col_lo + c];
        // same-panel pairs are written by k_hself
        val[b.off + e] = hpanel<KIND>(I, nt) == hpanel<KIND>(J, nt) ? z32{0.f, 0.f} : hentry<KIND>(v, nt, I, J, p);
    }
}

template <int KIND>
__global__ __launch_bounds__(256) void k_hnear(DevMesh m, HPar p, int64_t npairs, const int* __restrict__ pi,
                                               const int* __restrict__ pj, const int64_t* __restrict__ dst,
                                               z32* __restrict__ val) {
    const int lane = threadIdx.x & 63;
    const int64_t e = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (e >= npairs) return;
    const int nt = m.nt;
    const int I = pi[e], J = pj[e];
    const int i = hpanel<KIND>(I, nt), j = hpanel<KIND>(J, nt);
    const double x0 = m.cen[i], x1 = m.cen[nt + i], x2 = m.cen[2 * nt + i];
    const double nx0 = m.nrm[i], nx1 = m.nrm[nt + i], nx2 = m.nrm[2 * nt + i];
    const double ny0 = m.nrm[j], ny1 = m.nrm[nt + j], ny2 = m.nrm[2 * nt + j];
    const double c = nx0 * ny0 + nx1 * ny1 + nx2 * ny2;
    const double* V = m.vert + 9 * (size_t)j;
    const double aj = m.area[j];
    Acc4<double> a{}, b{};
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        const int pt = lane * 7 + q;
        const double l0 = c_near[pt][0], l1 = c_near[pt][1], l2 = c_near[pt][2];
        const double w = c_near[pt][3] * aj;
        const double y0 = l0 * V[0] + l1 * V[3] + l2 * V[6];
        const double y1 = l0 * V[1] + l1 * V[4] + l2 * V[7];
        const double y2 = l0 * V[2] + l1 * V[5] + l2 * V[8];
        kpoint<double, HNeed<KIND>::v>(p.k, x0 - y0, x1 - y1, x2 - y2, nx0, nx1, nx2, ny0, ny1, ny2, c, w, a);
        if (KIND == HK_MULLER)
            kpoint<double, 15>(p.ki, x0 - y0, x1 - y1, x2 - y2, nx0, nx1, nx2, ny0, ny1, ny2, c, w, b);
    }
    wave_sum_acc(a);
    if (KIND == HK_MULLER) wave_sum_acc(b);
    if (lane == 0)
        val[dst[e]] = KIND == HK_MULLER ? mcombine(a, b, I >= nt, J >= nt, p.g, false) : hcombine<KIND>(a, 1.0 / p.k, false);
}

// Pairs of unknowns on one flat panel: D_ii = K'_ii = 0; S_ii and the finite
// part H_ii in closed form.
template <int KIND>
__global__ __launch_bounds__(128) void k_hself(DevMesh m, HPar p, int64_t nself, const int* __restrict__ si,
                                               const int* __restrict__ sj, const int64_t* __restrict__ dst,
                                               z32* __restrict__ val) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nt = m.nt;
    if (e >= nself) return;
    const int I = si[e], J = sj[e];
    const int i = hpanel<KIND>(I, nt);
    Acc4<double> a{}, b{};
    if (KIND != HK_CBIE) {
        double V[9], x[3];
        for (int c = 0; c < 9; ++c) V[c] = m.vert[9 * (size_t)i + c];
        x[0] = m.cen[i];
        x[1] = m.cen[nt + i];
        x[2] = m.cen[2 * nt + i];
        self_terms(V, x, p.k, a.S, a.H);
        if (KIND == HK_MULLER) self_terms(V, x, p.ki, b.S, b.H);
    }
    val[dst[e]] = KIND == HK_MULLER ? mcombine(a, b, I >= nt, J >= nt, p.g, I == J)
                                    : hcombine<KIND>(a, 1.0 / p.k, true);
}

__device__ inline void add_corr(Acc4<double>& a, const z64* sd, const z64* kh, int64_t t) {
    const z64 dS = sd[2 * t], dD = sd[2 * t + 1];
    a.S.x += dS.x; a.S.y += dS.y;
    a.D.x += dD.x; a.D.y += dD.y;
    if (kh) {
        const z64 dK = kh[2 * t], dH = kh[2 * t + 1];
        a.K.x += dK.x; a.K.y += dK.y;
        a.H.x += dH.x; a.H.y += dH.y;
    }
}

template <int KIND>
__global__ __launch_bounds__(128) void k_hnys_self(DevMesh m, HPar p, int64_t nself, const int* __restrict__ si,
                                                   const int* __restrict__ sj, const int64_t* __restrict__ dst,
                                                   const z64* __restrict__ corr_sd, const z64* __restrict__ corr_kh,
                                                   const z64* __restrict__ corr_sd_i,
                                                   const z64* __restrict__ corr_kh_i, z32* __restrict__ val) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nself) return;
    const int nt = m.nt, nq = m.nq;
    const int I = si[e], J = sj[e];
    const int i = hpanel<KIND>(I, nt), j = hpanel<KIND>(J, nt);
    Acc4<double> a{}, b{};
    if (i != j) {
        const double nx0 = m.nrm[i], nx1 = m.nrm[nt + i], nx2 = m.nrm[2 * nt + i];
        const double ny0 = m.nrm[j], ny1 = m.nrm[nt + j], ny2 = m.nrm[2 * nt + j];
        const double rx = m.cen[i] - m.cen[j], ry = m.cen[nt + i] - m.cen[nt + j],
                     rz = m.cen[2 * nt + i] - m.cen[2 * nt + j];
        const double c = nx0 * ny0 + nx1 * ny1 + nx2 * ny2;
        kpoint<double, HNeed<KIND>::v>(p.k, rx, ry, rz, nx0, nx1, nx2, ny0, ny1, ny2, c, m.area[j], a);
        if (KIND == HK_MULLER) kpoint<double, 15>(p.ki, rx, ry, rz, nx0, nx1, nx2, ny0, ny1, ny2, c, m.area[j], b);
    }
    const int64_t t = (int64_t)i * nq + (j - (i / nq) * nq);
    add_corr(a, corr_sd, corr_kh, t);
    if (KIND == HK_MULLER) {
        add_corr(b, corr_sd_i, corr_kh_i, t);
        val[dst[e]] = mcombine(a, b, I >= nt, J >= nt, p.g, I == J);
    } else {
        val[dst[e]] = hcombine<KIND>(a, 1.0 / p.k, i == j);
    }
}

void launch_hnys_self(hipStream_t s, const DevMesh& m, int kind, const HPar& p, int64_t nself, const int* si,
                      const int* sj, const int64_t* dst, const z64* corr_sd, const z64* corr_kh,
                      const z64* corr_sd_i, const z64* corr_kh_i, z32* val) {
    if (nself <= 0) return;
    const dim3 grid((unsigned)((nself + 127) / 128));
#define BEM_HNS(K) hipLaunchKernelGGL((k_hnys_self<K>), grid, dim3(128), 0, s, m, p, nself, si, sj, dst, corr_sd, \
                                      corr_kh, corr_sd_i, corr_kh_i, val)
    switch (kind) {
    case HK_BM: BEM_HNS(HK_BM); break;
    case HK_CBIE: BEM_HNS(HK_CBIE); break;
    case HK_BM_RHS: BEM_HNS(HK_BM_RHS); break;
    case HK_MULLER: BEM_HNS(HK_MULLER); break;
    default: BEM_HNS(HK_CBIE_RHS); break;
    }
#undef BEM_HNS
}

static HParF hpar_f(const HPar& p) { return HParF{(float)p.k, (float)p.ki, (float)p.g}; }

void launch_hdense(hipStream_t s, const DevMesh& m, int kind, const HPar& p, const int* perm, int nblk,
                   const HBlk* blk, z32* val) {
    if (nblk <= 0) return;
#define BEM_HD(K) hipLaunchKernelGGL((k_hdense<K>), dim3(nblk), dim3(256), 0, s, m, hpar_f(p), perm, blk, val)
    switch (kind) {
    case HK_BM: BEM_HD(HK_BM); break;
    case HK_CBIE: BEM_HD(HK_CBIE); break;
    case HK_BM_RHS: BEM_HD(HK_BM_RHS); break;
    case HK_MULLER: BEM_HD(HK_MULLER); break;
    default: BEM_HD(HK_CBIE_RHS); break;
    }
#undef BEM_HD
}

template <int KIND>
static void launch_hnear_self_k(hipStream_t s, const DevMesh& m, const HPar& p, int64_t npairs, const int* pi,
                                const int* pj, const int64_t* dst, int64_t nself, const int* si, const int* sj,
                                const int64_t* self_dst, z32* val) {
    if (npairs > 0)
        hipLaunchKernelGGL((k_hnear<KIND>), dim3((unsigned)((npairs + 3) / 4)), dim3(256), 0, s, m, p, npairs, pi, pj,
                           dst, val);
    if (nself > 0)
        hipLaunchKernelGGL((k_hself<KIND>), dim3((unsigned)((nself + 127) / 128)), dim3(128), 0, s, m, p, nself, si,
                           sj, self_dst, val);
}
void launch_hnear_self(hipStream_t s, const DevMesh& m, int kind, const HPar& p, int64_t npairs, const int* pi,
                       const int* pj, const int64_t* dst, int64_t nself, const int* si, const int* sj,
                       const int64_t* self_dst, z32* val) {
    switch (kind) {
    case HK_BM: launch_hnear_self_k<HK_BM>(s, m, p, npairs, pi, pj, dst, nself, si, sj, self_dst, val); break;
    case HK_CBIE: launch_hnear_self_k<HK_CBIE>(s, m, p, npairs, pi, pj, dst, nself, si, sj, self_dst, val); break;
    case HK_BM_RHS: launch_hnear_self_k<HK_BM_RHS>(s, m, p, npairs, pi, pj, dst, nself, si, sj, self_dst, val); break;
    case HK_MULLER: launch_hnear_self_k<HK_MULLER>(s, m, p, npairs, pi, pj, dst, nself, si, sj, self_dst, val); break;
    default: launch_hnear_self_k<HK_CBIE_RHS>(s, m, p, npairs, pi, pj, dst, nself, si, sj, self_dst, val); break;
    }
}

// Block-wide reductions of the ACA kernel (NW waves).
// argmax of v >= 0 with ties to the smaller index; result in every thread.
template <int NW>
__device__ inline void aca_argmax(float& v, int& idx, float* sv, int* si) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const float ov = __shfl_xor(v, off, 64);
        const int oi = __shfl_xor(idx, off, 64);
        if (ov > v || (ov == v && oi < idx)) {
            v = ov;
            idx = oi;
        }
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        sv[threadIdx.x >> 6] = v;
        si[threadIdx.x >> 6] = idx;
    }
    __syncthreads();
    v = sv[0];
    idx = si[0];
#pragma unroll
    for (int w = 1; w < NW; ++w)
        if (sv[w] > v || (sv[w] == v && si[w] < idx)) {
            v = sv[w];
            idx = si[w];
        }
}
// sums of NV doubles per thread, fixed order (waves in pairs, then in
// order); results in every thread
template <int NW, int NV>
__device__ inline void aca_sums(double (&x)[NV], double* sd) {
#pragma unroll
    for (int q = 0; q < NV; ++q)
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) x[q] += __shfl_xor(x[q], off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) sd[(threadIdx.x >> 6) * NV + q] = x[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double t = 0;
#pragma unroll
        for (int w = 0; w < NW; w += 2) t += sd[w * NV + q] + sd[(w + 1) * NV + q];
        x[q] = t;
    }
}

// Partial-pivot adaptive cross approximation (Bebendorf, Numer. Math. 86
// (2000)) of one admissible block per workgroup, A ~ sum_l u_l v_l^T:
// residual row at the pivot row -> pivot column = its largest entry, v = row /
// pivot; residual column -> u; next pivot row = the largest |u| off the used
// row.  Stops when |u| |v| <= tol ||S||_F, with the Frobenius norm of the
// approximation updated exactly (cross terms included), or at the block's
// rank capacity (conv = 0).  NW waves per block: 4, or 8 for the largest
// blocks, whose rank steps stream every earlier U and V column (8 waves of
// 194 registers fill a CU; 16 would spill).
template <int KIND, int NW>
__global__ __launch_bounds__(64 * NW) void k_aca(DevMesh mesh, HParF p, const int* __restrict__ perm, const int* list,
                                                 HBlk* __restrict__ blks, z32* __restrict__ uv,
                                                 int* __restrict__ conv, float tol) {
    constexpr int NT = 64 * NW;
    const int bid = list ? list[blockIdx.x] : blockIdx.x;
    const HBlk b = blks[bid];
    const MeshView<float> v = view_of(mesh, 0.f);
    const int nt = mesh.nt, tid = threadIdx.x;
    z32* U = uv + b.off;
    z32* V = uv + b.off + (int64_t)b.rmax * b.m;
    __shared__ z32 coef[kAcaCap];
    __shared__ float sv[NW];
    __shared__ int si[NW];
    __shared__ double sd[NW * 32];
    double normS2 = 0;
    int r = 0, istar = 0, retries = 0, ok = 0;
    while (r < b.rmax) {
        // residual row istar -> V[:, r]
        if (tid < r) coef[tid] = U[(int64_t)tid * b.m + istar];
        __syncthreads();
        const int gi = perm[b.row_lo + istar];
        float bv = -1.f;
        int bj = 0;
        for (int j = tid; j < b.n; j += NT) {
            z32 a = hentry_aca<KIND>(v, nt, gi, perm[b.col_lo + j], p, b.sig);
            for (int l = 0; l < r; ++l) a = csub(a, cmul(coef[l], V[(int64_t)l * b.n + j]));
            V[(int64_t)r * b.n + j] = a;
            const float m2 = a.x * a.x + a.y * a.y;
            if (m2 > bv) {
                bv = m2;
                bj = j;
            }
        }
        aca_argmax<NW>(bv, bj, sv, si);
        if (!(bv > 0.f)) {
            // the residual vanishes on this row: try a few other rows first
            if (r == 0 && ++retries < 8 && retries < b.m) {
                istar = retries;
                continue;
            }
            ok = 1;
            break;
        }
        const z32 pivot = V[(int64_t)r * b.n + bj];
        const z32 pinv = cinv(pivot);
        // residual column bj -> U[:, r]
        if (tid < r) coef[tid] = V[(int64_t)tid * b.n + bj];
        __syncthreads();
        const int gj = perm[b.col_lo + bj];
        float bu = -1.f;
        int bi = 0;
        double acc[2] = {0.0, 0.0};  // |u|^2, |v|^2
        for (int i = tid; i < b.m; i += NT) {
            z32 a = hentry_aca<KIND>(v, nt, perm[b.row_lo + i], gj, p, b.sig);
            for (int l = 0; l < r; ++l) a = csub(a, cmul(U[(int64_t)l * b.m + i], coef[l]));
            U[(int64_t)r * b.m + i] = a;
            const float m2 = a.x * a.x + a.y * a.y;
            acc[0] += (double)m2;
            if (i != istar && m2 > bu) {
                bu = m2;
                bi = i;
            }
        }
        for (int j = tid; j < b.n; j += NT) {
            const z32 a = cmul(V[(int64_t)r * b.n + j], pinv);
            V[(int64_t)r * b.n + j] = a;
            acc[1] += (double)a.x * a.x + (double)a.y * a.y;
        }
        __syncthreads();
        aca_sums<NW, 2>(acc, sd);
        // cross terms 2 Re sum_{l<r} (u_l^H u_r)(v_l^H v_r), eight l at a time
        double cross = 0;
        for (int l0 = 0; l0 < r; l0 += 8) {
            double d[32];
#pragma unroll
            for (int q = 0; q < 32; ++q) d[q] = 0;
            for (int i = tid; i < b.m; i += NT) {
                const z32 ur = U[(int64_t)r * b.m + i];
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if (l0 + q < r) {
                        const z32 ul = U[(int64_t)(l0 + q) * b.m + i];
                        d[4 * q] += (double)ul.x * ur.x + (double)ul.y * ur.y;
                        d[4 * q + 1] += (double)ul.x * ur.y - (double)ul.y * ur.x;
                    }
            }
            for (int j = tid; j < b.n; j += NT) {
                const z32 vr = V[(int64_t)r * b.n + j];
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if (l0 + q < r) {
                        const z32 vl = V[(int64_t)(l0 + q) * b.n + j];
                        d[4 * q + 2] += (double)vl.x * vr.x + (double)vl.y * vr.y;
                        d[4 * q + 3] += (double)vl.x * vr.y - (double)vl.y * vr.x;
                    }
            }
            aca_sums<NW, 32>(d, sd);
#pragma unroll
            for (int q = 0; q < 8; ++q) cross += d[4 * q] * d[4 * q + 2] - d[4 * q + 1] * d[4 * q + 3];
        }
        normS2 += acc[0] * acc[1] + 2.0 * cross;
        r += 1;
        if (acc[0] * acc[1] <= (double)tol * tol * normS2) {
            ok = 1;
            break;
        }
        aca_argmax<NW>(bu, bi, sv, si);
        if (!(bu > 0.f)) {
            ok = 1;
            break;
        }
        istar = bi;
    }
    if (tid == 0) {
        blks[bid].rank = r;
        conv[bid] = ok;
    }
}

// Blocks list[0 .. nbig) (the largest; the list is sorted by m n) run with
// 8 waves per block on stream `s_big`, concurrently with the rest (4 waves
// per block) on `s`; `fork` and `join` order the two streams.
void launch_aca(hipStream_t s, hipStream_t s_big, hipEvent_t fork, hipEvent_t join, const DevMesh& m, int kind,
                const HPar& p, const int* perm, int nlist, int nbig, const int* list, HBlk* blk, z32* uv, int* conv,
                double tol) {
    if (nlist <= 0) return;
    if (!s_big) nbig = 0;
    nbig = std::min(nbig, nlist);
    if (nbig > 0) {
        (void)hipEventRecord(fork, s);
        (void)hipStreamWaitEvent(s_big, fork, 0);
    }
#define BEM_ACA(K)                                                                                                 \
    do {                                                                                                           \
        if (nbig > 0)                                                                                              \
            hipLaunchKernelGGL((k_aca<K, 8>), dim3(nbig), dim3(512), 0, s_big, m, hpar_f(p), perm, list, blk,      \
                               uv, conv, (float)tol);                                                              \
        if (nlist > nbig)                                                                                          \
            hipLaunchKernelGGL((k_aca<K, 4>), dim3(nlist - nbig), dim3(256), 0, s, m, hpar_f(p), perm, list + nbig, \
                               blk, uv, conv, (float)tol);                                                         \
    } while (0)
    switch (kind) {
    case HK_BM: BEM_ACA(HK_BM); break;
    case HK_CBIE: BEM_ACA(HK_CBIE); break;
    case HK_BM_RHS: BEM_ACA(HK_BM_RHS); break;
    case HK_MULLER: BEM_ACA(HK_MULLER); break;
    default: BEM_ACA(HK_CBIE_RHS); break;
    }
#undef BEM_ACA
    if (nbig > 0) {
        (void)hipEventRecord(join, s_big);
        (void)hipStreamWaitEvent(s, join, 0);
    }
}

}  // namespace bem
