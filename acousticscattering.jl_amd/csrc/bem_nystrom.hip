// Curved-element Nystrom operators with density-interpolation corrections
// (CDNA4 / gfx950).
//
// The reference's default full-3D discretisation (src/solvers/bem/full.jl:
// 89-97 over an Inti quadrature, src/geometry/surface.jl:331): unknowns at the
// quadrature nodes of curved triangles, layer operators w_j K(x_i, y_j) with
// the coincident entry zero, then a correction of each target's own-element
// block that makes the Green identity
//     -1/2 u(x) + S[du/dn](x) - D[u](x) = 0      (and its normal derivative
//     -1/2 du/dn(x) + K'[du/dn](x) - H[u](x) = 0)
// exact on an interpolation basis of Helmholtz solutions regular inside the
// body (src/solvers/bem/quadrature/fluid.jl:78-163; Faria, Perez-Arancibia
// and Bonnet 2021).  Four kernels:
//   k_nys_assemble  naive entries of the system, one lane per source node;
//   k_nys_basis     basis traces at the nodes (point sources or regular waves);
//   k_nys_defect    the identities' defects, S/D (or K'/H) evaluated on the fly
//                   against LDS-staged source chunks, one thread per target;
//   k_nys_weights   per element (one wave), the row-scaled minimum-norm fit of
//                   each target's defect by one-sided Jacobi SVD in FP64;
//   k_nys_apply     the corrections added into the element-diagonal blocks.
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "bem_device.h"
#include "bem_kernels.h"

namespace bem {

// ---------------------------------------------------------------------------
// naive entries: one lane per source node j (coalesced row-major stores),
// ROWS wave-uniform target nodes per wave; FP64 arithmetic, R storage
// ---------------------------------------------------------------------------
constexpr int kNysRows = 16;

template <typename R, int KIND>
__global__ __launch_bounds__(256) void k_nys_assemble(DevMesh m, AsmParams p, C2<R>* __restrict__ out0,
                                                      C2<R>* __restrict__ out1) {
    const int nt = m.nt;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + lane;
    const bool vj = j < nt;
    const int jj = vj ? j : nt - 1;
    const double y0 = m.cen[jj], y1 = m.cen[nt + jj], y2 = m.cen[2 * nt + jj];
    const double ny0 = m.nrm[jj], ny1 = m.nrm[nt + jj], ny2 = m.nrm[2 * nt + jj];
    const double wj = m.area[jj];
    const int i0 = (blockIdx.y * 4 + wv) * kNysRows;
    for (int r = 0; r < kNysRows; ++r) {
        const int i = i0 + r;
        if (i >= nt) break;
        const double x0 = m.cen[i], x1 = m.cen[nt + i], x2 = m.cen[2 * nt + i];
        const double nx0 = m.nrm[i], nx1 = m.nrm[nt + i], nx2 = m.nrm[2 * nt + i];
        const double c = nx0 * ny0 + nx1 * ny1 + nx2 * ny2;
        Acc4<double> e{}, f{};
        if (i != j) {
            const double rx = x0 - y0, ry = x1 - y1, rz = x2 - y2;
            kpoint<double, Need<KIND>::v>(p.k, rx, ry, rz, nx0, nx1, nx2, ny0, ny1, ny2, c, wj, e);
            if (Need<KIND>::two_media)
                kpoint<double, Need<KIND>::v>(p.kint, rx, ry, rz, nx0, nx1, nx2, ny0, ny1, ny2, c, wj, f);
        }
        if (vj) store_entries<R, KIND>(out0, out1, p.ld, nt, i, j, p.k, p.g, e, f, i == j);
    }
}

template <typename R, int KIND>
static void nys_assemble_k(hipStream_t s, const DevMesh& m, const AsmParams& p, C2<R>* o0, C2<R>* o1) {
    dim3 grid((m.nt + 63) / 64, (m.nt + 4 * kNysRows - 1) / (4 * kNysRows));
    hipLaunchKernelGGL((k_nys_assemble<R, KIND>), grid, dim3(256), 0, s, m, p, o0, o1);
}

template <typename R>
void launch_nys_assemble(hipStream_t s, const DevMesh& m, const AsmParams& p, C2<R>* o0, C2<R>* o1) {
    switch (p.kind) {
    case ASM_BM: nys_assemble_k<R, ASM_BM>(s, m, p, o0, o1); break;
    case ASM_CBIE_RS: nys_assemble_k<R, ASM_CBIE_RS>(s, m, p, o0, o1); break;
    case ASM_MULLER: nys_assemble_k<R, ASM_MULLER>(s, m, p, o0, o1); break;
    case ASM_SD: nys_assemble_k<R, ASM_SD>(s, m, p, o0, o1); break;
    default: nys_assemble_k<R, ASM_FLUID_CBIE>(s, m, p, o0, o1); break;
    }
}
template void launch_nys_assemble<double>(hipStream_t, const DevMesh&, const AsmParams&, z64*, z64*);
template void launch_nys_assemble<float>(hipStream_t, const DevMesh&, const AsmParams&, z32*, z32*);

// ---------------------------------------------------------------------------
// basis traces at the nodes, one thread per node
// ---------------------------------------------------------------------------
__device__ inline z64 zmul(z64 a, z64 b) { return z64{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline z64 zadd(z64 a, z64 b) { return z64{a.x + b.x, a.y + b.y}; }
__device__ inline z64 zsc(z64 a, double s) { return z64{a.x * s, a.y * s}; }

// Basis traces at one point x with normal n: g0[p * stride], g1[p * stride]
// for p < b.P (with n = 0 the fluxes are zero).
__device__ void basis_point(const NysBasis& b, const double* x, const double* n, z64* g0, z64* g1,
                            size_t stride) {
    if (b.kind == 0) {
        // G(x, z) = e^{ikR}/(4 pi R) and its derivative along n at x
        for (int p = 0; p < b.P; ++p) {
            const double rx = x[0] - b.zs[3 * p], ry = x[1] - b.zs[3 * p + 1], rz = x[2] - b.zs[3 * p + 2];
            const double R2 = rx * rx + ry * ry + rz * rz;
            const double iR = rsqrt(R2);
            double sn, cs;
            sincos(b.k * R2 * iR, &sn, &cs);
            const double gr = iR * 0.07957747154594767;
            const double Gx = gr * cs, Gy = gr * sn;
            const double a = (rx * n[0] + ry * n[1] + rz * n[2]) * iR;
            const double tx = -a * iR, ty = b.k * a;
            g0[(size_t)p * stride] = z64{Gx, Gy};
            g1[(size_t)p * stride] = z64{Gx * tx - Gy * ty, Gx * ty + Gy * tx};
        }
        return;
    }
    // regular solid harmonics (Cartesian recurrences) times the normalised
    // spherical Bessel series, in the order m = 0..L, l = m..L, conjugate
    // partner after each m > 0 term (src/solvers/bem/quadrature/fluid.jl:19-61)
    const double v[3] = {(x[0] - b.center[0]) / b.radius, (x[1] - b.center[1]) / b.radius,
                         (x[2] - b.center[2]) / b.radius};
    const double t = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const z64 xy{v[0], v[1]};
    z64 diag{1.0, 0.0};
    z64 grad[3] = {{0, 0}, {0, 0}, {0, 0}};
    const double kr2 = (b.k * b.radius) * (b.k * b.radius);
    int col = 0;
    for (int m = 0; m <= b.degree; ++m) {
        if (m > 0) {
            const double f = -(2.0 * m - 1.0);
            const z64 g0n = zsc(zadd(zmul(xy, grad[0]), diag), f);
            const z64 g1n = zsc(zadd(zmul(xy, grad[1]), z64{-diag.y, diag.x}), f);
            const z64 g2n = zsc(zmul(xy, grad[2]), f);
            grad[0] = g0n; grad[1] = g1n; grad[2] = g2n;
            diag = zsc(zmul(diag, xy), f);
        }
        z64 prev{0, 0}, pg[3] = {{0, 0}, {0, 0}, {0, 0}};
        z64 cur = diag, cg[3] = {grad[0], grad[1], grad[2]};
        for (int l = m; l <= b.degree; ++l) {
            if (l > m) {
                const double a1 = 2.0 * l - 1.0, a2 = (double)(l + m - 1), inv = 1.0 / (double)(l - m);
                const z64 nxt = zsc(zadd(zsc(cur, a1 * v[2]), zsc(prev, -a2 * t)), inv);
                z64 ng[3];
                for (int c = 0; c < 3; ++c) {
                    z64 term = zsc(cg[c], v[2]);
                    if (c == 2) term = zadd(term, cur);
                    const z64 back = zadd(zsc(pg[c], t), zsc(prev, 2.0 * v[c]));
                    ng[c] = zsc(zadd(zsc(term, a1), zsc(back, -a2)), inv);
                }
                prev = cur;
                for (int c = 0; c < 3; ++c) pg[c] = cg[c];
                cur = nxt;
                for (int c = 0; c < 3; ++c) cg[c] = ng[c];
            }
            double coef = 1.0, radial = 1.0, deriv = 0.0, tp = 1.0;  // tp = t^(p-1)
            for (int pp = 1; pp <= 20; ++pp) {
                coef *= -kr2 / (2.0 * pp * (2.0 * l + 2.0 * pp + 1.0));
                deriv += pp * coef * tp;
                tp *= t;
                radial += coef * tp;
            }
            double fac = 1.0;
            for (int q = l - m + 1; q <= l + m; ++q) fac *= q;
            const double nrm = sqrt((2.0 * l + 1.0) / (4.0 * M_PI) / fac);
            const z64 val = zsc(cur, nrm * radial);
            z64 fl{0, 0};
            for (int c = 0; c < 3; ++c)
                fl = zadd(fl, zsc(zadd(zsc(cg[c], radial), zsc(cur, 2.0 * v[c] * deriv)), n[c]));
            fl = zsc(fl, nrm / b.radius);
            g0[(size_t)col * stride] = val;
            g1[(size_t)col * stride] = fl;
            ++col;
            if (m > 0) {
                g0[(size_t)col * stride] = z64{val.x, -val.y};
                g1[(size_t)col * stride] = z64{fl.x, -fl.y};
                ++col;
            }
        }
    }
}

__global__ __launch_bounds__(128) void k_nys_basis(DevMesh m, NysBasis b, z64* __restrict__ g0,
                                                   z64* __restrict__ g1) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int nt = m.nt;
    if (j >= nt) return;
    const double x[3] = {m.cen[j], m.cen[nt + j], m.cen[2 * nt + j]};
    const double n[3] = {m.nrm[j], m.nrm[nt + j], m.nrm[2 * nt + j]};
    basis_point(b, x, n, g0 + j, g1 + j, (size_t)nt);
}

void launch_nys_basis(hipStream_t s, const DevMesh& m, const NysBasis& b, z64* g0, z64* g1) {
    hipLaunchKernelGGL(k_nys_basis, dim3((m.nt + 127) / 128), dim3(128), 0, s, m, b, g0, g1);
}

// ---------------------------------------------------------------------------
// Green-identity defects: one thread per target node, a slice of the sources
// per workgroup row (partial sums kept per slice, then summed in a fixed
// order), kDefPB basis functions per grid layer.  Source nodes and their
// basis traces are staged in LDS by chunks and read wave-uniformly.
// ---------------------------------------------------------------------------
constexpr int kDefPB = 18;
constexpr int kDefChunk = 64;
constexpr int kDefThreads = 256;

static int def_splits(int nt) {
    const int tb = (nt + kDefThreads - 1) / kDefThreads;
    return std::max(1, std::min(4096 / tb, std::max(1, nt / 128)));
}

size_t nys_defect_parts(int nt, int P) { return (size_t)def_splits(nt) * 2 * (size_t)P * nt; }

template <int FAM>
__global__ __launch_bounds__(kDefThreads) void k_nys_defect(DevMesh m, double k, int P, int per,
                                                           const z64* __restrict__ g0, const z64* __restrict__ g1,
                                                           z64* __restrict__ part) {
    __shared__ double sy[kDefChunk][8];
    __shared__ z64 sg0[kDefPB][kDefChunk], sg1[kDefPB][kDefChunk];
    constexpr int NEED = ((FAM & 1) ? 3 : 0) | ((FAM & 2) ? 12 : 0);
    const int nt = m.nt;
    const int i = blockIdx.x * kDefThreads + threadIdx.x;
    const bool vi = i < nt;
    const int ii = vi ? i : nt - 1;
    const double x0 = m.cen[ii], x1 = m.cen[nt + ii], x2 = m.cen[2 * nt + ii];
    const double nx0 = m.nrm[ii], nx1 = m.nrm[nt + ii], nx2 = m.nrm[2 * nt + ii];
    const int p0 = blockIdx.z * kDefPB;
    const int pb = min(kDefPB, P - p0);
    const int j0 = blockIdx.y * per, j1 = min(nt, j0 + per);
    z64 asd[kDefPB], akh[kDefPB];
#pragma unroll
    for (int q = 0; q < kDefPB; ++q) { asd[q] = z64{0, 0}; akh[q] = z64{0, 0}; }
    for (int c0 = j0; c0 < j1; c0 += kDefChunk) {
        const int cn = min(kDefChunk, j1 - c0);
        __syncthreads();
        for (int t = threadIdx.x; t < cn; t += kDefThreads) {
            const int j = c0 + t;
            sy[t][0] = m.cen[j]; sy[t][1] = m.cen[nt + j]; sy[t][2] = m.cen[2 * nt + j];
            sy[t][3] = m.nrm[j]; sy[t][4] = m.nrm[nt + j]; sy[t][5] = m.nrm[2 * nt + j];
            sy[t][6] = m.area[j];
        }
        // basis values beyond the pass's pb functions are staged as zeros, so
        // the products below run unpredicated
        for (int t = threadIdx.x; t < kDefPB * kDefChunk; t += kDefThreads) {
            const int q = t / kDefChunk, jj = t % kDefChunk;
            const bool ok = q < pb && jj < cn;
            sg0[q][jj] = ok ? g0[(size_t)(p0 + q) * nt + c0 + jj] : z64{0, 0};
            sg1[q][jj] = ok ? g1[(size_t)(p0 + q) * nt + c0 + jj] : z64{0, 0};
        }
        __syncthreads();
        for (int jj = 0; jj < cn; ++jj) {
            if (c0 + jj == i) continue;  // the coincident entry is zero
            const double rx = x0 - sy[jj][0], ry = x1 - sy[jj][1], rz = x2 - sy[jj][2];
            const double ny0 = sy[jj][3], ny1 = sy[jj][4], ny2 = sy[jj][5];
            const double c = nx0 * ny0 + nx1 * ny1 + nx2 * ny2;
            Acc4<double> v{};
            kpoint<double, NEED>(k, rx, ry, rz, nx0, nx1, nx2, ny0, ny1, ny2, c, sy[jj][6], v);
            // eight FMAs per basis function and family: += S b - D a (or K' b - H a)
#pragma unroll
            for (int q = 0; q < kDefPB; ++q) {
                const z64 a = sg0[q][jj], b = sg1[q][jj];
                if (FAM & 1) {
                    asd[q].x = fma(v.S.x, b.x, asd[q].x);
                    asd[q].x = fma(-v.S.y, b.y, asd[q].x);
                    asd[q].x = fma(-v.D.x, a.x, asd[q].x);
                    asd[q].x = fma(v.D.y, a.y, asd[q].x);
                    asd[q].y = fma(v.S.x, b.y, asd[q].y);
                    asd[q].y = fma(v.S.y, b.x, asd[q].y);
                    asd[q].y = fma(-v.D.x, a.y, asd[q].y);
                    asd[q].y = fma(-v.D.y, a.x, asd[q].y);
                }
                if (FAM & 2) {
                    akh[q].x = fma(v.K.x, b.x, akh[q].x);
                    akh[q].x = fma(-v.K.y, b.y, akh[q].x);
                    akh[q].x = fma(-v.H.x, a.x, akh[q].x);
                    akh[q].x = fma(v.H.y, a.y, akh[q].x);
                    akh[q].y = fma(v.K.x, b.y, akh[q].y);
                    akh[q].y = fma(v.K.y, b.x, akh[q].y);
                    akh[q].y = fma(-v.H.x, a.y, akh[q].y);
                    akh[q].y = fma(-v.H.y, a.x, akh[q].y);
                }
            }
        }
    }
    if (!vi) return;
    const size_t plane = (size_t)P * nt;
#pragma unroll
    for (int q = 0; q < kDefPB; ++q) {
        if (q < pb) {
            const size_t at = (size_t)(p0 + q) * nt + i;
            if (FAM & 1) part[(size_t)(2 * blockIdx.y) * plane + at] = asd[q];
            if (FAM & 2) part[(size_t)(2 * blockIdx.y + 1) * plane + at] = akh[q];
        }
    }
}

__global__ void k_nys_defect_sum(int nt, int P, int splits, int fam, const z64* __restrict__ part,
                                 const z64* __restrict__ g0, const z64* __restrict__ g1, z64* __restrict__ th_sd,
                                 z64* __restrict__ th_kh) {
    const size_t plane = (size_t)P * nt;
    const size_t at = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (at >= plane) return;
    for (int f = 0; f < 2; ++f) {
        if (!(fam & (1 << f))) continue;
        const z64 tr = f == 0 ? g0[at] : g1[at];
        z64 s{-0.5 * tr.x, -0.5 * tr.y};
        for (int sp = 0; sp < splits; ++sp) {
            const z64 v = part[(size_t)(2 * sp + f) * plane + at];
            s.x += v.x;
            s.y += v.y;
        }
        (f == 0 ? th_sd : th_kh)[at] = s;
    }
}

void launch_nys_defect(hipStream_t s, const DevMesh& m, double k, int P, int fam, const z64* g0, const z64* g1,
                       z64* part, z64* th_sd, z64* th_kh) {
    const int nt = m.nt;
    const int splits = def_splits(nt);
    const int per = (nt + splits - 1) / splits;
    dim3 grid((nt + kDefThreads - 1) / kDefThreads, splits, (P + kDefPB - 1) / kDefPB);
    if (fam == 1) hipLaunchKernelGGL(k_nys_defect<1>, grid, dim3(kDefThreads), 0, s, m, k, P, per, g0, g1, part);
    else if (fam == 2) hipLaunchKernelGGL(k_nys_defect<2>, grid, dim3(kDefThreads), 0, s, m, k, P, per, g0, g1, part);
    else hipLaunchKernelGGL(k_nys_defect<3>, grid, dim3(kDefThreads), 0, s, m, k, P, per, g0, g1, part);
    const size_t plane = (size_t)P * nt;
    hipLaunchKernelGGL(k_nys_defect_sum, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, s, nt, P, splits, fam,
                       part, g0, g1, th_sd, th_kh);
}

// ---------------------------------------------------------------------------
// per-element fits: one wave per element.  With T the element's traces
// (2 nq rows: pressures then fluxes; P basis columns) and s_c the row norms,
// the fit of a defect row th solves (T / s)^T w' = th in the minimum-norm
// sense over singular values above 1e-13 of the largest, and w = w' / s
// (src/solvers/bem/quadrature/fluid.jl:130-136).  The SVD is one-sided
// Jacobi (Hestenes) on the taller of A = (T / s)^T and A^H, rows on lanes.
// ---------------------------------------------------------------------------
__device__ inline void wave_sum4(double& a, double& b, double& c, double& d) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
        c += __shfl_xor(c, off, 64);
        d += __shfl_xor(d, off, 64);
    }
}

__device__ inline void wave_sum2(double& a, double& b) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
}

constexpr int kJacobiSweeps = 40;

__global__ __launch_bounds__(64) void k_nys_weights(int nt, int nq, int P, const z64* __restrict__ g0,
                                                    const z64* __restrict__ g1, const z64* __restrict__ th_sd,
                                                    const z64* __restrict__ th_kh, z64* __restrict__ corr_sd,
                                                    z64* __restrict__ corr_kh, z64* __restrict__ pinv) {
    extern __shared__ double lds[];
    const int e = blockIdx.x;
    const int lane = threadIdx.x;
    const int C = 2 * nq;
    const bool tall = P >= C;
    const int mr = tall ? P : C, nc = tall ? C : P;
    z64* B = (z64*)lds;               // mr x nc, column-major
    z64* Q = B + mr * nc;             // nc x nc, column-major
    z64* th = Q + nc * nc;            // P
    z64* coef = th + P;               // nc
    z64* wv = coef + nc;              // C
    double* scale = (double*)(wv + C);  // C
    double* sig2 = scale + C;           // nc
    const int base = e * nq;
    // row norms of the traces
    if (lane < C) {
        const z64* g = lane < nq ? g0 : g1;
        const int node = base + (lane < nq ? lane : lane - nq);
        double s = 0;
        for (int p = 0; p < P; ++p) {
            const z64 v = g[(size_t)p * nt + node];
            s += v.x * v.x + v.y * v.y;
        }
        scale[lane] = fmax(sqrt(s), DBL_EPSILON);
    }
    __syncthreads();
    for (int t = lane; t < P * C; t += 64) {
        const int p = t / C, c = t - p * C;
        const z64 v = (c < nq ? g0 : g1)[(size_t)p * nt + base + (c < nq ? c : c - nq)];
        const double is = 1.0 / scale[c];
        if (tall) B[c * mr + p] = z64{v.x * is, v.y * is};
        else B[p * mr + c] = z64{v.x * is, -v.y * is};
    }
    for (int t = lane; t < nc * nc; t += 64) Q[t] = z64{t % (nc + 1) == 0 ? 1.0 : 0.0, 0.0};
    __syncthreads();
    // one-sided Jacobi: each lane owns row `lane` of B and of Q; a pair is
    // rotated while |b_p^H b_q| exceeds mr * eps of |b_p| |b_q|
    const double tol = mr * DBL_EPSILON;
    for (int sweep = 0; sweep < kJacobiSweeps; ++sweep) {
        bool rotated = false;
        for (int p = 0; p < nc - 1; ++p) {
            for (int q = p + 1; q < nc; ++q) {
                const z64 bp = lane < mr ? B[p * mr + lane] : z64{0, 0};
                const z64 bq = lane < mr ? B[q * mr + lane] : z64{0, 0};
                double al = bp.x * bp.x + bp.y * bp.y, be = bq.x * bq.x + bq.y * bq.y;
                double gr = bp.x * bq.x + bp.y * bq.y, gi = bp.x * bq.y - bp.y * bq.x;  // conj(bp) bq
                wave_sum4(al, be, gr, gi);
                const double g = hypot(gr, gi);
                if (!(g > tol * sqrt(al * be)) || !(g > 0)) continue;
                rotated = true;
                const double zeta = (be - al) / (2.0 * g);
                const double tt = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double cs = 1.0 / sqrt(1.0 + tt * tt), sn = cs * tt;
                const z64 ph{gr / g, -gi / g};  // conj(gamma) / |gamma|
                if (lane < mr) {
                    const z64 eb = zmul(ph, bq);
                    B[p * mr + lane] = z64{cs * bp.x - sn * eb.x, cs * bp.y - sn * eb.y};
                    B[q * mr + lane] = z64{sn * bp.x + cs * eb.x, sn * bp.y + cs * eb.y};
                }
                if (lane < nc) {
                    const z64 qp = Q[p * nc + lane], qq = Q[q * nc + lane];
                    const z64 eq = zmul(ph, qq);
                    Q[p * nc + lane] = z64{cs * qp.x - sn * eq.x, cs * qp.y - sn * eq.y};
                    Q[q * nc + lane] = z64{sn * qp.x + cs * eq.x, sn * qp.y + cs * eq.y};
                }
            }
        }
        if (!rotated) break;
    }
    double smax = 0;
    for (int c = 0; c < nc; ++c) {
        const z64 b = lane < mr ? B[c * mr + lane] : z64{0, 0};
        double s = b.x * b.x + b.y * b.y, z = 0;
        wave_sum2(s, z);
        smax = fmax(smax, s);
        if (lane == 0) sig2[c] = s;
    }
    __syncthreads();
    const double cut = 1e-26 * smax;  // singular values above 1e-13 of the largest
    // the fits: of each target's defect (f = 0, 1), or of the unit defects
    // e_p, which give the pseudo-inverse's columns (f = 2)
    for (int f = 0; f < 3; ++f) {
        const z64* thf = f == 0 ? th_sd : th_kh;
        z64* out = f == 0 ? corr_sd : corr_kh;
        if (f < 2 ? !thf : !pinv) continue;
        for (int a = 0; a < (f < 2 ? nq : P); ++a) {
            const int i = base + a;
            if (lane < P) th[lane] = f < 2 ? thf[(size_t)lane * nt + i] : z64{lane == a ? 1.0 : 0.0, 0.0};
            __syncthreads();
            // coef_c = (u_c^H th) / sigma_c^2, u_c the c-th column of B (tall)
            // or of Q (wide), both of length P
            for (int c = 0; c < nc; ++c) {
                z64 u{0, 0};
                if (lane < P) u = tall ? B[c * mr + lane] : Q[c * nc + lane];
                const z64 tv = lane < P ? th[lane] : z64{0, 0};
                double sr = u.x * tv.x + u.y * tv.y, si = u.x * tv.y - u.y * tv.x;
                wave_sum2(sr, si);
                if (lane == 0) coef[c] = sig2[c] > cut ? z64{sr / sig2[c], si / sig2[c]} : z64{0, 0};
            }
            __syncthreads();
            // w' = sum_c v_c coef_c with v_c the c-th column of Q (tall) or B (wide), length C
            if (lane < C) {
                z64 w{0, 0};
                for (int c = 0; c < nc; ++c) w = zadd(w, zmul(tall ? Q[c * nc + lane] : B[c * mr + lane], coef[c]));
                wv[lane] = zsc(w, 1.0 / scale[lane]);
            }
            __syncthreads();
            if (f == 2) {
                if (lane < C) pinv[((size_t)blockIdx.x * C + lane) * P + a] = wv[lane];
            } else if (lane < nq) {
                const z64 d1 = wv[nq + lane], d2 = wv[lane];
                out[2 * ((size_t)i * nq + lane)] = z64{-d1.x, -d1.y};
                out[2 * ((size_t)i * nq + lane) + 1] = d2;
            }
            __syncthreads();
        }
    }
}

void launch_nys_weights(hipStream_t s, const DevMesh& m, int P, const z64* g0, const z64* g1, const z64* th_sd,
                        const z64* th_kh, z64* corr_sd, z64* corr_kh) {
    const int nq = m.nq, C = 2 * nq;
    const int mr = std::max(P, C), nc = std::min(P, C);
    const size_t bytes = sizeof(z64) * ((size_t)mr * nc + (size_t)nc * nc + P + nc + C) + sizeof(double) * (C + nc);
    hipLaunchKernelGGL(k_nys_weights, dim3(m.nt / nq), dim3(64), bytes, s, m.nt, nq, P, g0, g1, th_sd, th_kh,
                       corr_sd, corr_kh, (z64*)nullptr);
}

void launch_nys_pinv(hipStream_t s, const DevMesh& m, int P, const z64* g0, const z64* g1, z64* pinv) {
    const int nq = m.nq, C = 2 * nq;
    const int mr = std::max(P, C), nc = std::min(P, C);
    const size_t bytes = sizeof(z64) * ((size_t)mr * nc + (size_t)nc * nc + P + nc + C) + sizeof(double) * (C + nc);
    hipLaunchKernelGGL(k_nys_weights, dim3(m.nt / nq), dim3(64), bytes, s, m.nt, nq, P, g0, g1, (const z64*)nullptr,
                       (const z64*)nullptr, (z64*)nullptr, (z64*)nullptr, pinv);
}

// ---------------------------------------------------------------------------
// pressure at points (src/fields/pressure.jl:441-462)
// ---------------------------------------------------------------------------
// Winding numbers over the flat sub-triangles and the reference's region
// rule (pressure.jl:496-501), one wave per point, as the panel path.
__global__ __launch_bounds__(256) void k_nys_winding(DevMesh m, FieldParams fp, int npts,
                                                     const double* __restrict__ pts, int* __restrict__ region,
                                                     double* __restrict__ winding) {
    const int lane = threadIdx.x & 63;
    const int pt = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pt >= npts) return;
    const double x0 = pts[3 * pt], x1 = pts[3 * pt + 1], x2 = pts[3 * pt + 2];
    double om = 0.0;
    for (int j = lane; j < m.nsub; j += 64) om += tri_solid_angle(m.vert + 9 * (size_t)j, x0, x1, x2);
    om = __shfl(wave_sum(om), 0, 64);
    const double w = om * 0.07957747154594767;  // / (4 pi)
    const bool outside_or_on = w < 0.75, inside_or_on = w > 0.25;
    int reg = REGION_NONE;
    if (fp.field == FIELD_INCIDENT) reg = REGION_EXTERIOR;
    else if ((fp.field == FIELD_TOTAL || fp.field == FIELD_SCATTERED) && outside_or_on) reg = REGION_EXTERIOR;
    else if (fp.fluid && (fp.field == FIELD_TOTAL || fp.field == FIELD_INTERIOR) && inside_or_on)
        reg = REGION_INTERIOR;
    if (lane == 0) {
        region[pt] = reg;
        winding[pt] = w;
    }
}

void launch_nys_winding(hipStream_t s, const DevMesh& m, const FieldParams& fp, int npts, const double* pts,
                        int* region, double* winding) {
    hipLaunchKernelGGL(k_nys_winding, dim3((npts + 3) / 4), dim3(256), 0, s, m, fp, npts, pts, region, winding);
}

// One wave per point: the nearest node (lowest index on ties); the layer
// potentials over every node outside its element, accumulating beside them
// the identity's defect on the basis (kPtPB functions per pass); then the
// element's fit from its pseudo-inverse and the corrected contribution of the
// element's own nodes.
constexpr int kPtPB = 16;

__device__ inline void node_densities(const DevMesh& m, const NysPointEval& ev, const z64* p, const z64* q, int j,
                                      z64& u, z64& v) {
    const int nt = m.nt;
    u = q[j];
    v = p[j];
    if (ev.inside) {
        const double ph = ev.k0 * (ev.d[0] * m.cen[j] + ev.d[1] * m.cen[nt + j] + ev.d[2] * m.cen[2 * nt + j]);
        double sn, cs;
        sincos(ph, &sn, &cs);
        const double dn = ev.k0 * (ev.d[0] * m.nrm[j] + ev.d[1] * m.nrm[nt + j] + ev.d[2] * m.nrm[2 * nt + j]);
        v = z64{v.x + cs, v.y + sn};                               // p + p_inc
        u = z64{ev.g * (u.x - dn * sn), ev.g * (u.y + dn * cs)};   // g (q + q_inc)
    }
}

__global__ __launch_bounds__(64) void k_nys_points(DevMesh m, NysPointEval ev, int n, const int* __restrict__ idx,
                                                   const double* __restrict__ pts, const z64* __restrict__ p,
                                                   const z64* __restrict__ q, z64* __restrict__ out) {
    __shared__ z64 th[64], at0[64], at1[64], wv[64];
    const int lane = threadIdx.x;
    const int pt = idx[blockIdx.x];
    const int nt = m.nt, nq = m.nq, P = ev.basis.P, C = 2 * nq;
    const double x[3] = {pts[3 * pt], pts[3 * pt + 1], pts[3 * pt + 2]};
    // nearest node
    double best = INFINITY;
    int bj = nt;
    for (int j = lane; j < nt; j += 64) {
        const double r0 = x[0] - m.cen[j], r1 = x[1] - m.cen[nt + j], r2 = x[2] - m.cen[2 * nt + j];
        const double d2 = r0 * r0 + r1 * r1 + r2 * r2;
        if (d2 < best) { best = d2; bj = j; }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double ob = __shfl_xor(best, off, 64);
        const int oj = __shfl_xor(bj, off, 64);
        if (ob < best || (ob == best && oj < bj)) { best = ob; bj = oj; }
    }
    const bool corr = sqrt(best) <= ev.maxdist;
    const int e0 = corr ? (bj / nq) * nq : -1;  // first node of the excluded element
    const double zero3[3] = {0.0, 0.0, 0.0};
    if (corr && lane == 0) basis_point(ev.basis, x, zero3, at0, at1, 1);
    double sx = 0, sy = 0, dx = 0, dy = 0;
    const int passes = corr ? (P + kPtPB - 1) / kPtPB : 1;
    for (int pass = 0; pass < passes; ++pass) {
        const int p0 = pass * kPtPB, pb = corr ? min(kPtPB, P - p0) : 0;
        z64 acc[kPtPB];
#pragma unroll
        for (int c = 0; c < kPtPB; ++c) acc[c] = z64{0, 0};
        for (int j = lane; j < nt; j += 64) {
            if (corr && j >= e0 && j < e0 + nq) continue;  // the excluded element
            const double r0 = x[0] - m.cen[j], r1 = x[1] - m.cen[nt + j], r2 = x[2] - m.cen[2 * nt + j];
            Acc4<double> a{};
            kpoint<double, 3>(ev.k, r0, r1, r2, 0.0, 0.0, 0.0, m.nrm[j], m.nrm[nt + j], m.nrm[2 * nt + j], 0.0,
                              m.area[j], a);
            if (pass == 0) {
                z64 u, v;
                node_densities(m, ev, p, q, j, u, v);
                sx += a.S.x * u.x - a.S.y * u.y;
                sy += a.S.x * u.y + a.S.y * u.x;
                dx += a.D.x * v.x - a.D.y * v.y;
                dy += a.D.x * v.y + a.D.y * v.x;
            }
#pragma unroll
            for (int c = 0; c < kPtPB; ++c) {
                if (c < pb) {
                    const z64 b0 = ev.g0[(size_t)(p0 + c) * nt + j], b1 = ev.g1[(size_t)(p0 + c) * nt + j];
                    acc[c].x += a.S.x * b1.x - a.S.y * b1.y - (a.D.x * b0.x - a.D.y * b0.y);
                    acc[c].y += a.S.x * b1.y + a.S.y * b1.x - (a.D.x * b0.y + a.D.y * b0.x);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < kPtPB; ++c) {
            if (c < pb) {
                const double rx = wave_sum(acc[c].x), ry = wave_sum(acc[c].y);
                if (lane == 0) th[p0 + c] = z64{rx, ry};
            }
        }
    }
    sx = wave_sum(sx); sy = wave_sum(sy);
    dx = wave_sum(dx); dy = wave_sum(dy);
    z64 val = ev.inside ? z64{sx - dx, sy - dy} : z64{dx - sx, dy - sy};
    if (corr) {
        __syncthreads();
        // the identity's multiplier: -1 inside, 0 outside
        if (lane < P && ev.inside) th[lane] = z64{th[lane].x - at0[lane].x, th[lane].y - at0[lane].y};
        __syncthreads();
        // w = pinv th, lanes over the element's 2 nq trace rows
        if (lane < C) {
            const z64* pe = ev.pinv + ((size_t)(e0 / nq) * C + lane) * P;
            z64 w{0, 0};
            for (int c = 0; c < P; ++c) w = zadd(w, zmul(pe[c], th[c]));
            wv[lane] = w;
        }
        __syncthreads();
        // dS_b = -w[nq + b], dD_b = w[b]: D v - S u gains w_b v_b + w_{nq+b} u_b
        double cx = 0, cy = 0;
        if (lane < nq) {
            z64 u, v;
            node_densities(m, ev, p, q, e0 + lane, u, v);
            const z64 t = zadd(zmul(wv[lane], v), zmul(wv[nq + lane], u));
            cx = t.x;
            cy = t.y;
        }
        cx = wave_sum(cx);
        cy = wave_sum(cy);
        val = ev.inside ? z64{val.x - cx, val.y - cy} : z64{val.x + cx, val.y + cy};
    }
    if (lane == 0) {
        if (ev.add_incident) {
            double sn, cs;
            sincos(ev.k0 * (ev.d[0] * x[0] + ev.d[1] * x[1] + ev.d[2] * x[2]), &sn, &cs);
            val.x += cs;
            val.y += sn;
        }
        out[pt] = val;
    }
}

void launch_nys_points(hipStream_t s, const DevMesh& m, const NysPointEval& ev, int n, const int* idx,
                       const double* pts, const z64* p, const z64* q, z64* out) {
    if (n > 0) hipLaunchKernelGGL(k_nys_points, dim3(n), dim3(64), 0, s, m, ev, n, idx, pts, p, q, out);
}

// ---------------------------------------------------------------------------
// corrections into the system: the combinations of store_entries without the
// identity terms, added to the naive entries
// ---------------------------------------------------------------------------
template <typename R>
__device__ inline void add_to(C2<R>* a, double x, double y) {
    const C2<R> v = *a;
    *a = C2<R>{R((double)v.x + x), R((double)v.y + y)};
}

template <typename R, int KIND>
__global__ void k_nys_apply(int nt, int nq, AsmParams p, const z64* __restrict__ e_sd, const z64* __restrict__ e_kh,
                            const z64* __restrict__ i_sd, const z64* __restrict__ i_kh, C2<R>* __restrict__ out0,
                            C2<R>* __restrict__ out1) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nt * nq) return;
    const int i = (int)(t / nq), b = (int)(t - (int64_t)i * nq);
    const int j = (i / nq) * nq + b;
    const int64_t ld = p.ld;
    z64 eS{0, 0}, eD{0, 0}, eK{0, 0}, eH{0, 0}, fS{0, 0}, fD{0, 0}, fK{0, 0}, fH{0, 0};
    if (e_sd) { eS = e_sd[2 * t]; eD = e_sd[2 * t + 1]; }
    if (e_kh) { eK = e_kh[2 * t]; eH = e_kh[2 * t + 1]; }
    if (i_sd) { fS = i_sd[2 * t]; fD = i_sd[2 * t + 1]; }
    if (i_kh) { fK = i_kh[2 * t]; fH = i_kh[2 * t + 1]; }
    const double ik = 1.0 / p.k, g = p.g, ig = 1.0 / p.g;
    C2<R>* a00 = out0 + (int64_t)i * ld + j;
    if (KIND == ASM_BM) {
        add_to(a00, -eD.x + eH.y * ik, -eD.y - eH.x * ik);
        add_to(out1 + (int64_t)i * ld + j, eS.x - eK.y * ik, eS.y + eK.x * ik);
    } else if (KIND == ASM_CBIE_RS) {
        add_to(a00, -eD.x, -eD.y);
        add_to(out1 + (int64_t)i * ld + j, eS.x, eS.y);
    } else if (KIND == ASM_SD) {
        add_to(a00, eS.x, eS.y);
        add_to(out1 + (int64_t)i * ld + j, eD.x, eD.y);
    } else if (KIND == ASM_MULLER) {
        add_to(a00, -eD.x + fD.x, -eD.y + fD.y);
        add_to(a00 + nt, eS.x - g * fS.x, eS.y - g * fS.y);
        add_to(out0 + (int64_t)(i + nt) * ld + j, -eH.x + fH.x * ig, -eH.y + fH.y * ig);
        add_to(out0 + (int64_t)(i + nt) * ld + j + nt, eK.x - fK.x, eK.y - fK.y);
    } else {
        add_to(a00, -eD.x, -eD.y);
        add_to(a00 + nt, eS.x, eS.y);
        add_to(out0 + (int64_t)(i + nt) * ld + j, fD.x, fD.y);
        add_to(out0 + (int64_t)(i + nt) * ld + j + nt, -g * fS.x, -g * fS.y);
    }
}

template <typename R, int KIND>
static void nys_apply_k(hipStream_t s, const DevMesh& m, const AsmParams& p, const z64* a, const z64* b,
                        const z64* c, const z64* d, C2<R>* o0, C2<R>* o1) {
    const int64_t count = (int64_t)m.nt * m.nq;
    hipLaunchKernelGGL((k_nys_apply<R, KIND>), dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, m.nt, m.nq, p,
                       a, b, c, d, o0, o1);
}

template <typename R>
void launch_nys_apply(hipStream_t s, const DevMesh& m, const AsmParams& p, const z64* e_sd, const z64* e_kh,
                      const z64* i_sd, const z64* i_kh, C2<R>* o0, C2<R>* o1) {
    switch (p.kind) {
    case ASM_BM: nys_apply_k<R, ASM_BM>(s, m, p, e_sd, e_kh, i_sd, i_kh, o0, o1); break;
    case ASM_CBIE_RS: nys_apply_k<R, ASM_CBIE_RS>(s, m, p, e_sd, e_kh, i_sd, i_kh, o0, o1); break;
    case ASM_MULLER: nys_apply_k<R, ASM_MULLER>(s, m, p, e_sd, e_kh, i_sd, i_kh, o0, o1); break;
    case ASM_SD: nys_apply_k<R, ASM_SD>(s, m, p, e_sd, e_kh, i_sd, i_kh, o0, o1); break;
    default: nys_apply_k<R, ASM_FLUID_CBIE>(s, m, p, e_sd, e_kh, i_sd, i_kh, o0, o1); break;
    }
}
template void launch_nys_apply<double>(hipStream_t, const DevMesh&, const AsmParams&, const z64*, const z64*,
                                       const z64*, const z64*, z64*, z64*);
template void launch_nys_apply<float>(hipStream_t, const DevMesh&, const AsmParams&, const z64*, const z64*,
                                      const z64*, const z64*, z32*, z32*);

}  // namespace bem
