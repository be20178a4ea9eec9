/*
 * CPU FP64 oracle for the full-3D BEM hot path.  TEST INFRASTRUCTURE ONLY:
 * tests/, the smoke check and the bench's cpu_baseline leg use it as the
 * checker; the product library never links or calls it.
 *
 * It restates, in plain C with OpenMP, the discretisation the GPU library
 * runs: flat triangles, one unknown per panel collocated at the centroid,
 * a 7-point degree-5 triangle rule (Radon 1948 / Dunavant 1985) on regular
 * panel pairs, 4^3-fold midpoint subdivision of the same rule on near pairs,
 * and closed-form radial integration with Gauss-Legendre in the polar angle
 * for the self panel (single layer, and the Hadamard finite part of the
 * hypersingular operator).
 *
 * Reference behaviour followed (paths relative to the reference repo):
 *   kernel convention  G = e^{ikr}/(4 pi r)     docs/src/models/boundary_methods.md:9-13
 *   Burton-Miller rigid/soft systems and RHS    src/solvers/bem/full.jl:68-128,144-153
 *   Mueller / CBIE fluid 2Nx2N systems          src/solvers/bem/full.jl:289-347
 *   CBIE fluid forcing                          src/solvers/bem/full.jl:442-454
 *   plane-wave traces                           src/core/incident.jl:64-76
 *   incidence direction                         src/solvers/bem/full.jl:15-18
 *   far field                                   src/solvers/bem/full.jl:509-519
 *   mesh topology checks                        src/geometry/surface.jl:341-405
 *
 * Parity of the assembled entries is pinned in tests by: the polyhedral
 * Laplace identities sum_j D0_ij = -1/2 and sum_j H0_ij = 0, independent
 * Duffy-transformed quadrature of the self-panel integrals, convergence of
 * the near-field subdivision, and whole-solve agreement with the sphere
 * modal series and the reference's golden amplitudes.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double _Complex zc;

#define FOUR_PI (4.0 * M_PI)
#define NEAR_ETA 2.0   /* pair (i,j) is near when |c_i - c_j| < NEAR_ETA * hmax_j */
#define NEAR_LEVEL 3   /* near pairs: 4^3 = 64 sub-triangles x 7 points       */
#define SELF_GL 20     /* Gauss-Legendre points per edge for the self panel  */

/* ------------------------------------------------------------------------ */
/* quadrature rules                                                          */
/* ------------------------------------------------------------------------ */
static double q7_l[7][3];
static double q7_w[7];
static int nsub_pts;
static double (*sub_l)[3];
static double *sub_w;
static double gl_x[SELF_GL], gl_w[SELF_GL];
static int rules_ready = 0;

static void init_q7(void) {
    /* Radon's 7-point degree-5 rule; weights sum to 1 (multiply by area). */
    double s15 = sqrt(15.0);
    double a = (6.0 - s15) / 21.0, b = (6.0 + s15) / 21.0;
    double wa = (155.0 - s15) / 1200.0, wb = (155.0 + s15) / 1200.0;
    double L[7][3] = {{1.0 / 3, 1.0 / 3, 1.0 / 3},
                      {a, a, 1 - 2 * a}, {a, 1 - 2 * a, a}, {1 - 2 * a, a, a},
                      {b, b, 1 - 2 * b}, {b, 1 - 2 * b, b}, {1 - 2 * b, b, b}};
    double W[7] = {9.0 / 40, wa, wa, wa, wb, wb, wb};
    for (int q = 0; q < 7; ++q) {
        for (int c = 0; c < 3; ++c) q7_l[q][c] = L[q][c];
        q7_w[q] = W[q];
    }
}

/* Midpoint subdivision of the reference triangle (barycentric vertices). */
static int subdivide(const double t[3][3], int level, double (*out)[3][3], int pos) {
    if (level == 0) {
        memcpy(out[pos], t, sizeof(double) * 9);
        return pos + 1;
    }
    double m01[3], m12[3], m20[3];
    for (int c = 0; c < 3; ++c) {
        m01[c] = 0.5 * (t[0][c] + t[1][c]);
        m12[c] = 0.5 * (t[1][c] + t[2][c]);
        m20[c] = 0.5 * (t[2][c] + t[0][c]);
    }
    double c0[3][3], c1[3][3], c2[3][3], c3[3][3];
    for (int c = 0; c < 3; ++c) {
        c0[0][c] = t[0][c]; c0[1][c] = m01[c]; c0[2][c] = m20[c];
        c1[0][c] = m01[c]; c1[1][c] = t[1][c]; c1[2][c] = m12[c];
        c2[0][c] = m20[c]; c2[1][c] = m12[c]; c2[2][c] = t[2][c];
        c3[0][c] = m12[c]; c3[1][c] = m20[c]; c3[2][c] = m01[c];
    }
    pos = subdivide((const double(*)[3])c0, level - 1, out, pos);
    pos = subdivide((const double(*)[3])c1, level - 1, out, pos);
    pos = subdivide((const double(*)[3])c2, level - 1, out, pos);
    pos = subdivide((const double(*)[3])c3, level - 1, out, pos);
    return pos;
}

static void init_gl(void) {
    /* Gauss-Legendre nodes on [-1,1] by Newton iteration on P_n. */
    int n = SELF_GL;
    for (int i = 0; i < n; ++i) {
        double x = cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2.0 * k - 1) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1; p1 = p2;
            }
            dp = n * (x * p1 - p0) / (x * x - 1.0);
            double dx = p1 / dp;
            x -= dx;
            if (fabs(dx) < 1e-16) break;
        }
        double p0 = 1.0, p1 = x;
        for (int k = 2; k <= n; ++k) {
            double p2 = ((2.0 * k - 1) * x * p1 - (k - 1.0) * p0) / k;
            p0 = p1; p1 = p2;
        }
        dp = n * (x * p1 - p0) / (x * x - 1.0);
        gl_x[i] = x;
        gl_w[i] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
}

static void init_rules(void) {
    if (rules_ready) return;
#pragma omp critical(orc_rules)
    {
        if (!rules_ready) {
            init_q7();
            init_gl();
            int nsub = 1;
            for (int l = 0; l < NEAR_LEVEL; ++l) nsub *= 4;
            double (*tris)[3][3] = malloc(sizeof(double) * 9 * nsub);
            double ref[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
            subdivide((const double(*)[3])ref, NEAR_LEVEL, tris, 0);
            nsub_pts = nsub * 7;
            sub_l = malloc(sizeof(double) * 3 * nsub_pts);
            sub_w = malloc(sizeof(double) * nsub_pts);
            for (int s = 0; s < nsub; ++s)
                for (int q = 0; q < 7; ++q) {
                    int p = s * 7 + q;
                    for (int c = 0; c < 3; ++c)
                        sub_l[p][c] = q7_l[q][0] * tris[s][0][c] + q7_l[q][1] * tris[s][1][c] +
                                      q7_l[q][2] * tris[s][2][c];
                    sub_w[p] = q7_w[q] / nsub;
                }
            free(tris);
            rules_ready = 1;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* mesh                                                                      */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t nt;
    double *v;    /* 9*nt: vertex coordinates v0,v1,v2 per panel */
    double *c;    /* 3*nt centroids                               */
    double *n;    /* 3*nt unit outward normals                    */
    double *area; /* nt                                           */
    double *h;    /* nt maximum edge length                       */
} orc_mesh;

void orc_mesh_free(orc_mesh *m) {
    if (!m) return;
    free(m->v); free(m->c); free(m->n); free(m->area); free(m->h);
    free(m);
}

/* nodes: 3 x nv column-major (x,y,z interleaved); tri: 3 x nt, 0-based. */
orc_mesh *orc_mesh_create(int64_t nv, const double *nodes, int64_t nt, const int64_t *tri) {
    init_rules();
    orc_mesh *m = calloc(1, sizeof(orc_mesh));
    m->nt = nt;
    m->v = malloc(sizeof(double) * 9 * nt);
    m->c = malloc(sizeof(double) * 3 * nt);
    m->n = malloc(sizeof(double) * 3 * nt);
    m->area = malloc(sizeof(double) * nt);
    m->h = malloc(sizeof(double) * nt);
    for (int64_t t = 0; t < nt; ++t) {
        const double *a = nodes + 3 * tri[3 * t], *b = nodes + 3 * tri[3 * t + 1],
                     *c = nodes + 3 * tri[3 * t + 2];
        for (int d = 0; d < 3; ++d) {
            m->v[9 * t + d] = a[d];
            m->v[9 * t + 3 + d] = b[d];
            m->v[9 * t + 6 + d] = c[d];
            m->c[3 * t + d] = (a[d] + b[d] + c[d]) / 3.0;
        }
        double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        double e3[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
        double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                        e1[0] * e2[1] - e1[1] * e2[0]};
        double len = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        for (int d = 0; d < 3; ++d) m->n[3 * t + d] = cr[d] / len;
        m->area[t] = 0.5 * len;
        double l1 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
        double l2 = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
        double l3 = sqrt(e3[0] * e3[0] + e3[1] * e3[1] + e3[2] * e3[2]);
        double hm = l1 > l2 ? l1 : l2;
        m->h[t] = hm > l3 ? hm : l3;
    }
    return m;
}

int64_t orc_mesh_size(const orc_mesh *m) { return m->nt; }

void orc_mesh_geometry(const orc_mesh *m, double *c, double *n, double *area, double *h) {
    memcpy(c, m->c, sizeof(double) * 3 * m->nt);
    memcpy(n, m->n, sizeof(double) * 3 * m->nt);
    memcpy(area, m->area, sizeof(double) * m->nt);
    memcpy(h, m->h, sizeof(double) * m->nt);
}

/* Gauss nodes of the 7-point rule: coords 3 x 7nt, weights 7nt (area-scaled). */
void orc_mesh_quadrature(const orc_mesh *m, double *xyz, double *w) {
    for (int64_t t = 0; t < m->nt; ++t)
        for (int q = 0; q < 7; ++q) {
            for (int d = 0; d < 3; ++d)
                xyz[3 * (7 * t + q) + d] = q7_l[q][0] * m->v[9 * t + d] +
                                           q7_l[q][1] * m->v[9 * t + 3 + d] +
                                           q7_l[q][2] * m->v[9 * t + 6 + d];
            w[7 * t + q] = q7_w[q] * m->area[t];
        }
}

static inline int is_near(const orc_mesh *m, int64_t i, int64_t j) {
    const double *ci = m->c + 3 * i, *cj = m->c + 3 * j;
    double dx = ci[0] - cj[0], dy = ci[1] - cj[1], dz = ci[2] - cj[2];
    double d2 = dx * dx + dy * dy + dz * dz;
    double lim = NEAR_ETA * m->h[j];
    return d2 < lim * lim;
}

int orc_is_near(const orc_mesh *m, int64_t i, int64_t j) { return i != j && is_near(m, i, j); }

/* ------------------------------------------------------------------------ */
/* kernels                                                                   */
/* ------------------------------------------------------------------------ */
/* Accumulate w * {G, dG/dn_y, dG/dn_x, d2G/dn_x dn_y} at target x (normal nx),
 * source y (normal ny).  With r = x - y, R = |r|, a = r.nx/R, b = r.ny/R,
 * c = nx.ny:
 *   G       = e^{ikR}/(4 pi R)
 *   dG/dny  = G (1/R - ik) b
 *   dG/dnx  = G (ik - 1/R) a
 *   d2G     = G [ k^2 ab + ab (3ik/R - 3/R^2) + c (1/R^2 - ik/R) ]          */
static inline void kern4(double k, const double *x, const double *nx, const double *y,
                         const double *ny, double w, zc acc[4]) {
    double rx = x[0] - y[0], ry = x[1] - y[1], rz = x[2] - y[2];
    double R = sqrt(rx * rx + ry * ry + rz * rz);
    double iR = 1.0 / R;
    double a = (rx * nx[0] + ry * nx[1] + rz * nx[2]) * iR;
    double b = (rx * ny[0] + ry * ny[1] + rz * ny[2]) * iR;
    double c = nx[0] * ny[0] + nx[1] * ny[1] + nx[2] * ny[2];
    double s = sin(k * R), co = cos(k * R);
    zc G = (co + I * s) * (w * iR / FOUR_PI);
    zc ik = I * k;
    acc[0] += G;
    acc[1] += G * (iR - ik) * b;
    acc[2] += G * (ik - iR) * a;
    acc[3] += G * (k * k * a * b + a * b * (3.0 * ik * iR - 3.0 * iR * iR) + c * (iR * iR - ik * iR));
}

/* Self panel: x at the centroid of panel t, in its plane.  In polar
 * coordinates about x, R(theta) = d / cos(theta - theta_e) per edge and
 *   S_tt = (1/4pi) sum_e int (e^{ikR} - 1)/(ik) dtheta
 *   H_tt = (1/4pi) [ 2 pi i k - sum_e int e^{ikR}/R dtheta ]   (Hadamard f.p.)
 *   D_tt = K'_tt = 0  (flat panel).                                          */
static void self_terms(const orc_mesh *m, int64_t t, double k, zc *S, zc *H) {
    const double *x = m->c + 3 * t;
    zc sS = 0, sH = 0;
    for (int e = 0; e < 3; ++e) {
        const double *P = m->v + 9 * t + 3 * e, *Q = m->v + 9 * t + 3 * ((e + 1) % 3);
        double tv[3] = {Q[0] - P[0], Q[1] - P[1], Q[2] - P[2]};
        double L = sqrt(tv[0] * tv[0] + tv[1] * tv[1] + tv[2] * tv[2]);
        for (int d = 0; d < 3; ++d) tv[d] /= L;
        double px[3] = {P[0] - x[0], P[1] - x[1], P[2] - x[2]};
        double sP = px[0] * tv[0] + px[1] * tv[1] + px[2] * tv[2];
        double fx[3] = {px[0] - sP * tv[0], px[1] - sP * tv[1], px[2] - sP * tv[2]};
        double dist = sqrt(fx[0] * fx[0] + fx[1] * fx[1] + fx[2] * fx[2]);
        double sQ = sP + L;
        double th0 = atan2(sP, dist), th1 = atan2(sQ, dist);
        double half = 0.5 * (th1 - th0), mid = 0.5 * (th1 + th0);
        for (int g = 0; g < SELF_GL; ++g) {
            double th = mid + half * gl_x[g];
            double wg = half * gl_w[g];
            double R = dist / cos(th);
            double xr = k * R;
            double sn = sin(xr), cs = cos(xr), sh = sin(0.5 * xr);
            /* (e^{ikR}-1)/(ik) = (sin kR + 2i sin^2(kR/2))/k ; -> R as k -> 0 */
            zc fS = (k > 0) ? (sn + 2.0 * I * sh * sh) / k : R;
            sS += wg * fS;
            sH += wg * (cs + I * sn) / R;
        }
    }
    *S = sS / FOUR_PI;
    *H = (2.0 * M_PI * I * k - sH) / FOUR_PI;
}

/* The four operator entries for target panel i, source panel j. */
static void pair_ops(const orc_mesh *m, int64_t i, int64_t j, double k, zc out[4]) {
    out[0] = out[1] = out[2] = out[3] = 0;
    const double *x = m->c + 3 * i, *nx = m->n + 3 * i, *ny = m->n + 3 * j;
    const double *V = m->v + 9 * j;
    if (i == j) {
        self_terms(m, i, k, &out[0], &out[3]);
        return;
    }
    int near = is_near(m, i, j);
    int np = near ? nsub_pts : 7;
    for (int q = 0; q < np; ++q) {
        const double *l = near ? sub_l[q] : q7_l[q];
        double w = (near ? sub_w[q] : q7_w[q]) * m->area[j];
        double y[3];
        for (int d = 0; d < 3; ++d) y[d] = l[0] * V[d] + l[1] * V[3 + d] + l[2] * V[6 + d];
        kern4(k, x, nx, y, ny, w, out);
    }
}

/* out: 8 doubles = S, D, K', H (complex) for the pair (i, j). */
void orc_pair_ops(const orc_mesh *m, int64_t i, int64_t j, double k, double *out) {
    zc o[4];
    pair_ops(m, i, j, k, o);
    for (int q = 0; q < 4; ++q) {
        out[2 * q] = creal(o[q]);
        out[2 * q + 1] = cimag(o[q]);
    }
}

/* Same pair with an explicit subdivision level (tests: near-field convergence). */
void orc_pair_ops_level(const orc_mesh *m, int64_t i, int64_t j, double k, int level,
                        double *out) {
    init_rules();
    int nsub = 1;
    for (int l = 0; l < level; ++l) nsub *= 4;
    double (*tris)[3][3] = malloc(sizeof(double) * 9 * nsub);
    double ref[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    subdivide((const double(*)[3])ref, level, tris, 0);
    zc acc[4] = {0, 0, 0, 0};
    const double *x = m->c + 3 * i, *nx = m->n + 3 * i, *ny = m->n + 3 * j;
    const double *V = m->v + 9 * j;
    for (int s = 0; s < nsub; ++s)
        for (int q = 0; q < 7; ++q) {
            double l[3];
            for (int c = 0; c < 3; ++c)
                l[c] = q7_l[q][0] * tris[s][0][c] + q7_l[q][1] * tris[s][1][c] +
                       q7_l[q][2] * tris[s][2][c];
            double y[3];
            for (int d = 0; d < 3; ++d) y[d] = l[0] * V[d] + l[1] * V[3 + d] + l[2] * V[6 + d];
            kern4(k, x, nx, y, ny, q7_w[q] / nsub * m->area[j], acc);
        }
    free(tris);
    for (int q = 0; q < 4; ++q) {
        out[2 * q] = creal(acc[q]);
        out[2 * q + 1] = cimag(acc[q]);
    }
}

/* ------------------------------------------------------------------------ */
/* assembly                                                                  */
/* ------------------------------------------------------------------------ */
enum {
    ORC_RIGID_BM = 0,   /* 1/2 I - D - alpha H            (full.jl:103-111)  */
    ORC_SOFT_BM = 1,    /* S + alpha (1/2 I + K')         (full.jl:114-121)  */
    ORC_RIGID_CBIE = 2, /* 1/2 I - D                                         */
    ORC_SOFT_CBIE = 3,  /* S                                                 */
    ORC_FLUID_MULLER = 4,
    ORC_FLUID_CBIE = 5,
    ORC_OPERATOR_S = 6, /* raw operators, for tests                          */
    ORC_OPERATOR_D = 7,
    ORC_OPERATOR_KP = 8,
    ORC_OPERATOR_H = 9
};

/* A: complex column-major, leading dimension lda (complex elements).
 * Fluid kinds fill a 2N x 2N matrix; the others N x N.                      */
int orc_assemble(const orc_mesh *m, int kind, double k, double g, double h, double *Ad,
                 int64_t lda) {
    init_rules();
    if (!(k > 0) || !isfinite(k)) return 1;
    int64_t n = m->nt;
    zc *A = (zc *)Ad;
    zc alpha = I / k;
    double kint = (h > 0) ? k / h : 0.0;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t j = 0; j < n; ++j) {
        for (int64_t i = 0; i < n; ++i) {
            zc e[4];
            pair_ops(m, i, j, k, e);
            double dij = (i == j) ? 1.0 : 0.0;
            switch (kind) {
            case ORC_RIGID_BM: A[i + j * lda] = 0.5 * dij - e[1] - alpha * e[3]; break;
            case ORC_SOFT_BM: A[i + j * lda] = e[0] + alpha * (0.5 * dij + e[2]); break;
            case ORC_RIGID_CBIE: A[i + j * lda] = 0.5 * dij - e[1]; break;
            case ORC_SOFT_CBIE: A[i + j * lda] = e[0]; break;
            case ORC_OPERATOR_S: A[i + j * lda] = e[0]; break;
            case ORC_OPERATOR_D: A[i + j * lda] = e[1]; break;
            case ORC_OPERATOR_KP: A[i + j * lda] = e[2]; break;
            case ORC_OPERATOR_H: A[i + j * lda] = e[3]; break;
            case ORC_FLUID_MULLER:
            case ORC_FLUID_CBIE: {
                zc f[4];
                pair_ops(m, i, j, kint, f);
                if (kind == ORC_FLUID_MULLER) {
                    /* [I-De+Di, Se-gSi; -He+Hi/g, I+K'e-K'i]  full.jl:341 */
                    A[i + j * lda] = dij - e[1] + f[1];
                    A[i + (j + n) * lda] = e[0] - g * f[0];
                    A[(i + n) + j * lda] = -e[3] + f[3] / g;
                    A[(i + n) + (j + n) * lda] = dij + e[2] - f[2];
                } else {
                    /* [1/2 I - De, Se; 1/2 I + Di, -g Si]  full.jl:344 */
                    A[i + j * lda] = 0.5 * dij - e[1];
                    A[i + (j + n) * lda] = e[0];
                    A[(i + n) + j * lda] = 0.5 * dij + f[1];
                    A[(i + n) + (j + n) * lda] = -g * f[0];
                }
            } break;
            default: break;
            }
        }
    }
    return 0;
}

/* Selected rows of the same operators, for checks at sizes whose full matrix
 * is out of reach (a 200k-panel surface): rows[r] indexes the n x n (or
 * 2N x 2N fluid) matrix of orc_assemble; out is row-major, nr x ncols
 * complex.                                                                   */
int orc_assemble_rows(const orc_mesh *m, int kind, double k, double g, double h, int64_t nr,
                      const int64_t *rows, double *outd) {
    init_rules();
    if (!(k > 0) || !isfinite(k)) return 1;
    int64_t n = m->nt;
    int fluid = (kind == ORC_FLUID_MULLER || kind == ORC_FLUID_CBIE);
    int64_t ncols = fluid ? 2 * n : n;
    for (int64_t r = 0; r < nr; ++r)
        if (rows[r] < 0 || rows[r] >= ncols) return 2;
    zc *out = (zc *)outd;
    zc alpha = I / k;
    double kint = (h > 0) ? k / h : 0.0;
#pragma omp parallel for schedule(dynamic, 64) collapse(2)
    for (int64_t r = 0; r < nr; ++r) {
        for (int64_t j = 0; j < n; ++j) {
            int64_t row = rows[r];
            int64_t i = fluid ? row % n : row;
            zc e[4];
            pair_ops(m, i, j, k, e);
            double dij = (i == j) ? 1.0 : 0.0;
            zc *o = out + r * ncols;
            switch (kind) {
            case ORC_RIGID_BM: o[j] = 0.5 * dij - e[1] - alpha * e[3]; break;
            case ORC_SOFT_BM: o[j] = e[0] + alpha * (0.5 * dij + e[2]); break;
            case ORC_RIGID_CBIE: o[j] = 0.5 * dij - e[1]; break;
            case ORC_SOFT_CBIE: case ORC_OPERATOR_S: o[j] = e[0]; break;
            case ORC_OPERATOR_D: o[j] = e[1]; break;
            case ORC_OPERATOR_KP: o[j] = e[2]; break;
            case ORC_OPERATOR_H: o[j] = e[3]; break;
            case ORC_FLUID_MULLER:
            case ORC_FLUID_CBIE: {
                zc f[4];
                pair_ops(m, i, j, kint, f);
                int top = row < n;
                if (kind == ORC_FLUID_MULLER) {
                    o[j] = top ? dij - e[1] + f[1] : -e[3] + f[3] / g;
                    o[j + n] = top ? e[0] - g * f[0] : dij + e[2] - f[2];
                } else {
                    o[j] = top ? 0.5 * dij - e[1] : 0.5 * dij + f[1];
                    o[j + n] = top ? e[0] : -g * f[0];
                }
            } break;
            default: break;
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* pressure at points (src/fields/pressure.jl:404-472)                       */
/* ------------------------------------------------------------------------ */
/* Solid angle of a flat triangle seen from x (Van Oosterom and Strackee,
 * IEEE Trans. Biomed. Eng. 30 (1983) 125-126).                              */
static double solid_angle(const double *V, const double *x) {
    double a[3], b[3], c[3];
    for (int d = 0; d < 3; ++d) {
        a[d] = V[d] - x[d];
        b[d] = V[3 + d] - x[d];
        c[d] = V[6 + d] - x[d];
    }
    double la = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    double lb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    double lc = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    double num = a[0] * (b[1] * c[2] - b[2] * c[1]) + a[1] * (b[2] * c[0] - b[0] * c[2]) +
                 a[2] * (b[0] * c[1] - b[1] * c[0]);
    double den = la * lb * lc + (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) * lc +
                 (a[0] * c[0] + a[1] * c[1] + a[2] * c[2]) * lb +
                 (b[0] * c[0] + b[1] * c[1] + b[2] * c[2]) * la;
    return 2.0 * atan2(num, den);
}

/* field: 0 total, 1 scattered, 2 incident, 3 interior.  region[i]: 0
 * exterior, 1 interior, 2 none (NaN value).  p, q: scattered traces at the
 * centroids; d: incidence direction.  Exterior D p - S q at k (+ incident
 * for the total field); interior S (g (q + q_inc)) - D (p + p_inc) at kint. */
void orc_pressure(const orc_mesh *m, double k, double kint, double g, int fluid, int field,
                  const double *dir, int64_t npts, const double *pts, const double *pd,
                  const double *qd, double *outd, int *region, double *winding) {
    init_rules();
    const zc *p = (const zc *)pd, *q = (const zc *)qd;
    zc *out = (zc *)outd;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t i = 0; i < npts; ++i) {
        const double *x = pts + 3 * i;
        zc inc = cexp(I * k * (dir[0] * x[0] + dir[1] * x[1] + dir[2] * x[2]));
        if (field == 2) {
            out[i] = inc;
            region[i] = 0;
            winding[i] = 0.0;
            continue;
        }
        double om = 0.0;
        for (int64_t j = 0; j < m->nt; ++j) om += solid_angle(m->v + 9 * j, x);
        double w = om / FOUR_PI;
        winding[i] = w;
        int outside_or_on = w < 0.75, inside_or_on = w > 0.25;
        int reg = 2;
        if ((field == 0 || field == 1) && outside_or_on) reg = 0;
        else if (fluid && (field == 0 || field == 3) && inside_or_on) reg = 1;
        region[i] = reg;
        if (reg == 2) {
            out[i] = NAN + NAN * I;
            continue;
        }
        double kk = reg == 1 ? kint : k;
        zc S = 0, D = 0;
        for (int64_t j = 0; j < m->nt; ++j) {
            const double *c = m->c + 3 * j, *ny = m->n + 3 * j, *V = m->v + 9 * j;
            zc u = q[j], v = p[j];
            if (reg == 1) {
                zc pin = cexp(I * k * (dir[0] * c[0] + dir[1] * c[1] + dir[2] * c[2]));
                double dn = k * (dir[0] * ny[0] + dir[1] * ny[1] + dir[2] * ny[2]);
                v = v + pin;
                u = g * (u + I * dn * pin);
            }
            double rx = x[0] - c[0], ry = x[1] - c[1], rz = x[2] - c[2];
            double lim = NEAR_ETA * m->h[j];
            int near = rx * rx + ry * ry + rz * rz < lim * lim;
            int np = near ? nsub_pts : 7;
            zc acc[4] = {0, 0, 0, 0};
            double nx[3] = {0, 0, 0};
            for (int t = 0; t < np; ++t) {
                const double *l = near ? sub_l[t] : q7_l[t];
                double wq = (near ? sub_w[t] : q7_w[t]) * m->area[j];
                double y[3];
                for (int d = 0; d < 3; ++d) y[d] = l[0] * V[d] + l[1] * V[3 + d] + l[2] * V[6 + d];
                kern4(kk, x, nx, y, ny, wq, acc);
            }
            S += acc[0] * u;
            D += acc[1] * v;
        }
        zc val = reg == 1 ? S - D : D - S;
        if (reg == 0 && field == 0) val += inc;
        out[i] = val;
    }
}

/* Polar angle from +x, azimuth from +y toward +z (full.jl:15-18). */
void orc_direction(double beta, double alpha, double *d) {
    d[0] = cos(beta);
    d[1] = sin(beta) * cos(alpha);
    d[2] = sin(beta) * sin(alpha);
}

/* p = e^{ik d.c_i}, q = ik (d.n_i) p at the collocation points (incident.jl:64-76). */
void orc_incident(const orc_mesh *m, double k, const double *d, double *pd, double *qd) {
    zc *p = (zc *)pd, *q = (zc *)qd;
    for (int64_t i = 0; i < m->nt; ++i) {
        const double *c = m->c + 3 * i, *n = m->n + 3 * i;
        double ph = k * (d[0] * c[0] + d[1] * c[1] + d[2] * c[2]);
        zc e = cos(ph) + I * sin(ph);
        p[i] = e;
        q[i] = I * k * (d[0] * n[0] + d[1] * n[1] + d[2] * n[2]) * e;
    }
}

/* f(xhat) = -(1/4pi) sum_j sum_g w_jg (ik (xhat.n_j) p_j + q_j) e^{-ik xhat.y_jg}
 * (full.jl:509-519), with the 7-point rule over each panel. */
void orc_far_field(const orc_mesh *m, double k, int64_t ndir, const double *dirs,
                   const double *pd, const double *qd, double *outd) {
    const zc *p = (const zc *)pd, *q = (const zc *)qd;
    zc *out = (zc *)outd;
    for (int64_t d = 0; d < ndir; ++d) {
        const double *xh = dirs + 3 * d;
        zc total = 0;
        for (int64_t j = 0; j < m->nt; ++j) {
            const double *V = m->v + 9 * j, *n = m->n + 3 * j;
            double xn = xh[0] * n[0] + xh[1] * n[1] + xh[2] * n[2];
            zc coef = I * k * xn * p[j] + q[j];
            zc sum = 0;
            for (int g = 0; g < 7; ++g) {
                double y[3];
                for (int c = 0; c < 3; ++c)
                    y[c] = q7_l[g][0] * V[c] + q7_l[g][1] * V[3 + c] + q7_l[g][2] * V[6 + c];
                double ph = -k * (xh[0] * y[0] + xh[1] * y[1] + xh[2] * y[2]);
                sum += q7_w[g] * (cos(ph) + I * sin(ph));
            }
            total += coef * sum * m->area[j];
        }
        out[d] = -total / FOUR_PI;
    }
}

/* ------------------------------------------------------------------------ */
/* mesh topology checks (restates surface.jl:341-405)                        */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t a, b; /* a < b */
    int64_t tri;
    int sign;
} edge_rec;

static int edge_cmp(const void *x, const void *y) {
    const edge_rec *p = x, *q = y;
    if (p->a != q->a) return p->a < q->a ? -1 : 1;
    if (p->b != q->b) return p->b < q->b ? -1 : 1;
    if (p->tri != q->tri) return p->tri < q->tri ? -1 : 1;
    return 0;
}

/* Returns 0 when valid; otherwise a code and a message.  Codes:
 * 1 repeated node, 2 degenerate, 3 edge count != 2, 4 inconsistent orientation,
 * 5 disconnected, 6 nonmanifold vertex, 7 non-positive volume, 8 bad index.  */
int orc_validate(int64_t nv, const double *nodes, int64_t nt, const int64_t *tri,
                 double *resolution, char *msg, int msglen) {
    if (nt <= 0) { snprintf(msg, msglen, "surface must contain triangles"); return 9; }
    for (int64_t i = 0; i < 3 * nt; ++i)
        if (tri[i] < 0 || tri[i] >= nv) {
            snprintf(msg, msglen, "triangle connectivity contains an invalid node index");
            return 8;
        }
    const double *p0 = nodes;
    double extent = 0;
    for (int64_t v = 0; v < nv; ++v) {
        double dx = nodes[3 * v] - p0[0], dy = nodes[3 * v + 1] - p0[1], dz = nodes[3 * v + 2] - p0[2];
        double e = sqrt(dx * dx + dy * dy + dz * dz);
        if (e > extent) extent = e;
    }
    double tol = 100 * 2.220446049250313e-16 * extent * extent;
    double res = 0, vol = 0;
    const double *o = nodes + 3 * tri[0];
    edge_rec *E = malloc(sizeof(edge_rec) * 3 * nt);
    for (int64_t t = 0; t < nt; ++t) {
        int64_t a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
        if (a == b || b == c || a == c) {
            snprintf(msg, msglen, "triangle %lld repeats a node", (long long)t + 1);
            free(E);
            return 1;
        }
        const double *A = nodes + 3 * a, *B = nodes + 3 * b, *C = nodes + 3 * c;
        double u[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
        double w[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
        double cr[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
        if (!(sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]) > tol)) {
            snprintf(msg, msglen, "triangle %lld is degenerate", (long long)t + 1);
            free(E);
            return 2;
        }
        double ao[3] = {A[0] - o[0], A[1] - o[1], A[2] - o[2]};
        double bo[3] = {B[0] - o[0], B[1] - o[1], B[2] - o[2]};
        double co[3] = {C[0] - o[0], C[1] - o[1], C[2] - o[2]};
        double bc[3] = {bo[1] * co[2] - bo[2] * co[1], bo[2] * co[0] - bo[0] * co[2],
                        bo[0] * co[1] - bo[1] * co[0]};
        vol += (ao[0] * bc[0] + ao[1] * bc[1] + ao[2] * bc[2]) / 6.0;
        int64_t vs[3] = {a, b, c};
        for (int e = 0; e < 3; ++e) {
            int64_t x = vs[e], y = vs[(e + 1) % 3];
            edge_rec r = {x < y ? x : y, x < y ? y : x, t, x < y ? 1 : -1};
            E[3 * t + e] = r;
            const double *X = nodes + 3 * x, *Y = nodes + 3 * y;
            double l = sqrt((X[0] - Y[0]) * (X[0] - Y[0]) + (X[1] - Y[1]) * (X[1] - Y[1]) +
                            (X[2] - Y[2]) * (X[2] - Y[2]));
            if (l > res) res = l;
        }
    }
    qsort(E, 3 * nt, sizeof(edge_rec), edge_cmp);
    /* neighbours: 3 per triangle on a closed manifold */
    int64_t *nb = malloc(sizeof(int64_t) * 3 * nt);
    int *nbc = calloc(nt, sizeof(int));
    int code = 0;
    for (int64_t s = 0; s < 3 * nt && !code;) {
        int64_t e = s;
        while (e < 3 * nt && E[e].a == E[s].a && E[e].b == E[s].b) ++e;
        if (e - s != 2) {
            snprintf(msg, msglen, "edge (%lld, %lld) has %lld incident triangles; expected two",
                     (long long)E[s].a + 1, (long long)E[s].b + 1, (long long)(e - s));
            code = 3;
            break;
        }
        if (E[s].sign != -E[s + 1].sign) {
            snprintf(msg, msglen, "triangles %lld and %lld have inconsistent orientation",
                     (long long)E[s].tri + 1, (long long)E[s + 1].tri + 1);
            code = 4;
            break;
        }
        int64_t i = E[s].tri, j = E[s + 1].tri;
        if (nbc[i] < 3) nb[3 * i + nbc[i]++] = j;
        if (nbc[j] < 3) nb[3 * j + nbc[j]++] = i;
        s = e;
    }
    if (!code) {
        char *vis = calloc(nt, 1);
        int64_t *stack = malloc(sizeof(int64_t) * (3 * nt + 1));
        int64_t sp = 0, seen = 0;
        stack[sp++] = 0;
        while (sp) {
            int64_t i = stack[--sp];
            if (vis[i]) continue;
            vis[i] = 1;
            ++seen;
            for (int q = 0; q < nbc[i]; ++q) stack[sp++] = nb[3 * i + q];
        }
        if (seen != nt) {
            snprintf(msg, msglen, "supply one connected surface; nested/disjoint regions need a multi-region solve");
            code = 5;
        }
        /* vertex fans */
        if (!code) {
            int64_t *cnt = calloc(nv + 1, sizeof(int64_t));
            for (int64_t t = 0; t < nt; ++t)
                for (int e = 0; e < 3; ++e) cnt[tri[3 * t + e] + 1]++;
            for (int64_t v = 0; v < nv; ++v) cnt[v + 1] += cnt[v];
            int64_t *vf = malloc(sizeof(int64_t) * 3 * nt);
            int64_t *fill = calloc(nv, sizeof(int64_t));
            for (int64_t t = 0; t < nt; ++t)
                for (int e = 0; e < 3; ++e) {
                    int64_t v = tri[3 * t + e];
                    vf[cnt[v] + fill[v]++] = t;
                }
            for (int64_t v = 0; v < nv && !code; ++v) {
                int64_t s0 = cnt[v], s1 = cnt[v + 1];
                if (s1 == s0) continue;
                /* flood fill over neighbours restricted to the incident set */
                int64_t deg = s1 - s0;
                char *in = calloc(deg, 1);
                int64_t st[256];
                int64_t ssp = 0, got = 0;
                if (deg > 256) { free(in); continue; }
                st[ssp++] = 0;
                while (ssp) {
                    int64_t a = st[--ssp];
                    if (in[a]) continue;
                    in[a] = 1;
                    ++got;
                    int64_t ta = vf[s0 + a];
                    for (int q = 0; q < nbc[ta]; ++q)
                        for (int64_t b = 0; b < deg; ++b)
                            if (!in[b] && vf[s0 + b] == nb[3 * ta + q] && ssp < 256) st[ssp++] = b;
                }
                free(in);
                if (got != deg) {
                    snprintf(msg, msglen, "vertex %lld is nonmanifold", (long long)v + 1);
                    code = 6;
                }
            }
            free(cnt); free(vf); free(fill);
        }
        free(vis); free(stack);
    }
    if (!code && !(vol > 100 * 2.220446049250313e-16 * extent * extent * extent)) {
        snprintf(msg, msglen, "surface must enclose positive volume with outward orientation");
        code = 7;
    }
    free(E); free(nb); free(nbc);
    if (resolution) *resolution = res;
    if (!code && msglen > 0) msg[0] = 0;
    return code;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
