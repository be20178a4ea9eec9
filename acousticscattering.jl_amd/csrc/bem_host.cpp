// Host-side mesh generators, topology checks, panel geometry and near lists.
// Compiled with -ffp-contract=off so the geometry and the near-pair predicate
// are reproducible bit for bit across compilers.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <unordered_map>

#include "bem_internal.h"

namespace bem {

void q7_rule(double l[7][3], double w[7]) {
    // Radon (1948) degree-5 rule; Dunavant (1985) tabulates the same points.
    const double s15 = std::sqrt(15.0);
    const double a = (6.0 - s15) / 21.0, b = (6.0 + s15) / 21.0;
    const double wa = (155.0 - s15) / 1200.0, wb = (155.0 + s15) / 1200.0;
    const double L[7][3] = {{1.0 / 3, 1.0 / 3, 1.0 / 3},
                            {a, a, 1 - 2 * a}, {a, 1 - 2 * a, a}, {1 - 2 * a, a, a},
                            {b, b, 1 - 2 * b}, {b, 1 - 2 * b, b}, {1 - 2 * b, b, b}};
    const double W[7] = {9.0 / 40, wa, wa, wa, wb, wb, wb};
    for (int q = 0; q < 7; ++q) {
        for (int c = 0; c < 3; ++c) l[q][c] = L[q][c];
        w[q] = W[q];
    }
}

static void subdivide(const double t[3][3], int level, std::vector<double>& out) {
    if (level == 0) {
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) out.push_back(t[r][c]);
        return;
    }
    double m01[3], m12[3], m20[3];
    for (int c = 0; c < 3; ++c) {
        m01[c] = 0.5 * (t[0][c] + t[1][c]);
        m12[c] = 0.5 * (t[1][c] + t[2][c]);
        m20[c] = 0.5 * (t[2][c] + t[0][c]);
    }
    double ch[4][3][3];
    for (int c = 0; c < 3; ++c) {
        ch[0][0][c] = t[0][c]; ch[0][1][c] = m01[c]; ch[0][2][c] = m20[c];
        ch[1][0][c] = m01[c]; ch[1][1][c] = t[1][c]; ch[1][2][c] = m12[c];
        ch[2][0][c] = m20[c]; ch[2][1][c] = m12[c]; ch[2][2][c] = t[2][c];
        ch[3][0][c] = m12[c]; ch[3][1][c] = m20[c]; ch[3][2][c] = m01[c];
    }
    for (int s = 0; s < 4; ++s) subdivide(ch[s], level - 1, out);
}

void near_rule(std::vector<double>& l3, std::vector<double>& w) {
    double l[7][3], qw[7];
    q7_rule(l, qw);
    std::vector<double> tris;
    const double ref[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    subdivide(ref, kNearLevel, tris);
    const int nsub = (int)(tris.size() / 9);
    l3.assign(3 * nsub * 7, 0.0);
    w.assign(nsub * 7, 0.0);
    for (int s = 0; s < nsub; ++s)
        for (int q = 0; q < 7; ++q) {
            const int p = s * 7 + q;
            for (int c = 0; c < 3; ++c)
                l3[3 * p + c] = l[q][0] * tris[9 * s + c] + l[q][1] * tris[9 * s + 3 + c] +
                                l[q][2] * tris[9 * s + 6 + c];
            w[p] = qw[q] / nsub;
        }
}

void gauss_legendre(int n, double* xs, double* ws) {
    for (int i = 0; i < n; ++i) {
        double x = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2.0 * k - 1) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1; p1 = p2;
            }
            double dp = n * (x * p1 - p0) / (x * x - 1.0);
            double dx = p1 / dp;
            x -= dx;
            if (std::fabs(dx) < 1e-16) break;
        }
        double p0 = 1.0, p1 = x;
        for (int k = 2; k <= n; ++k) {
            double p2 = ((2.0 * k - 1) * x * p1 - (k - 1.0) * p0) / k;
            p0 = p1; p1 = p2;
        }
        double dp = n * (x * p1 - p0) / (x * x - 1.0);
        xs[i] = x;
        ws[i] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
}

// ---------------------------------------------------------------------------
// icosphere: 12 golden-ratio vertices, 20 faces, midpoint subdivision with
// vertices created at the first encounter of each edge, projected to |x| = 1.
// ---------------------------------------------------------------------------
static void normalise3(double* p) {
    double n = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    p[0] = p[0] / n; p[1] = p[1] / n; p[2] = p[2] / n;
}

void icosphere(int level, double ax, double ay, double az, std::vector<double>& nodes,
               std::vector<int64_t>& tri0) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    const double V[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0},
                             {0, -1, t}, {0, 1, t}, {0, -1, -t}, {0, 1, -t},
                             {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    const int F[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11},
                          {1, 5, 9}, {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                          {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9},
                          {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    std::vector<double> v;
    for (int i = 0; i < 12; ++i) {
        double p[3] = {V[i][0], V[i][1], V[i][2]};
        normalise3(p);
        v.insert(v.end(), p, p + 3);
    }
    std::vector<int64_t> f;
    for (int i = 0; i < 20; ++i) f.insert(f.end(), F[i], F[i] + 3);
    for (int lv = 0; lv < level; ++lv) {
        std::unordered_map<uint64_t, int64_t> cache;
        cache.reserve(f.size());
        std::vector<int64_t> nf;
        nf.reserve(f.size() * 4);
        auto mid = [&](int64_t a, int64_t b) -> int64_t {
            uint64_t key = a < b ? ((uint64_t)a << 32) | (uint64_t)b : ((uint64_t)b << 32) | (uint64_t)a;
            auto it = cache.find(key);
            if (it != cache.end()) return it->second;
            double m[3] = {0.5 * (v[3 * a] + v[3 * b]), 0.5 * (v[3 * a + 1] + v[3 * b + 1]),
                           0.5 * (v[3 * a + 2] + v[3 * b + 2])};
            normalise3(m);
            v.insert(v.end(), m, m + 3);
            int64_t idx = (int64_t)(v.size() / 3) - 1;
            cache.emplace(key, idx);
            return idx;
        };
        for (size_t s = 0; s < f.size(); s += 3) {
            int64_t a = f[s], b = f[s + 1], c = f[s + 2];
            int64_t ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
            const int64_t add[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nf.insert(nf.end(), add, add + 12);
        }
        f.swap(nf);
    }
    nodes.resize(v.size());
    for (size_t i = 0; i < v.size(); i += 3) {
        nodes[i] = v[i] * ax;
        nodes[i + 1] = v[i + 1] * ay;
        nodes[i + 2] = v[i + 2] * az;
    }
    tri0 = f;
}

// ---------------------------------------------------------------------------
// topology checks
// ---------------------------------------------------------------------------
int validate_topology(int64_t nv, const double* nodes, int64_t nt, const int64_t* tri,
                      double* resolution, double* volume, std::string& msg) {
    char buf[256];
    if (nt <= 0) { msg = "surface must contain triangles"; return 9; }
    for (int64_t i = 0; i < 3 * nt; ++i)
        if (tri[i] < 0 || tri[i] >= nv) {
            msg = "triangle connectivity contains an invalid node index";
            return 8;
        }
    for (int64_t i = 0; i < 3 * nv; ++i)
        if (!std::isfinite(nodes[i])) { msg = "node coordinates must be finite"; return 10; }
    double extent = 0;
    for (int64_t v = 0; v < nv; ++v) {
        double dx = nodes[3 * v] - nodes[0], dy = nodes[3 * v + 1] - nodes[1],
               dz = nodes[3 * v + 2] - nodes[2];
        extent = std::max(extent, std::sqrt(dx * dx + dy * dy + dz * dz));
    }
    const double eps = 2.220446049250313e-16;
    const double tol = 100 * eps * extent * extent;
    double res = 0, vol = 0;
    const double* o = nodes + 3 * tri[0];
    struct E { int64_t a, b, t; int s; };
    std::vector<E> edges;
    edges.reserve(3 * nt);
    for (int64_t t = 0; t < nt; ++t) {
        int64_t a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
        if (a == b || b == c || a == c) {
            snprintf(buf, sizeof buf, "triangle %lld repeats a node", (long long)t + 1);
            msg = buf;
            return 1;
        }
        const double *A = nodes + 3 * a, *B = nodes + 3 * b, *C = nodes + 3 * c;
        double u[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
        double w[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
        double cr[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
        if (!(std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]) > tol)) {
            snprintf(buf, sizeof buf, "triangle %lld is degenerate", (long long)t + 1);
            msg = buf;
            return 2;
        }
        double ao[3] = {A[0] - o[0], A[1] - o[1], A[2] - o[2]};
        double bo[3] = {B[0] - o[0], B[1] - o[1], B[2] - o[2]};
        double co[3] = {C[0] - o[0], C[1] - o[1], C[2] - o[2]};
        double bc[3] = {bo[1] * co[2] - bo[2] * co[1], bo[2] * co[0] - bo[0] * co[2],
                        bo[0] * co[1] - bo[1] * co[0]};
        vol += (ao[0] * bc[0] + ao[1] * bc[1] + ao[2] * bc[2]) / 6.0;
        const int64_t vs[3] = {a, b, c};
        for (int e = 0; e < 3; ++e) {
            int64_t x = vs[e], y = vs[(e + 1) % 3];
            edges.push_back({std::min(x, y), std::max(x, y), t, x < y ? 1 : -1});
            const double *X = nodes + 3 * x, *Y = nodes + 3 * y;
            double l = std::sqrt((X[0] - Y[0]) * (X[0] - Y[0]) + (X[1] - Y[1]) * (X[1] - Y[1]) +
                                 (X[2] - Y[2]) * (X[2] - Y[2]));
            res = std::max(res, l);
        }
    }
    std::sort(edges.begin(), edges.end(), [](const E& p, const E& q) {
        if (p.a != q.a) return p.a < q.a;
        if (p.b != q.b) return p.b < q.b;
        return p.t < q.t;
    });
    std::vector<int64_t> nb(3 * nt, -1);
    std::vector<int> nbc(nt, 0);
    for (size_t s = 0; s < edges.size();) {
        size_t e = s;
        while (e < edges.size() && edges[e].a == edges[s].a && edges[e].b == edges[s].b) ++e;
        if (e - s != 2) {
            snprintf(buf, sizeof buf, "edge (%lld, %lld) has %lld incident triangles; expected two",
                     (long long)edges[s].a + 1, (long long)edges[s].b + 1, (long long)(e - s));
            msg = buf;
            return 3;
        }
        if (edges[s].s != -edges[s + 1].s) {
            snprintf(buf, sizeof buf, "triangles %lld and %lld have inconsistent orientation",
                     (long long)edges[s].t + 1, (long long)edges[s + 1].t + 1);
            msg = buf;
            return 4;
        }
        int64_t i = edges[s].t, j = edges[s + 1].t;
        if (nbc[i] < 3) nb[3 * i + nbc[i]++] = j;
        if (nbc[j] < 3) nb[3 * j + nbc[j]++] = i;
        s = e;
    }
    {
        std::vector<char> vis(nt, 0);
        std::vector<int64_t> st{0};
        int64_t seen = 0;
        while (!st.empty()) {
            int64_t i = st.back();
            st.pop_back();
            if (vis[i]) continue;
            vis[i] = 1;
            ++seen;
            for (int q = 0; q < nbc[i]; ++q) st.push_back(nb[3 * i + q]);
        }
        if (seen != nt) {
            msg = "supply one connected surface; nested/disjoint regions need a multi-region solve";
            return 5;
        }
    }
    {
        std::vector<int64_t> cnt(nv + 1, 0), fill(nv, 0), vf(3 * nt);
        for (int64_t t = 0; t < nt; ++t)
            for (int e = 0; e < 3; ++e) cnt[tri[3 * t + e] + 1]++;
        for (int64_t v = 0; v < nv; ++v) cnt[v + 1] += cnt[v];
        for (int64_t t = 0; t < nt; ++t)
            for (int e = 0; e < 3; ++e) {
                int64_t v = tri[3 * t + e];
                vf[cnt[v] + fill[v]++] = t;
            }
        for (int64_t v = 0; v < nv; ++v) {
            int64_t s0 = cnt[v], deg = cnt[v + 1] - s0;
            if (deg == 0) continue;
            std::vector<char> in(deg, 0);
            std::vector<int64_t> st{0};
            int64_t got = 0;
            while (!st.empty()) {
                int64_t a = st.back();
                st.pop_back();
                if (in[a]) continue;
                in[a] = 1;
                ++got;
                int64_t ta = vf[s0 + a];
                for (int q = 0; q < nbc[ta]; ++q)
                    for (int64_t b = 0; b < deg; ++b)
                        if (!in[b] && vf[s0 + b] == nb[3 * ta + q]) st.push_back(b);
            }
            if (got != deg) {
                snprintf(buf, sizeof buf, "vertex %lld is nonmanifold", (long long)v + 1);
                msg = buf;
                return 6;
            }
        }
    }
    if (!(vol > 100 * eps * extent * extent * extent)) {
        msg = "surface must enclose positive volume with outward orientation";
        return 7;
    }
    if (resolution) *resolution = res;
    if (volume) *volume = vol;
    msg.clear();
    return 0;
}

void build_geometry(HostMesh& m) {
    const int64_t nt = m.nt;
    m.vert.resize(9 * nt);
    m.cen.resize(3 * nt);
    m.nrm.resize(3 * nt);
    m.area.resize(nt);
    m.hmax.resize(nt);
    m.gauss.resize(21 * nt);
    double l[7][3], w[7];
    q7_rule(l, w);
    for (int64_t t = 0; t < nt; ++t) {
        const double* a = &m.nodes[3 * m.tri[3 * t]];
        const double* b = &m.nodes[3 * m.tri[3 * t + 1]];
        const double* c = &m.nodes[3 * m.tri[3 * t + 2]];
        for (int d = 0; d < 3; ++d) {
            m.vert[9 * t + d] = a[d];
            m.vert[9 * t + 3 + d] = b[d];
            m.vert[9 * t + 6 + d] = c[d];
            m.cen[3 * t + d] = (a[d] + b[d] + c[d]) / 3.0;
        }
        double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        double e3[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
        double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                        e1[0] * e2[1] - e1[1] * e2[0]};
        double len = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        for (int d = 0; d < 3; ++d) m.nrm[3 * t + d] = cr[d] / len;
        m.area[t] = 0.5 * len;
        double l1 = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
        double l2 = std::sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
        double l3 = std::sqrt(e3[0] * e3[0] + e3[1] * e3[1] + e3[2] * e3[2]);
        double hm = l1 > l2 ? l1 : l2;
        m.hmax[t] = hm > l3 ? hm : l3;
        for (int q = 0; q < 7; ++q)
            for (int d = 0; d < 3; ++d)
                m.gauss[3 * (7 * t + q) + d] = l[q][0] * a[d] + l[q][1] * b[d] + l[q][2] * c[d];
    }
}

// Near pairs: j != i with |c_i - c_j|^2 < (kNearEta * hmax_j)^2.  A uniform
// grid of cell size kNearEta * max(hmax) finds candidates in the 27 cells
// around each target; the predicate matches the oracle's operation order.
void build_near_list(HostMesh& m) {
    const int64_t nt = m.nt;
    double hm = 0, lo[3] = {1e300, 1e300, 1e300};
    for (int64_t t = 0; t < nt; ++t) {
        hm = std::max(hm, m.hmax[t]);
        for (int d = 0; d < 3; ++d) lo[d] = std::min(lo[d], m.cen[3 * t + d]);
    }
    const double cell = kNearEta * hm;
    auto key = [&](int64_t x, int64_t y, int64_t z) -> uint64_t {
        return ((uint64_t)(x & 0x1FFFFF) << 42) | ((uint64_t)(y & 0x1FFFFF) << 21) | (uint64_t)(z & 0x1FFFFF);
    };
    std::unordered_map<uint64_t, std::vector<int32_t>> grid;
    grid.reserve(nt);
    std::vector<int64_t> cx(nt), cy(nt), cz(nt);
    for (int64_t t = 0; t < nt; ++t) {
        cx[t] = (int64_t)std::floor((m.cen[3 * t] - lo[0]) / cell);
        cy[t] = (int64_t)std::floor((m.cen[3 * t + 1] - lo[1]) / cell);
        cz[t] = (int64_t)std::floor((m.cen[3 * t + 2] - lo[2]) / cell);
        grid[key(cx[t], cy[t], cz[t])].push_back((int32_t)t);
    }
    m.near_ptr.assign(nt + 1, 0);
    m.near_col.clear();
    std::vector<int32_t> row;
    for (int64_t i = 0; i < nt; ++i) {
        row.clear();
        const double* ci = &m.cen[3 * i];
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    auto it = grid.find(key(cx[i] + dx, cy[i] + dy, cz[i] + dz));
                    if (it == grid.end()) continue;
                    for (int32_t j : it->second) {
                        if (j == i) continue;
                        const double* cj = &m.cen[3 * j];
                        double ddx = ci[0] - cj[0], ddy = ci[1] - cj[1], ddz = ci[2] - cj[2];
                        double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                        double lim = kNearEta * m.hmax[j];
                        if (d2 < lim * lim) row.push_back(j);
                    }
                }
        std::sort(row.begin(), row.end());
        m.near_col.insert(m.near_col.end(), row.begin(), row.end());
        m.near_ptr[i + 1] = (int32_t)m.near_col.size();
    }
}

}  // namespace bem

// ---------------------------------------------------------------------------
// Hierarchical-matrix plan (cluster tree + block tree); rules in bem_internal.h.
// ---------------------------------------------------------------------------
namespace bem {

namespace {

void tight_box(const double* pts, const std::vector<int64_t>& perm, int64_t lo, int64_t hi, HNode& nd) {
    for (int c = 0; c < 3; ++c) {
        nd.bmin[c] = INFINITY;
        nd.bmax[c] = -INFINITY;
    }
    for (int64_t q = lo; q < hi; ++q) {
        const double* x = pts + 3 * perm[q];
        for (int c = 0; c < 3; ++c) {
            nd.bmin[c] = std::min(nd.bmin[c], x[c]);
            nd.bmax[c] = std::max(nd.bmax[c], x[c]);
        }
    }
}

double box_diam(const HNode& a) {
    double s = 0;
    for (int c = 0; c < 3; ++c) s += (a.bmax[c] - a.bmin[c]) * (a.bmax[c] - a.bmin[c]);
    return std::sqrt(s);
}

double box_dist(const HNode& a, const HNode& b) {
    double s = 0;
    for (int c = 0; c < 3; ++c) {
        const double g = std::max({0.0, b.bmin[c] - a.bmax[c], a.bmin[c] - b.bmax[c]});
        s += g * g;
    }
    return std::sqrt(s);
}

}  // namespace

void hmatrix_plan(const double* pts, int64_t n, int leaf, double eta, double near_radius, HPlan& P) {
    P.n = n;
    P.leaf = leaf;
    P.eta = eta;
    P.near_radius = near_radius;
    P.perm.resize(n);
    for (int64_t i = 0; i < n; ++i) P.perm[i] = i;
    P.nodes.clear();
    P.dense.clear();
    P.lowrank.clear();
    HNode root;
    root.lo = 0;
    root.hi = n;
    tight_box(pts, P.perm, 0, n, root);
    P.nodes.push_back(root);
    std::vector<int32_t> stack{0};
    std::vector<int64_t> tmp;
    while (!stack.empty()) {
        const int32_t id = stack.back();
        stack.pop_back();
        const int64_t lo = P.nodes[id].lo, hi = P.nodes[id].hi;
        if (hi - lo <= leaf) continue;
        int axis = 0;
        double ext = -1;
        for (int c = 0; c < 3; ++c) {
            const double e = P.nodes[id].bmax[c] - P.nodes[id].bmin[c];
            if (e > ext) {
                ext = e;
                axis = c;
            }
        }
        const double mid = 0.5 * (P.nodes[id].bmin[axis] + P.nodes[id].bmax[axis]);
        tmp.clear();
        for (int64_t q = lo; q < hi; ++q)
            if (pts[3 * P.perm[q] + axis] < mid) tmp.push_back(P.perm[q]);
        int64_t split = lo + (int64_t)tmp.size();
        for (int64_t q = lo; q < hi; ++q)
            if (!(pts[3 * P.perm[q] + axis] < mid)) tmp.push_back(P.perm[q]);
        std::copy(tmp.begin(), tmp.end(), P.perm.begin() + lo);
        if (split == lo || split == hi) split = lo + (hi - lo) / 2;
        HNode L, R;
        L.lo = lo;
        L.hi = split;
        R.lo = split;
        R.hi = hi;
        tight_box(pts, P.perm, L.lo, L.hi, L);
        tight_box(pts, P.perm, R.lo, R.hi, R);
        const int32_t li = (int32_t)P.nodes.size();
        P.nodes.push_back(L);
        P.nodes.push_back(R);
        P.nodes[id].child[0] = li;
        P.nodes[id].child[1] = li + 1;
        // depth first, left before right
        stack.push_back(li + 1);
        stack.push_back(li);
    }
    P.iperm.resize(n);
    for (int64_t q = 0; q < n; ++q) P.iperm[P.perm[q]] = q;
    // block tree
    std::vector<std::pair<int32_t, int32_t>> bs{{0, 0}};
    while (!bs.empty()) {
        const auto pr = bs.back();
        bs.pop_back();
        const HNode& a = P.nodes[pr.first];
        const HNode& b = P.nodes[pr.second];
        const double dist = box_dist(a, b);
        if (dist > near_radius && std::min(box_diam(a), box_diam(b)) < eta * dist) {
            P.lowrank.push_back({pr.first, pr.second});
            continue;
        }
        const bool la = a.child[0] < 0, lb = b.child[0] < 0;
        if (la && lb) {
            P.dense.push_back({pr.first, pr.second});
        } else if (la) {
            bs.push_back({pr.first, b.child[1]});
            bs.push_back({pr.first, b.child[0]});
        } else if (lb) {
            bs.push_back({a.child[1], pr.second});
            bs.push_back({a.child[0], pr.second});
        } else {
            bs.push_back({a.child[1], b.child[1]});
            bs.push_back({a.child[1], b.child[0]});
            bs.push_back({a.child[0], b.child[1]});
            bs.push_back({a.child[0], b.child[0]});
        }
    }
}

}  // namespace bem
