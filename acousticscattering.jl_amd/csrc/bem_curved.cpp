// Curved-element Nystrom meshes on the host: triangle rules by order, Gmsh
// Lagrange shape functions, the curved topology checks, the quadrature nodes
// and the curved icosphere mesher.
//
// The reference builds an Inti quadrature of order `qorder` on Gmsh's 3-, 6-
// or 10-node triangles (src/geometry/surface.jl:271-339,
// src/geometry/canonical_mesh.jl:38-58); its unknowns are the quadrature
// nodes, numbered element by element.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <unordered_map>

#include "bem_internal.h"

namespace bem {

// ---------------------------------------------------------------------------
// triangle rules (Dunavant 1985, IJNME 21:1129; Radon 1948 for degree 5)
// ---------------------------------------------------------------------------
namespace {
struct Orbit {
    int kind;  // 0 centroid, 1 (a, a, 1-2a), 2 (a, b, 1-a-b) and permutations
    double a, b, w;
};

int rule_degree(int qorder) {
    static const int deg[9] = {0, 1, 2, 4, 4, 5, 6, 8, 8};
    return qorder >= 1 && qorder <= 8 ? deg[qorder] : 0;
}

std::vector<Orbit> orbits(int degree) {
    switch (degree) {
    case 1: return {{0, 0, 0, 1.0}};
    case 2: return {{1, 1.0 / 6.0, 0, 1.0 / 3.0}};
    case 4: return {{1, 0.445948490915965, 0, 0.223381589678011}, {1, 0.091576213509771, 0, 0.109951743655322}};
    case 5: {
        const double s15 = std::sqrt(15.0);
        return {{0, 0, 0, 9.0 / 40.0},
                {1, (6.0 - s15) / 21.0, 0, (155.0 - s15) / 1200.0},
                {1, (6.0 + s15) / 21.0, 0, (155.0 + s15) / 1200.0}};
    }
    case 6:
        return {{1, 0.249286745170910, 0, 0.116786275726379},
                {1, 0.063089014491502, 0, 0.050844906370207},
                {2, 0.053145049844817, 0.310352451033784, 0.082851075618374}};
    case 8:
        return {{0, 0, 0, 0.144315607677787},
                {1, 0.459292588292723, 0, 0.095091634267285},
                {1, 0.170569307751760, 0, 0.103217370534718},
                {1, 0.050547228317031, 0, 0.032458497623198},
                {2, 0.008394777409958, 0.263112829634638, 0.027230314174435}};
    default: return {};
    }
}
}  // namespace

int triangle_rule(int qorder, std::vector<double>& bary, std::vector<double>& w) {
    bary.clear();
    w.clear();
    const int deg = rule_degree(qorder);
    if (!deg) return 0;
    auto put = [&](double l0, double l1, double l2, double wt) {
        bary.push_back(l0);
        bary.push_back(l1);
        bary.push_back(l2);
        w.push_back(0.5 * wt);
    };
    for (const Orbit& o : orbits(deg)) {
        if (o.kind == 0) {
            put(1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, o.w);
        } else if (o.kind == 1) {
            const double c = 1.0 - 2.0 * o.a;
            put(o.a, o.a, c, o.w);
            put(o.a, c, o.a, o.w);
            put(c, o.a, o.a, o.w);
        } else {
            const double a = o.a, b = o.b, c = 1.0 - a - b;
            put(a, b, c, o.w);
            put(a, c, b, o.w);
            put(b, a, c, o.w);
            put(b, c, a, o.w);
            put(c, a, b, o.w);
            put(c, b, a, o.w);
        }
    }
    return (int)w.size();
}

// ---------------------------------------------------------------------------
// Lagrange shape functions, Gmsh order: vertices, the edge nodes along
// 0->1, 1->2, 2->0 (nearer the edge's first vertex first), the interior node
// ---------------------------------------------------------------------------
void lagrange_shape(int npt, const double* l, double* N, double* Nu, double* Nv) {
    // derivatives with respect to (l0, l1, l2) as independent variables,
    // then d/du = d/dl1 - d/dl0 and d/dv = d/dl2 - d/dl0
    double d[10][3] = {};
    const double l0 = l[0], l1 = l[1], l2 = l[2];
    if (npt == 3) {
        N[0] = l0; N[1] = l1; N[2] = l2;
        d[0][0] = 1; d[1][1] = 1; d[2][2] = 1;
    } else if (npt == 6) {
        N[0] = l0 * (2 * l0 - 1); N[1] = l1 * (2 * l1 - 1); N[2] = l2 * (2 * l2 - 1);
        N[3] = 4 * l0 * l1; N[4] = 4 * l1 * l2; N[5] = 4 * l2 * l0;
        d[0][0] = 4 * l0 - 1; d[1][1] = 4 * l1 - 1; d[2][2] = 4 * l2 - 1;
        d[3][0] = 4 * l1; d[3][1] = 4 * l0;
        d[4][1] = 4 * l2; d[4][2] = 4 * l1;
        d[5][2] = 4 * l0; d[5][0] = 4 * l2;
    } else {
        const double L[3] = {l0, l1, l2};
        for (int v = 0; v < 3; ++v) {
            const double x = L[v];
            N[v] = 0.5 * x * (3 * x - 1) * (3 * x - 2);
            d[v][v] = 0.5 * (27 * x * x - 18 * x + 2);
        }
        // node nearer a on edge a-b: 9/2 a b (3a - 1)
        static const int E[6][2] = {{0, 1}, {1, 0}, {1, 2}, {2, 1}, {2, 0}, {0, 2}};
        for (int s = 0; s < 6; ++s) {
            const int ia = E[s][0], ib = E[s][1];
            const double a = L[ia], b = L[ib];
            N[3 + s] = 4.5 * a * b * (3 * a - 1);
            d[3 + s][ia] = 4.5 * b * (6 * a - 1);
            d[3 + s][ib] = 4.5 * a * (3 * a - 1);
        }
        N[9] = 27 * l0 * l1 * l2;
        d[9][0] = 27 * l1 * l2; d[9][1] = 27 * l0 * l2; d[9][2] = 27 * l0 * l1;
    }
    for (int a = 0; a < npt; ++a) {
        Nu[a] = d[a][1] - d[a][0];
        Nv[a] = d[a][2] - d[a][0];
    }
}

// ---------------------------------------------------------------------------
// topology of curved triangles
// ---------------------------------------------------------------------------
int validate_curved(int64_t nv, const double* nodes, int64_t nt, int npt, const int64_t* tri, double* resolution,
                    double* volume, std::string& msg) {
    char buf[256];
    if (nt <= 0) { msg = "surface must contain triangles"; return 9; }
    if (npt != 3 && npt != 6 && npt != 10) {
        msg = "triangles must have 3, 6 or 10 rows in Gmsh ordering";
        return 11;
    }
    for (int64_t i = 0; i < (int64_t)npt * nt; ++i)
        if (tri[i] < 0 || tri[i] >= nv) {
            msg = "triangle connectivity contains an invalid node index";
            return 8;
        }
    for (int64_t t = 0; t < nt; ++t) {
        const int64_t* T = tri + (int64_t)npt * t;
        for (int a = 0; a < npt; ++a)
            for (int b = a + 1; b < npt; ++b)
                if (T[a] == T[b]) {
                    snprintf(buf, sizeof buf, "triangle %lld repeats a node", (long long)t + 1);
                    msg = buf;
                    return 1;
                }
    }
    std::vector<int64_t> corners(3 * (size_t)nt);
    for (int64_t t = 0; t < nt; ++t)
        for (int c = 0; c < 3; ++c) corners[3 * t + c] = tri[(int64_t)npt * t + c];
    int code = validate_topology(nv, nodes, nt, corners.data(), resolution, volume, msg);
    if (code || npt == 3) return code;
    // conforming edge nodes: the middle nodes of an edge, read from its lower
    // to its higher vertex, are the same from both triangles
    const int per = npt == 6 ? 1 : 2;
    std::map<std::pair<int64_t, int64_t>, std::vector<int64_t>> seen;
    for (int64_t t = 0; t < nt; ++t) {
        const int64_t* T = tri + (int64_t)npt * t;
        for (int j = 0; j < 3; ++j) {
            const int64_t u = T[j], v = T[(j + 1) % 3];
            std::vector<int64_t> mid(T + 3 + j * per, T + 3 + (j + 1) * per);
            if (u > v) std::reverse(mid.begin(), mid.end());
            const auto key = std::make_pair(std::min(u, v), std::max(u, v));
            auto it = seen.find(key);
            if (it == seen.end()) {
                seen.emplace(key, mid);
            } else if (it->second != mid) {
                snprintf(buf, sizeof buf, "nonconforming curved nodes at edge (%lld, %lld)", (long long)key.first + 1,
                         (long long)key.second + 1);
                msg = buf;
                return 12;
            }
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// quadrature nodes
// ---------------------------------------------------------------------------
int build_nystrom(HostMesh& m, std::string& msg) {
    std::vector<double> bary, w;
    const int nq = triangle_rule(m.qorder, bary, w);
    if (!nq) {
        msg = "qorder must be between 1 and 8";
        return 1;
    }
    const int npt = m.npt;
    m.nq = nq;
    m.nt = m.nel * nq;
    m.cen.assign(3 * (size_t)m.nt, 0.0);
    m.nrm.assign(3 * (size_t)m.nt, 0.0);
    m.area.assign((size_t)m.nt, 0.0);
    std::vector<double> N(npt * nq), Nu(npt * nq), Nv(npt * nq);
    for (int q = 0; q < nq; ++q) lagrange_shape(npt, &bary[3 * q], &N[npt * q], &Nu[npt * q], &Nv[npt * q]);
    for (int64_t e = 0; e < m.nel; ++e) {
        const int64_t* T = &m.elem[(size_t)npt * e];
        for (int q = 0; q < nq; ++q) {
            double x[3] = {0, 0, 0}, xu[3] = {0, 0, 0}, xv[3] = {0, 0, 0};
            for (int a = 0; a < npt; ++a) {
                const double* P = &m.nodes[3 * T[a]];
                for (int c = 0; c < 3; ++c) {
                    x[c] += N[npt * q + a] * P[c];
                    xu[c] += Nu[npt * q + a] * P[c];
                    xv[c] += Nv[npt * q + a] * P[c];
                }
            }
            const double cr[3] = {xu[1] * xv[2] - xu[2] * xv[1], xu[2] * xv[0] - xu[0] * xv[2],
                                  xu[0] * xv[1] - xu[1] * xv[0]};
            const double J = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
            if (!(J > 0) || !std::isfinite(J)) {
                msg = "surface has a singular quadrature Jacobian";
                return 2;
            }
            const int64_t p = e * nq + q;
            for (int c = 0; c < 3; ++c) {
                m.cen[3 * p + c] = x[c];
                m.nrm[3 * p + c] = cr[c] / J;
            }
            m.area[p] = w[q] * J;
        }
    }
    // size figures of src/solvers/bem/quadrature/fluid.jl:1-7
    double c[3] = {0, 0, 0}, wsum = 0, wd2 = 0, r2 = 0;
    for (int64_t p = 0; p < m.nt; ++p)
        for (int d = 0; d < 3; ++d) c[d] += m.cen[3 * p + d];
    for (int d = 0; d < 3; ++d) c[d] /= (double)m.nt;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t p = 0; p < m.nt; ++p) {
        double d2 = 0;
        for (int d = 0; d < 3; ++d) {
            const double x = m.cen[3 * p + d];
            d2 += (x - c[d]) * (x - c[d]);
            lo[d] = std::min(lo[d], x);
            hi[d] = std::max(hi[d], x);
        }
        r2 = std::max(r2, d2);
        wsum += m.area[p];
        wd2 += m.area[p] * d2;
    }
    for (int d = 0; d < 3; ++d) m.qcenter[d] = c[d];
    m.qradius = std::sqrt(r2);
    m.qrms = std::sqrt(wd2 / wsum);
    // point sources: golden-angle spiral (a Fibonacci lattice)
    const int ns = (int)std::ceil(kDimOversample * nq);
    const double diag = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                                  (hi[2] - lo[2]) * (hi[2] - lo[2]));
    const double R = kDimRadius * 0.5 * diag;
    const double golden = M_PI * (3.0 - std::sqrt(5.0));
    m.sources.assign(3 * (size_t)ns, 0.0);
    for (int i = 0; i < ns; ++i) {
        const double y = ns > 1 ? 1.0 - 2.0 * i / (ns - 1) : 0.0;
        const double rho = std::sqrt(std::max(0.0, 1.0 - y * y));
        const double th = golden * i;
        m.sources[3 * i] = std::cos(th) * rho * R + 0.5 * (lo[0] + hi[0]);
        m.sources[3 * i + 1] = y * R + 0.5 * (lo[1] + hi[1]);
        m.sources[3 * i + 2] = std::sin(th) * rho * R + 0.5 * (lo[2] + hi[2]);
    }
    return 0;
}

void sub_triangles(const HostMesh& m, std::vector<double>& out) {
    static const int T1[1][3] = {{0, 1, 2}};
    static const int T2[4][3] = {{0, 3, 5}, {3, 1, 4}, {5, 4, 2}, {3, 4, 5}};
    // the cubic element's nodes on the lattice (i, j), u = i/3, v = j/3:
    // up triangles (i,j),(i+1,j),(i,j+1) and down ones (i+1,j),(i+1,j+1),(i,j+1)
    static const int T3[9][3] = {{0, 3, 8}, {3, 9, 8}, {3, 4, 9}, {4, 5, 9}, {4, 1, 5},
                                 {8, 9, 7}, {9, 6, 7}, {9, 5, 6}, {7, 6, 2}};
    const int ns = m.npt == 3 ? 1 : m.npt == 6 ? 4 : 9;
    const int(*T)[3] = m.npt == 3 ? T1 : m.npt == 6 ? T2 : T3;
    out.assign(9 * (size_t)ns * m.nel, 0.0);
    for (int64_t e = 0; e < m.nel; ++e) {
        const int64_t* E = &m.elem[(size_t)m.npt * e];
        for (int t = 0; t < ns; ++t)
            for (int v = 0; v < 3; ++v)
                for (int c = 0; c < 3; ++c) out[9 * ((size_t)e * ns + t) + 3 * v + c] = m.nodes[3 * E[T[t][v]] + c];
    }
}

// ---------------------------------------------------------------------------
// curved icosphere: the flat icosphere's vertices and faces, then per face
// (in face order) the nodes of its edges (a,b), (b,c), (c,a) at their first
// encounter, placed from the lower to the higher vertex index (order 2: the
// midpoint; order 3: (2 lo + hi)/3 and (lo + 2 hi)/3), then for order 3 the
// face's interior node ((a+b)+c)/3, all projected to the unit sphere; the
// nodes are scaled per axis last.
// ---------------------------------------------------------------------------
void curved_icosphere(int level, double ax, double ay, double az, int order, std::vector<double>& nodes,
                      std::vector<int64_t>& tri0) {
    std::vector<double> v;
    std::vector<int64_t> f;
    icosphere(level, 1.0, 1.0, 1.0, v, f);
    const int64_t nf = (int64_t)(f.size() / 3);
    auto proj = [&](double x, double y, double z) {
        const double n = std::sqrt(x * x + y * y + z * z);
        v.push_back(x / n);
        v.push_back(y / n);
        v.push_back(z / n);
        return (int64_t)(v.size() / 3) - 1;
    };
    const int npt = order == 1 ? 3 : order == 2 ? 6 : 10;
    tri0.assign((size_t)npt * nf, 0);
    std::unordered_map<uint64_t, std::pair<int64_t, int64_t>> cache;
    auto edge = [&](int64_t a, int64_t b, int64_t* out) {
        const int64_t lo = std::min(a, b), hi = std::max(a, b);
        const uint64_t key = ((uint64_t)lo << 32) | (uint64_t)hi;
        auto it = cache.find(key);
        std::pair<int64_t, int64_t> ids;
        if (it != cache.end()) {
            ids = it->second;
        } else {
            const double A[3] = {v[3 * lo], v[3 * lo + 1], v[3 * lo + 2]};
            const double B[3] = {v[3 * hi], v[3 * hi + 1], v[3 * hi + 2]};
            if (order == 2) {
                ids.first = proj(0.5 * (A[0] + B[0]), 0.5 * (A[1] + B[1]), 0.5 * (A[2] + B[2]));
                ids.second = -1;
            } else {
                ids.first = proj((2.0 * A[0] + B[0]) / 3.0, (2.0 * A[1] + B[1]) / 3.0, (2.0 * A[2] + B[2]) / 3.0);
                ids.second = proj((A[0] + 2.0 * B[0]) / 3.0, (A[1] + 2.0 * B[1]) / 3.0, (A[2] + 2.0 * B[2]) / 3.0);
            }
            cache.emplace(key, ids);
        }
        if (order == 2) {
            out[0] = ids.first;
        } else if (a < b) {
            out[0] = ids.first;
            out[1] = ids.second;
        } else {
            out[0] = ids.second;
            out[1] = ids.first;
        }
    };
    const int per = order - 1;
    for (int64_t t = 0; t < nf; ++t) {
        int64_t* T = &tri0[(size_t)npt * t];
        const int64_t a = f[3 * t], b = f[3 * t + 1], c = f[3 * t + 2];
        T[0] = a; T[1] = b; T[2] = c;
        if (order == 1) continue;
        edge(a, b, T + 3);
        edge(b, c, T + 3 + per);
        edge(c, a, T + 3 + 2 * per);
        if (order == 3) {
            const double x = ((v[3 * a] + v[3 * b]) + v[3 * c]) / 3.0;
            const double y = ((v[3 * a + 1] + v[3 * b + 1]) + v[3 * c + 1]) / 3.0;
            const double z = ((v[3 * a + 2] + v[3 * b + 2]) + v[3 * c + 2]) / 3.0;
            T[9] = proj(x, y, z);
        }
    }
    nodes.resize(v.size());
    for (size_t i = 0; i < v.size(); i += 3) {
        nodes[i] = v[i] * ax;
        nodes[i + 1] = v[i + 1] * ay;
        nodes[i + 2] = v[i + 2] * az;
    }
}

}  // namespace bem
