// Internal declarations shared by the host code and the HIP translation units.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace bem {

// Near-pair rule: (i, j), j != i, is integrated with subdivided quadrature when
// |c_i - c_j| < kNearEta * hmax_j.
constexpr double kNearEta = 2.0;
constexpr int kNearLevel = 3;                 // 4^3 sub-triangles
constexpr int kNearSub = 64;                  // sub-triangles per near pair
constexpr int kNearPts = kNearSub * 7;        // 448 points
constexpr int kSelfGL = 20;                   // Gauss-Legendre points per edge (self panel)

// Host-side panel geometry (FP64), derived from nodes + 0-based triangles.
struct HostMesh {
    int64_t nv = 0, nt = 0;
    std::vector<double> nodes;    // 3*nv
    std::vector<int64_t> tri;     // 3*nt, 0-based
    std::vector<double> vert;     // 9*nt: v0, v1, v2
    std::vector<double> cen;      // 3*nt
    std::vector<double> nrm;      // 3*nt
    std::vector<double> area;     // nt
    std::vector<double> hmax;     // nt
    std::vector<double> gauss;    // 3 * 7 * nt, point q of panel j at 3*(7j+q)
    std::vector<int32_t> near_ptr;  // nt+1 CSR over target rows
    std::vector<int32_t> near_col;  // source panels, ascending per row
    double resolution = 0, volume = 0;
    // Curved-element Nystrom meshes (nq > 0): the unknowns are the quadrature
    // nodes, element e owning nodes e*nq .. e*nq+nq-1; cen, nrm and area hold
    // their coordinates, unit normals and weights, and nt their count.  The
    // nel elements of npt Gmsh-ordered nodes sit in `elem` (0-based).
    int npt = 3, nq = 0, qorder = 0;
    int64_t nel = 0;
    std::vector<int64_t> elem;
    double qcenter[3] = {0, 0, 0}, qradius = 0, qrms = 0;  // node mean, max and area-weighted rms distance
    std::vector<double> sources;                          // 3 x ns point sources of the interpolation basis
};

// ---- curved-element Nystrom quadrature --------------------------------------
// Fully symmetric triangle rule exact to degree >= qorder (1..8): nq
// barycentric triples and weights summing to 1/2, the reference triangle's
// area.  Returns nq, or 0 for an unsupported order.
int triangle_rule(int qorder, std::vector<double>& bary, std::vector<double>& w);
// Lagrange shape functions of Gmsh-ordered triangles (npt = 3, 6, 10) at
// barycentric l[3]: values N[npt] and derivatives along u = l1, v = l2.
void lagrange_shape(int npt, const double* l, double* N, double* Nu, double* Nv);
// Topology checks of curved triangles (src/geometry/surface.jl:341-419):
// node indices, repeated nodes, the corner checks of validate_topology, and
// conforming edge nodes.  `tri0` is npt x nt, 0-based.
int validate_curved(int64_t nv, const double* nodes, int64_t nt, int npt, const int64_t* tri0,
                    double* resolution, double* volume, std::string& msg);
// Quadrature nodes of m.elem (m.nodes, m.npt, m.qorder set): fills nq, nt,
// cen, nrm, area, the size figures and the point sources.  Returns 0, or
// nonzero with msg set for a singular Jacobian.
int build_nystrom(HostMesh& m, std::string& msg);
// Flat triangles through every node of each element (1, 4 or 9 per element,
// oriented like it), 9 doubles each: the surface whose winding number
// classifies field points.
void sub_triangles(const HostMesh& m, std::vector<double>& out);
// Point sources of the general density interpolation: ceil(3 nq) points of a
// golden-angle spiral on the sphere of radius 1.5 x half the diagonal of the
// nodes' bounding box, about its centre.
constexpr double kDimOversample = 3.0, kDimRadius = 1.5;
// Icosphere of `level` with order-2 or order-3 nodes projected to the unit
// sphere and scaled per axis (npt x nt elements, 0-based, Gmsh order).
void curved_icosphere(int level, double ax, double ay, double az, int order, std::vector<double>& nodes,
                      std::vector<int64_t>& tri0);

// Radon 7-point rule: barycentric coordinates and weights (sum 1).
void q7_rule(double l[7][3], double w[7]);
// 448-point subdivided rule (level kNearLevel) on the reference triangle.
void near_rule(std::vector<double>& l3, std::vector<double>& w);
// Gauss-Legendre nodes/weights on [-1, 1].
void gauss_legendre(int n, double* x, double* w);

// Topology checks (restating src/geometry/surface.jl:341-405).  Returns 0 or
// a nonzero code with `msg` set.
int validate_topology(int64_t nv, const double* nodes, int64_t nt, const int64_t* tri0,
                      double* resolution, double* volume, std::string& msg);
void build_geometry(HostMesh& m);
void build_near_list(HostMesh& m);

void icosphere(int level, double ax, double ay, double az, std::vector<double>& nodes,
               std::vector<int64_t>& tri0);

// ---- hierarchical-matrix plan ---------------------------------------------
// Cluster tree: a node covers tree positions [lo, hi); it is split when it
// holds more than `leaf` points, along the longest side of its tight bounding
// box at the box midpoint (stable partition: coordinate < midpoint goes left;
// an empty side falls back to splitting the count in half).  Nodes are
// numbered in creation order (root 0, then depth-first, left before right).
struct HNode {
    int64_t lo = 0, hi = 0;
    int32_t child[2] = {-1, -1};
    double bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
};
// Block tree over (row node, column node) pairs, depth first from (root,
// root): admissible pairs, min(diam) < eta * dist and dist > near_radius,
// become low-rank blocks; pairs of two leaves become dense blocks; otherwise
// the non-leaf side(s) are split (both when both are non-leaves).
struct HBlock {
    int32_t row = 0, col = 0;
};
struct HPlan {
    int64_t n = 0;
    int leaf = 64;
    double eta = 2.0, near_radius = 0.0;
    std::vector<int64_t> perm;   // tree position -> point
    std::vector<int64_t> iperm;  // point -> tree position
    std::vector<HNode> nodes;
    std::vector<HBlock> dense, lowrank;
};
void hmatrix_plan(const double* points, int64_t n, int leaf, double eta, double near_radius, HPlan& plan);

}  // namespace bem
