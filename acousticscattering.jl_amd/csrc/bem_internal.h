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
};

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

}  // namespace bem
