// Device-side types and launcher declarations shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace bem {

// Interleaved complex value (re, im); same layout as Julia ComplexF64/ComplexF32.
template <typename R>
struct alignas(2 * sizeof(R)) C2 {
    R x, y;
};
using z64 = C2<double>;
using z32 = C2<float>;

template <typename R>
__host__ __device__ inline C2<R> cmk(R a, R b) { return C2<R>{a, b}; }
template <typename R>
__host__ __device__ inline C2<R> cadd(C2<R> a, C2<R> b) { return C2<R>{a.x + b.x, a.y + b.y}; }
template <typename R>
__host__ __device__ inline C2<R> csub(C2<R> a, C2<R> b) { return C2<R>{a.x - b.x, a.y - b.y}; }
template <typename R>
__host__ __device__ inline C2<R> cmul(C2<R> a, C2<R> b) {
    return C2<R>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R>
__host__ __device__ inline C2<R> cscale(C2<R> a, R s) { return C2<R>{a.x * s, a.y * s}; }
// 1/z, scaled to avoid overflow in |z|^2 (Smith's method, 1962).
template <typename R>
__host__ __device__ inline C2<R> cinv(C2<R> z) {
    R a = z.x, b = z.y;
    if ((a < 0 ? -a : a) >= (b < 0 ? -b : b)) {
        R r = b / a, d = a + b * r;
        return C2<R>{R(1) / d, -r / d};
    }
    R r = a / b, d = b + a * r;
    return C2<R>{r / d, R(-1) / d};
}
template <typename R>
__host__ __device__ inline R cabs1(C2<R> z) {
    return (z.x < 0 ? -z.x : z.x) + (z.y < 0 ? -z.y : z.y);
}

// Panel geometry resident on the device.  SoA arrays: component d of panel j
// at [d*nt + j]; Gauss point q, component d of panel j at [(3q+d)*nt + j].
struct DevMesh {
    int nt = 0;
    double *cen = nullptr, *nrm = nullptr, *area = nullptr, *gp = nullptr, *vert = nullptr;
    float *cen32 = nullptr, *nrm32 = nullptr, *area32 = nullptr, *gp32 = nullptr;
    int *near_ptr = nullptr, *near_col = nullptr, *near_row = nullptr;
    int64_t nnear = 0;
    // Curved-element Nystrom meshes: nq > 0 quadrature nodes per element
    // (element e owns nodes e*nq ...), cen / nrm / area hold the nodes'
    // coordinates, normals and weights, nt their count; the panel-only
    // arrays above are null except `vert`, which holds nsub flat triangles
    // through every node of each element (the surface of the winding number).
    int nq = 0;
    int nsub = 0;
};

// System kinds written by the assembly kernels.
enum AsmKind : int {
    ASM_BM = 0,          // out0 = 1/2 I - D - a H ; out1 = S + a (1/2 I + K'), a = i/k
    ASM_CBIE_RS = 1,     // out0 = 1/2 I - D       ; out1 = S
    ASM_MULLER = 2,      // out0 = 2N x 2N Mueller system
    ASM_FLUID_CBIE = 3,  // out0 = 2N x 2N CBIE fluid system
    ASM_SD = 4,          // out0 = S, out1 = D (one medium, wavenumber k)
};

struct AsmParams {
    int kind;
    double k, kint, g;
    int64_t ld;  // leading dimension (row-major) of the outputs
};

// ---- bem_assembly.hip ----
int upload_rules();  // constant tables (7-point rule, near rule, Gauss-Legendre)
// Regular 7-point pairs for every (i, j); precision R; row-major outputs.
template <typename R>
void launch_assemble(hipStream_t s, const DevMesh& m, const AsmParams& p, C2<R>* out0, C2<R>* out1);
// FP64 near pairs (subdivided) and self panels, overwriting those entries.
template <typename R>
void launch_near_self(hipStream_t s, const DevMesh& m, const AsmParams& p, C2<R>* out0, C2<R>* out1);
// Plane-wave traces at the centroids for nrhs directions (3 x nrhs doubles).
void launch_incident(hipStream_t s, const DevMesh& m, double k, int nrhs, const double* dirs,
                     z64* p, z64* q);
// Far-field amplitudes: out[r*ndir + d] for nrhs x ndir; partial buffer of
// size nrhs*ndir*kFarBlocks.
constexpr int kFarBlocks = 256;
void launch_far_field(hipStream_t s, const DevMesh& m, double k, int nrhs, int ndir,
                      const double* dirs, const z64* p, const z64* q, z64* partial, z64* out);

// Pressure at points (reference src/fields/pressure.jl:404-472): field codes
// and the region each point was evaluated in.
enum FieldKind : int { FIELD_TOTAL = 0, FIELD_SCATTERED = 1, FIELD_INCIDENT = 2, FIELD_INTERIOR = 3 };
enum FieldRegion : int { REGION_EXTERIOR = 0, REGION_INTERIOR = 1, REGION_NONE = 2 };
struct FieldParams {
    double k, kint, g;   // exterior and interior wavenumbers, density contrast
    double d[3];         // incidence direction of the plane wave
    int field;           // FieldKind
    int fluid;           // 1 when the body has a fluid interior
};
// out[i] for points pts[3i..3i+2]; region[i] = FieldRegion; winding[i] the
// solid-angle winding number of the surface about the point.
void launch_pressure(hipStream_t s, const DevMesh& m, const FieldParams& fp, int npts, const double* pts,
                     const z64* p, const z64* q, z64* out, int* region, double* winding);

// ---- hierarchical matrices (bem_assembly.hip: entries and ACA; bem_hsolve.hip: apply, GMRES) ----
// A block of the compressed operator, in cluster-tree positions.  Dense: m x n
// row-major complex FP32 values at `off`.  Low-rank A ~ U V^T: U (m x rmax) and
// then V (n x rmax), both column-major, at `off`.
struct HBlk {
    int row_lo, m, col_lo, n;
    int rmax, rank;
    int64_t off;
    // Mueller low-rank blocks: ACA runs on T A S, T and S scaling the
    // normal-derivative rows by 1/sig and columns by sig (sig = k + 1/dist),
    // which equalises the quadrants' magnitudes (S ~ 1, D and K' ~ sig,
    // H ~ sig^2), so the relative tolerance holds for every quadrant; the
    // stored factors are unscaled afterwards.
    float sig = 1.f;
};
// Operators of the hierarchical path: for rigid surfaces the system matrix
// and the right-hand-side operator of the scattered-field equations
// (reference src/solvers/bem/full.jl:103-111,144-148); for fluid bodies the
// 2N x 2N Mueller system (full.jl:341), whose unknown I < N is the pressure
// and N + i the normal derivative on panel i.
enum HKind : int {
    HK_BM = 0,        // 1/2 I - D - (i/k) H
    HK_CBIE = 1,      // 1/2 I - D
    HK_BM_RHS = 2,    // S + (i/k)(1/2 I + K')
    HK_CBIE_RHS = 3,  // S
    HK_MULLER = 4,    // [I - De + Di, Se - g Si; -He + Hi / g, I + K'e - K'i]
};
// Wavenumbers and density contrast of an operator (ki and g: Mueller only).
struct HPar {
    double k, ki, g;
};
constexpr int kAcaCap = 256;  // largest rank of one ACA block
void launch_hdense(hipStream_t s, const DevMesh& m, int kind, const HPar& p, const int* perm, int nblk,
                   const HBlk* blk, z32* val);
// near pairs (unknown indices pi, pj) and pairs of unknowns on one panel (si,
// sj) into dense-block slots
// Nystrom meshes (rigid kinds): the pairs of nodes of one element, whose slots
// take the naive entry (zero on the diagonal) plus the identity terms and the
// density-interpolation corrections corr_sd / corr_kh (launch_nys_weights
// layout; corr_kh may be null), in FP64.
// Mueller (HK_MULLER) takes the interior medium's corrections in corr_sd_i /
// corr_kh_i and the unknowns' two quadrants (I >= nt: normal derivative).
void launch_hnys_self(hipStream_t s, const DevMesh& m, int kind, const HPar& p, int64_t nself, const int* si,
                      const int* sj, const int64_t* dst, const z64* corr_sd, const z64* corr_kh,
                      const z64* corr_sd_i, const z64* corr_kh_i, z32* val);
void launch_hnear_self(hipStream_t s, const DevMesh& m, int kind, const HPar& p, int64_t npairs, const int* pi,
                       const int* pj, const int64_t* dst, int64_t nself, const int* si, const int* sj,
                       const int64_t* self_dst, z32* val);
// partial-pivot ACA of the listed low-rank blocks (all when list is null)
// Blocks list[0 .. nbig) (the largest) run on s_big with wider workgroups,
// concurrently with the rest on s (fork / join: events ordering the streams).
void launch_aca(hipStream_t s, hipStream_t s_big, hipEvent_t fork, hipEvent_t join, const DevMesh& m, int kind,
                const HPar& p, const int* perm, int nlist, int nbig, const int* list, HBlk* blk, z32* uv, int* conv,
                double tol);
// ACA blocks with min(m, n) at or above this run on the wide workgroups and
// start with rank capacity kAcaWideCap instead of 64.
constexpr int kAcaWide = 4096, kAcaWideCap = 128;

// ---- bem_nystrom.hip: curved-element Nystrom operators ----------------------
// Interpolation basis of the density-interpolation correction at wavenumber
// k: point sources G(., z_p) (kind 0) or regular solid harmonics times the
// normalised spherical Bessel series (kind 1, degree `degree`).
struct NysBasis {
    int kind = 0, P = 0, degree = 0;
    double k = 0;
    double center[3] = {0, 0, 0}, radius = 1;
    const double* zs = nullptr;  // device, 3 x P (kind 0)
};
// Naive Nystrom entries w_j K(x_i, y_j) (zero at i = j) of the system `p.kind`
// with its identity terms, row-major like launch_assemble.
template <typename R>
void launch_nys_assemble(hipStream_t s, const DevMesh& m, const AsmParams& p, C2<R>* out0, C2<R>* out1);
// Basis traces at the nodes: g0[p * nt + j] (pressure), g1 (normal flux).
void launch_nys_basis(hipStream_t s, const DevMesh& m, const NysBasis& b, z64* g0, z64* g1);
// Green-identity defects of the naive layers at wavenumber k on the basis:
// fam bit 0: th_sd = -1/2 g0 + S g1 - D g0; bit 1: th_kh = -1/2 g1 + K' g1 - H g0
// (P x nt each, [p * nt + i]).  `part` holds nys_defect_parts(nt, P) complex.
size_t nys_defect_parts(int nt, int P);
void launch_nys_defect(hipStream_t s, const DevMesh& m, double k, int P, int fam, const z64* g0, const z64* g1,
                       z64* part, z64* th_sd, z64* th_kh);
// Per element, the scaled minimum-norm fit of each target's defect on the
// element's basis traces: corr_*[(i * nq + b)] = {dS or dK', dD or dH} of
// target node i and source node (i / nq) * nq + b.
void launch_nys_weights(hipStream_t s, const DevMesh& m, int P, const z64* g0, const z64* g1, const z64* th_sd,
                        const z64* th_kh, z64* corr_sd, z64* corr_kh);
// With `pinv` set, launch_nys_weights writes instead each element's scaled
// pseudo-inverse: pinv[(e * 2nq + c) * P + p], the fit w_c of the unit defect
// e_p (w = pinv th).
void launch_nys_pinv(hipStream_t s, const DevMesh& m, int P, const z64* g0, const z64* g1, z64* pinv);
// Pressure of a Nystrom solution (src/fields/pressure.jl:441-462): winding
// numbers and regions of the points (as launch_pressure), then, per region
// and basis, the layer potentials at the points in `idx` with the nearest
// element excluded and recovered by density interpolation.
void launch_nys_winding(hipStream_t s, const DevMesh& m, const FieldParams& fp, int npts, const double* pts,
                        int* region, double* winding);
struct NysPointEval {
    int inside;          // 1: S u - D v with multiplier -1; 0: D v - S u, multiplier 0
    double k;            // wavenumber of the region
    double maxdist;      // corrections apply when the nearest node is this close
    NysBasis basis;      // interpolation basis at k (traces g0, g1 at the nodes; pinv per element)
    const z64* g0;
    const z64* g1;
    const z64* pinv;
    // densities: u multiplies S, v multiplies D.  Exterior: u = q, v = p; interior:
    // v = p + p_inc, u = g (q + q_inc) with the incident wave of direction d, k0
    double g, k0, d[3];
    int add_incident;    // exterior total field
};
void launch_nys_points(hipStream_t s, const DevMesh& m, const NysPointEval& ev, int n, const int* idx,
                       const double* pts, const z64* p, const z64* q, z64* out);
// Adds the corrections of the exterior (e_*) and interior (i_*; Mueller and
// fluid CBIE) media into the element-diagonal blocks of the system.
template <typename R>
void launch_nys_apply(hipStream_t s, const DevMesh& m, const AsmParams& p, const z64* e_sd, const z64* e_kh,
                      const z64* i_sd, const z64* i_kh, C2<R>* out0, C2<R>* out1);

// ---- bem_linalg.hip ----
template <typename R>
void launch_row_stats(hipStream_t s, const C2<R>* A, int64_t lda, int n, double* rows, double* rowsum);
template <typename R>
void launch_col_max(hipStream_t s, const C2<R>* A, int64_t lda, int n, const double* rows, double* cols);
void launch_fix_scale(hipStream_t s, int n, double* v);
template <typename Ri, typename Ro>
void launch_scale_copy(hipStream_t s, const C2<Ri>* A, int64_t lda, int n, const double* rows,
                       const double* cols, C2<Ro>* out, int64_t ldo);
// y = A x (FP64 arithmetic); when b != null also r = b - A x and den = |b| + |A||x|.
template <typename R>
void launch_matvec(hipStream_t s, const C2<R>* A, int64_t lda, int rows, int cols, const z64* x, z64* y,
                   const z64* b, double* den);

// Workspace of the tournament-pivoting panel (one per stream).
// Device bytes of PanelWS::rec: the tournament root's multipliers for 64 steps
// x 256 rows, its 64 pivot-row windows in rows of 128 (complex FP64, the
// larger case) and 64 slots.
constexpr size_t kPanelRecBytes = 16 * (64 * 256 + 64 * 128) + 4 * 64;

struct PanelWS {
    int* status;       // [0] first zero pivot (INT_MAX if none), [1] inconsistent pivot selection,
                       // [2] a triangular-solve wait expired
    int* tickets;      // [max_nodes] merge tickets, zero between launches
    int* node_cnt;     // [max_nodes] rows nominated by each tree node
    int* node_rows;    // [max_nodes][64] their global row indices
    int* root;         // [0] count, [1..65) the panel's pivot rows in order
    int max_blocks;    // largest leaf count (256 rows per leaf)
    int max_nodes;
    void* rec = nullptr;  // the tournament root's record of the diagonal block's factors (kPanelRecBytes)
    unsigned* tri_ctr = nullptr;             // [2] triangular-solve tickets, zero between launches
    unsigned long long* tri_prog = nullptr;  // [2] forward / backward progress words
    unsigned tri_epoch = 0;                  // host: triangular solves launched on this workspace
    long long spin_ticks = 0;                // wall-clock ticks one device-side wait may last
};
// Streams of one LU: main, and an optional side stream with two events for
// the panel look-ahead (side = null: everything on main).
struct LuStreams {
    hipStream_t main = nullptr, side = nullptr;
    hipEvent_t ready = nullptr, panel = nullptr;
};
// ---- axisymmetric path ----------------------------------------------------
// Meridian panels on the device (SoA, n panels): endpoints, midpoints,
// lengths and outward normals (-dz, drho) / L.
struct AxiPanels {
    const double *r1, *z1, *r2, *z2, *rm, *zm, *L, *nr, *nz;
    int n;
};
// Gauss-Legendre rules on [0, 1].
struct AxiRules {
    double x8[8], w8[8], x16[16], w16[16];
};
constexpr int kAxiModes = 8;  // Fourier modes per operator launch
// K[c], V[c] (row-major n x n, c = 0 .. nm - 1) of modes m0 + c; with axial,
// also RK[i][j], RV[i][j]: panel j's mode-0 kernels against e^{ikz} and
// ik n_z e^{ikz} along the panel.
void launch_axi_ops(hipStream_t s, const AxiPanels& P, const AxiRules& R, double k, int m0, int nm, int axial, z64* K,
                    z64* V, z64* RK, z64* RV);
// out (leading dimension ldo) = a I + b X + c Y over an n x n block.
void launch_axi_block(hipStream_t s, int n, z64 a, z64 b, const z64* X, z64 c, const z64* Y, z64* out, int64_t ldo);
void launch_axi_rowsum(hipStream_t s, int n, const z64* M, z64* y);
// Far field of the mode traces (modes 0 .. nm - 1, nm <= 256) at ndir
// directions (theta, phi pairs).
void launch_axi_far(hipStream_t s, const AxiPanels& P, const AxiRules& R, double k, int nm, const z64* Pm,
                    const z64* Dm, int ndir, const double* dirs, z64* out);

// A synthetic n x n matrix for timing factorizations (diagonally weighted).
template <typename R>
void launch_fill_test(hipStream_t s, C2<R>* A, int n);
// Shape of one factorization, chosen per size class by the context's tuner.
struct LuTune {
    int outer = 512;  // outer block width = K of the outer trailing updates (256 or 512)
    int group = 8;    // tile rows per L2 group in the trailing updates' tile order (4, 8 or 16)
    int split = 0;    // complex FP32 only: outer trailing updates as bf16-split 3M products
    int ahead = 1;    // panels of block j + 1 on the side stream during block j's update
    int graph = 0;    // host side: replay the launches as a captured HIP graph (lu_factor ignores it)
};
// Device bytes of the bf16 split planes of an n-unknown factorization
// (L21 and U12 of one outer block: 9 planes each, rows padded to 128).
inline size_t lu_split_bytes(int64_t n) { return (size_t)2 * 9 * 512 * 2 * (size_t)((n + 127) / 128 * 128); }
// Returns 0, 2 (panel workspace too small) or 3 (HIP stream/event failure).
// `split` (lu_split_bytes(n), complex FP32 with tune.split) holds the bf16
// planes of the outer trailing updates.
template <typename R>
int lu_factor(const LuStreams& st, C2<R>* A, int64_t lda, int n, int* ipiv, C2<R>* inv, PanelWS& ws,
              void* prof, const LuTune& tune = LuTune{}, void* split = nullptr);
// x = inv(A) b (one launch; ws.status[2] is raised if a device-side wait
// expires, and x then holds NaN).  b, work and x must not overlap.
template <typename R>
void lu_solve(hipStream_t s, const C2<R>* LU, int64_t lda, int n, const int* perm,
              const C2<R>* inv, const C2<R>* b, C2<R>* work, C2<R>* x, PanelWS& ws, void* prof);

// C = A B (complex, row-major).  C may alias B when M <= 128: a workgroup
// reads all of its B columns before it writes its C tile, and a single
// tile row means no other workgroup reads those columns.
template <typename R>
void gemm_assign(hipStream_t s, const C2<R>* A, int64_t lda, const C2<R>* B, int64_t ldb, C2<R>* C, int64_t ldc,
                 int M, int N, int K);
// X = inv(A) X for P A = L U and ncols right-hand sides (row-major n x ncols,
// leading dimension ldx), in place.
template <typename R>
void lu_solve_many(hipStream_t s, const C2<R>* LU, int64_t lda, int n, const int* ipiv, const C2<R>* inv, C2<R>* X,
                   int64_t ldx, int ncols, void* prof);
// X[i][i] += v for i < n.
void launch_add_diag(hipStream_t s, int n, z64* X, int64_t ldx, double v);
// Bottom block row of the low-frequency Mueller system from the Calderon
// blocks of both media: Me = [K'e | He], Mi = [K'i | Hi] (nt x 2nt each):
// A[nt + i][j] = -He + Hi / g, A[nt + i][nt + j] = d_ij + K'e - K'i.
template <typename R>
void launch_calderon_bottom(hipStream_t s, int nt, const z64* Me, const z64* Mi, double g, C2<R>* A);

// x = inv(A)^H b with the same factors (conjugate-transpose solves).
template <typename R>
void lu_solve_h(hipStream_t s, const C2<R>* LU, int64_t lda, int n, const int* perm, const C2<R>* inv,
                const C2<R>* b, C2<R>* work, C2<R>* x, PanelWS& ws, void* prof);

// Refinement helpers (FP64 vectors).
template <typename R>
void launch_prep_rhs(hipStream_t s, int n, const z64* b, const double* rows, double inv_mag, C2<R>* out);
template <typename R>
void launch_finish(hipStream_t s, int n, const C2<R>* v, double mag, const double* cols, z64* x,
                   int accumulate, int* changed);
// Per-block maxima for backward errors: writes nblk x 5 doubles
// {max|r|, max ratio, max|x|, max|b|, nonfinite flag}.
constexpr int kRedBlocks = 128;
void launch_bwd_stats(hipStream_t s, int n, const z64* r, const double* den, const z64* x,
                      const z64* b, double* out);
void launch_max_abs_scaled(hipStream_t s, int n, const z64* b, const double* rows, double* out);
// Per-block sums of |r|^2, |b|^2, |r/rows|^2, |b/rows|^2 (nblk x 4).
void launch_norm_stats(hipStream_t s, int n, const z64* r, const z64* b, const double* rows,
                       double* out);
void launch_fill_int(hipStream_t s, int* p, int n, int v);
void launch_fill_double(hipStream_t s, double* p, int64_t n, double v);
void launch_copy_z(hipStream_t s, int n, const z64* src, z64* dst, double scale, const z64* sub);

}  // namespace bem
