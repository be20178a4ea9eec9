// Hierarchical-matrix apply and batched GMRES: device structures and launchers.
#pragma once
#include "bem_device.h"

namespace bem {

constexpr int kHRhs = 64;  // right-hand sides per batched apply (stride of T)

// Tile of the low-rank projection: up to kHProjRows consecutive rows of the
// stacked V factors of low-rank blocks that share one column cluster, over
// the columns [col_lo, col_lo + n) of the tree order, which are columns voff
// .. voff + n - 1 of each V row.  Row s of the tile is V row vrow[row0 + s]
// (offset into the factor store).  With part < 0 the tile covers the whole
// cluster and its product lands in T row row0 + s (stride kHRhs); clusters
// longer than kHProjSeg columns are cut into segments, each writing slab
// `part` of P (kHProjRows x kHRhs), and an HSplit sums the nseg slabs
// part0 .. part0 + nseg - 1 into T rows row0 .. row0 + rows - 1.
constexpr int kHProjRows = 64;
constexpr int kHProjSeg = 4096;
struct HTile {
    int64_t row0;
    int rows, col_lo, n, voff, part;
};
struct HSplit {
    int64_t row0;
    int rows, part0, nseg;
};
// A low-rank block covering a leaf row cluster, at row offset o in the block.
struct HRef {
    int blk, o;
};
// Leaf row clusters and, per leaf, its dense blocks and the low-rank blocks
// covering it (CSR), plus its diagonal dense block.
struct HLeafs {
    int count = 0;
    const int* lo = nullptr;
    const int* m = nullptr;
    const int* diag = nullptr;
    const int* dptr = nullptr;
    const int* dlist = nullptr;
    const int* lptr = nullptr;
    const HRef* llist = nullptr;
    const int* order = nullptr;  // leaves by decreasing apply work (launch order)
};

// Y = A X for nr <= kHRhs right-hand sides (tree-position order): the
// projections T = V X of the low-rank blocks (T: kHRhs complex doubles per
// stacked rank row; the segments of long clusters through P and the split
// sums), then per leaf row the dense blocks times X plus U T of the
// low-rank blocks covering it.
void launch_hproj(hipStream_t s, int ntiles, const HTile* tiles, const int64_t* vrow, const z32* uv, const z64* X,
                  int nr, z64* T, int nsplit, const HSplit* splits, z64* P);
void launch_hrows(hipStream_t s, const HLeafs& L, const HBlk* blk, const z32* dval, const z32* uv,
                  const int64_t* toff, const z64* T, const z64* X, int nr, z64* Y);
void launch_hjacobi_inv(hipStream_t s, const HLeafs& L, const HBlk* blk, const z32* dval, z64* minv, int* status);
void launch_hjacobi_apply(hipStream_t s, const HLeafs& L, const z64* minv, const z64* X, int nr, z64* Y);

// out[i][c] = sum_p conj(V_i[p][c]) W[p][c] for i < nb (partial: 512 * nb * nr)
void launch_gm_dots(hipStream_t s, int64_t n, int nr, int nb, const z64* V, int64_t vstride, const z64* W,
                    z64* partial, z64* out);
// Shared basis Q (nb columns of n): out[i][c] = Q_i^H X_c (partial: 512 * nb * nr),
// Y_c += Q coef_c, and out = M w for a small nb x nb row-major M.
void launch_basis_dots(hipStream_t s, int64_t n, int nr, int nb, const z64* Q, const z64* X, z64* partial,
                       z64* out);
void launch_basis_combine(hipStream_t s, int64_t n, int nr, int nb, const z64* Q, const z64* coef, z64* Y);
void launch_small_apply(hipStream_t s, int nb, int nr, const z64* M, const z64* w, z64* out);
// out = a X + b Y + c Z over tot complex entries (Z may be null)
void launch_lincomb3(hipStream_t s, int64_t tot, double a, const z64* X, double b, const z64* Y, double c,
                     const z64* Z, z64* out);
// local 2m x 2m block inverses of the low-frequency Mueller system on [p; q]
void launch_hblock2_apply(hipStream_t s, const HLeafs& L, int64_t nt, const z64* lblk, const z64* X, int nr,
                          z64* Y);
// out = column c of X (n rows, nr columns interleaved)
void launch_gather_col(hipStream_t s, int64_t n, int nr, int c, const z64* X, z64* out);
// W += sign * sum_i V_i coef[i]
void launch_gm_combine(hipStream_t s, int64_t n, int nr, int nb, const z64* V, int64_t vstride, const z64* coef,
                       double sign, z64* W);
void launch_gm_scale(hipStream_t s, int64_t n, int nr, const z64* W, const double* scale, z64* out);
void launch_gm_sub(hipStream_t s, int64_t tot, const z64* a, const z64* b, z64* out);

// q_inc in tree order (nr directions)
void launch_hqinc(hipStream_t s, const DevMesh& m, double k, int nr, const double* dirs, const int* perm, z64* Q);
void launch_htraces(hipStream_t s, const DevMesh& m, double k, int nr, const double* dirs, const int* perm,
                    const z64* X, z64* P, z64* Q);
// Mueller: [p_inc; q_inc] in tree order over 2 nt unknowns, and the
// scattered traces P, Q (nt x nr) from the total-field solution X
void launch_hfluid_rhs(hipStream_t s, const DevMesh& m, double k, int nr, const double* dirs, const int* perm,
                       z64* B);
void launch_hfluid_traces(hipStream_t s, const DevMesh& m, double k, int nr, const double* dirs, const int* perm,
                          const z64* X, z64* P, z64* Q);
// undo the Mueller ACA scaling (HBlk::sig) of the listed compacted blocks
void launch_hunscale(hipStream_t s, int nblk, const HBlk* blk, const int* list, const int* perm, int nt, z32* uv);
void launch_hcompact(hipStream_t s, int nblk, const HBlk* oldb, const HBlk* newb, const z32* src, z32* dst);

}  // namespace bem
