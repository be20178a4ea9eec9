// Hierarchical-matrix apply and batched GMRES kernels (CDNA4 / gfx950).
//
// The compressed operator (reference: HMatrices.jl blocks built in
// src/solvers/bem/compression/assembly.jl, applied by the batched workspace of
// compression/workspace.jl:3-122) is stored in complex FP32: dense leaf blocks
// and low-rank blocks U V^T.  It is applied to FP64 vectors with FP64
// arithmetic, so the GMRES residual is an FP64 residual of the FP32-stored
// operator.  Vectors hold nr right-hand sides per tree position:
// X[pos * nr + c].  Apply = (1) T_b = V_b^T X_cols(b) for every low-rank
// block, then (2) for every leaf row cluster, the sum of its dense blocks
// times X and of U_b (its rows) times T_b for the low-rank blocks covering it,
// in a fixed order (no atomics: the result is deterministic).
#include <algorithm>

#include "bem_device.h"
#include "bem_hsolve.h"

namespace bem {

// ---------------------------------------------------------------------------
// Both phases run on the FP64 matrix cores (v_mfma_f64_16x16x4f64: A[i][k]
// from lane i + 16k, B[k][j] from lane j + 16k, result slot r of lane l at row
// l/16 + 4r, column l%16), complex as four real products.  FP32 block values
// are widened to FP64 as they are read: FP64 arithmetic throughout.
// ---------------------------------------------------------------------------
typedef double dv4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ dv4 mfma64(double a, double b, dv4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
// Both kernels stage K chunks of KC columns of their two operands in LDS
// (block values as stored, complex FP32; vectors complex FP64) and keep the
// next chunk's loads in flight in registers while the matrix cores work on
// the current one.  A workgroup is 4 waves; wave w owns right-hand-side
// columns 16w..16w+15 and up to four 16-row tiles of the output, so each wave
// holds up to 8 independent accumulators and issues the real and imaginary
// products of consecutive MFMAs into different ones.
//
// LDS strides (64 banks of 4 bytes): a block-operand row is 34 z32 = 68 words,
// so the 16 lanes of one 128-bit store pass (4 rows x 4 quarter rows) and the
// 32 lanes of one 64-bit read pass (16 rows x 2 k) touch distinct banks; a
// vector row is 72 z64 = 288 words (= 32 mod 64), so the 16 lanes of a
// 128-bit pass (8 columns x 2 rows, or 16 columns of one row) do too.
template <int KC>
struct HStage {
    static constexpr int kAPad = KC + 2;  // block-operand row stride (z32)
    static constexpr int kBPad = 72;      // vector row stride (z64)
    z32 sA[64][kAPad];  // sA[row][k]
    z64 sB[KC][kBPad];  // sB[k][rhs]
};

// Registers of one thread's share of a chunk of KC columns (64 x KC block
// values and KC x 64 vector values over 256 threads).
template <int KC>
struct HChunkRegs {
    z32 a[KC / 4];
    z64 b[KC / 4];
};

// Vector operand rows r0 + kb0 + t / (256 / KC), columns t % (256 / KC) +
// (256 / KC) u; rows at or beyond kmax and columns at or beyond nr are zero.
template <int KC>
__device__ __forceinline__ void load_vec(const z64* __restrict__ V, int64_t r0, int stride, int kb0, int kmax,
                                         int nr, int t, z64 (&b)[KC / 4]) {
    constexpr int TPR = 256 / KC;
    const int kb = t / TPR, cb = t % TPR;
    const bool ok = kb0 + kb < kmax;
    const z64* src = V + (r0 + kb0 + kb) * (int64_t)stride + cb;
#pragma unroll
    for (int u = 0; u < KC / 4; ++u) b[u] = (ok && cb + TPR * u < nr) ? src[TPR * u] : z64{0.0, 0.0};
}
template <int KC>
__device__ __forceinline__ void store_vec(HStage<KC>& S, int t, const z64 (&b)[KC / 4]) {
    constexpr int TPR = 256 / KC;
    const int kb = t / TPR, cb = t % TPR;
#pragma unroll
    for (int u = 0; u < KC / 4; ++u) S.sB[kb][cb + TPR * u] = b[u];
}
// Row-major block operand row (a dense leaf row, or one stacked V row) at
// src: columns k0 + (t % 4) * KC / 4 .. + KC / 4 (< kmax) of row t / 4.
template <int KC>
__device__ __forceinline__ void load_rowmajor(const z32* __restrict__ src, bool row_ok, int k0, int kmax, int t,
                                              z32 (&a)[KC / 4]) {
    const int j = (t & 3) * (KC / 4);
#pragma unroll
    for (int u = 0; u < KC / 4; ++u) a[u] = (row_ok && k0 + j + u < kmax) ? src[k0 + j + u] : z32{0.f, 0.f};
}
template <int KC>
__device__ __forceinline__ void store_rowmajor(HStage<KC>& S, int t, const z32 (&a)[KC / 4]) {
    const int i = t >> 2, j = (t & 3) * (KC / 4);
#pragma unroll
    for (int u = 0; u < KC / 4; ++u) S.sA[i][j + u] = a[u];
}
// U factor rows of a leaf (U[l][o + i] = base[l * ld + o + i]): row t % 64
// (< rows), ranks k0 + (t / 64) * KC / 4 .. + KC / 4 (< kmax); stored
// transposed.
template <int KC>
__device__ __forceinline__ void load_colmajor(const z32* __restrict__ base, int ld, int rows, int k0, int kmax,
                                              int t, z32 (&a)[KC / 4]) {
    const int i = t & 63, l = (t >> 6) * (KC / 4);
    const z32* src = base + (int64_t)(k0 + l) * ld + i;
#pragma unroll
    for (int u = 0; u < KC / 4; ++u)
        a[u] = (i < rows && k0 + l + u < kmax) ? src[(int64_t)u * ld] : z32{0.f, 0.f};
}
template <int KC>
__device__ __forceinline__ void store_colmajor(HStage<KC>& S, int t, const z32 (&a)[KC / 4]) {
    const int i = t & 63, l = (t >> 6) * (KC / 4);
#pragma unroll
    for (int u = 0; u < KC / 4; ++u) S.sA[i][l + u] = a[u];
}

// acc[rt] += sA[16 rt + i][k] * sB[k][16 w + j] over the first kn (a
// multiple of 4, <= KC; the same for the whole workgroup) columns of one
// staged chunk, for the row tiles rt < NT.  Columns beyond the operand's end
// are staged as zeros, so kn only skips products of zeros.
template <int NT, int KC>
__device__ __forceinline__ void chunk_products(const HStage<KC>& S, int kn, int w, int li, int lk, dv4 (&cr)[NT],
                                               dv4 (&ci)[NT]) {
#pragma unroll 2
    for (int kk = 0; kk < KC; kk += 4) {
        if (kk < kn) {
            const int k = kk + lk;
            const z64 b = S.sB[k][16 * w + li];
            double ar[NT], ai[NT];
#pragma unroll
            for (int rt = 0; rt < NT; ++rt) {
                const z32 a = S.sA[16 * rt + li][k];
                ar[rt] = a.x;
                ai[rt] = a.y;
            }
#pragma unroll
            for (int rt = 0; rt < NT; ++rt) {
                cr[rt] = mfma64(ar[rt], b.x, cr[rt]);
                ci[rt] = mfma64(ar[rt], b.y, ci[rt]);
            }
#pragma unroll
            for (int rt = 0; rt < NT; ++rt) {
                cr[rt] = mfma64(-ai[rt], b.y, cr[rt]);
                ci[rt] = mfma64(ai[rt], b.x, ci[rt]);
            }
        }
    }
}

template <int NT>
__device__ __forceinline__ void store_tiles(const dv4 (&cr)[NT], const dv4 (&ci)[NT], z64* __restrict__ out,
                                            int64_t row_stride, int rows, int nr, int w, int li, int lk) {
    const int c = 16 * w + li;
#pragma unroll
    for (int rt = 0; rt < NT; ++rt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = 16 * rt + lk + 4 * r;
            if (i < rows && c < nr) out[(int64_t)i * row_stride + c] = z64{cr[rt][r], ci[rt][r]};
        }
}

// multiple of 4 at or above min(KC, kmax - k0)
template <int KC>
__device__ __forceinline__ int chunk_width(int k0, int kmax) {
    return (min(KC, kmax - k0) + 3) & ~3;
}

// (1) Projection T = V X_cols for the stacked V rows of one tile (up to 64
// rows of low-rank blocks that share a column cluster) over its column range
// (the whole cluster, or one segment of a long one): one X chunk in LDS
// serves every stacked row, so X is read once per tile rather than once per
// block.  A segment writes its partial product to its own slab of P, summed
// into T afterwards in segment order.  (A variant in which each workgroup
// streams several tiles, with the next tile's first chunk in flight across
// the boundary, needed 184 VGPRs, two workgroups per CU instead of three,
// and ran 8% slower.)
constexpr int kProjKC = 32;  // K chunk of the projection

template <int NT>
__device__ void hproj_tile(const HTile& tl, const int64_t* __restrict__ vrow, const z32* __restrict__ uv,
                           const z64* __restrict__ X, int nr, z64* __restrict__ T, z64* __restrict__ P,
                           HStage<kProjKC>& S) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int ia = tid >> 2;
    const bool row_ok = ia < tl.rows;
    const z32* vsrc = row_ok ? uv + vrow[tl.row0 + ia] + tl.voff : uv;
    dv4 cr[NT], ci[NT];
#pragma unroll
    for (int q = 0; q < NT; ++q) cr[q] = ci[q] = dv4{0, 0, 0, 0};
    HChunkRegs<kProjKC> R;
    load_rowmajor<kProjKC>(vsrc, row_ok, 0, tl.n, tid, R.a);
    load_vec<kProjKC>(X, tl.col_lo, nr, 0, tl.n, nr, tid, R.b);
    for (int k0 = 0; k0 < tl.n; k0 += kProjKC) {
        __syncthreads();
        store_rowmajor(S, tid, R.a);
        store_vec(S, tid, R.b);
        __syncthreads();
        if (k0 + kProjKC < tl.n) {
            load_rowmajor<kProjKC>(vsrc, row_ok, k0 + kProjKC, tl.n, tid, R.a);
            load_vec<kProjKC>(X, tl.col_lo, nr, k0 + kProjKC, tl.n, nr, tid, R.b);
        }
        chunk_products<NT, kProjKC>(S, chunk_width<kProjKC>(k0, tl.n), w, li, lk, cr, ci);
    }
    z64* out = tl.part < 0 ? T + tl.row0 * kHRhs : P + (int64_t)tl.part * kHProjRows * kHRhs;
    store_tiles<NT>(cr, ci, out, kHRhs, tl.rows, nr, w, li, lk);
}

__global__ __launch_bounds__(256) void k_hproj(const HTile* __restrict__ tiles, const int64_t* __restrict__ vrow,
                                               const z32* __restrict__ uv, const z64* __restrict__ X, int nr,
                                               z64* __restrict__ T, z64* __restrict__ P) {
    const HTile tl = tiles[blockIdx.x];
    __shared__ HStage<kProjKC> S;
    switch ((tl.rows + 15) >> 4) {
    case 1: hproj_tile<1>(tl, vrow, uv, X, nr, T, P, S); break;
    case 2: hproj_tile<2>(tl, vrow, uv, X, nr, T, P, S); break;
    case 3: hproj_tile<3>(tl, vrow, uv, X, nr, T, P, S); break;
    default: hproj_tile<4>(tl, vrow, uv, X, nr, T, P, S); break;
    }
}

// T rows of each split tile = the sum of its segments' slabs, in segment
// order (fixed summation order: deterministic).
__global__ __launch_bounds__(256) void k_hproj_sum(const HSplit* __restrict__ splits, const z64* __restrict__ P,
                                                   int nr, z64* __restrict__ T) {
    const HSplit sp = splits[blockIdx.x];
    for (int idx = threadIdx.x; idx < sp.rows * nr; idx += 256) {
        const int i = idx / nr, c = idx - i * nr;
        const z64* src = P + ((int64_t)sp.part0 * kHProjRows + i) * kHRhs + c;
        z64 acc = src[0];
        for (int q = 1; q < sp.nseg; ++q) {
            const z64 v = src[(int64_t)q * kHProjRows * kHRhs];
            acc.x += v.x;
            acc.y += v.y;
        }
        T[(sp.row0 + i) * kHRhs + c] = acc;
    }
}

// (2) one workgroup per leaf row cluster (leaves in `order`, most work
// first): its dense blocks times X, then U_b T_b for the low-rank blocks
// covering it, chunk after chunk in list order (fixed summation order).
constexpr int kRowsKC = 16;  // K chunk of the row kernel

template <int NT>
__device__ void hrows_leaf(const HLeafs& L, int leaf, const HBlk* __restrict__ blk, const z32* __restrict__ dval,
                           const z32* __restrict__ uv, const int64_t* __restrict__ toff, const z64* __restrict__ T,
                           const z64* __restrict__ X, int nr, z64* __restrict__ Y, HStage<kRowsKC>& S) {
    const int lo = L.lo[leaf], m = L.m[leaf];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int d0 = L.dptr[leaf], nd = L.dptr[leaf + 1] - d0, l0 = L.lptr[leaf];
    const int nitems = nd + (L.lptr[leaf + 1] - l0);
    dv4 cr[NT], ci[NT];
#pragma unroll
    for (int q = 0; q < NT; ++q) cr[q] = ci[q] = dv4{0, 0, 0, 0};
    // item e < nd: dense block dlist[d0 + e]; else low-rank ref llist[l0 + e - nd]
    auto kdim = [&](int e) { return e < nd ? blk[L.dlist[d0 + e]].n : blk[L.llist[l0 + e - nd].blk].rank; };
    auto load = [&](int e, int k0, HChunkRegs<kRowsKC>& R) {
        if (e < nd) {
            const HBlk b = blk[L.dlist[d0 + e]];
            const int ia = tid >> 2;
            load_rowmajor<kRowsKC>(dval + b.off + (int64_t)ia * b.n, ia < m, k0, b.n, tid, R.a);
            load_vec<kRowsKC>(X, b.col_lo, nr, k0, b.n, nr, tid, R.b);
        } else {
            const HRef ref = L.llist[l0 + e - nd];
            const HBlk b = blk[ref.blk];
            load_colmajor<kRowsKC>(uv + b.off + ref.o, b.m, m, k0, b.rank, tid, R.a);
            load_vec<kRowsKC>(T + toff[ref.blk], 0, kHRhs, k0, b.rank, nr, tid, R.b);
        }
    };
    int e = 0, k0 = 0, kmax = 0;
    while (e < nitems && (kmax = kdim(e)) <= 0) ++e;
    HChunkRegs<kRowsKC> R;
    if (e < nitems) load(e, 0, R);
    while (e < nitems) {
        const bool dense = e < nd;
        __syncthreads();
        if (dense) store_rowmajor(S, tid, R.a);
        else store_colmajor(S, tid, R.a);
        store_vec(S, tid, R.b);
        __syncthreads();
        const int kn = chunk_width<kRowsKC>(k0, kmax);
        // next chunk: the rest of this item, else the next item with work
        k0 += kRowsKC;
        if (k0 >= kmax) {
            k0 = 0;
            ++e;
            while (e < nitems && (kmax = kdim(e)) <= 0) ++e;
        }
        if (e < nitems) load(e, k0, R);
        chunk_products<NT, kRowsKC>(S, kn, w, li, lk, cr, ci);
    }
    store_tiles<NT>(cr, ci, Y + (int64_t)lo * nr, nr, m, nr, w, li, lk);
}

// Four waves per SIMD (four workgroups per CU): at a 16-wide K chunk the
// staging registers fit 128 VGPRs, and the extra resident workgroups hide
// the chunk loads better than a 32-wide chunk at three per CU (7.3 against
// 8.3 ms per 64-RHS apply on the 200k-panel bent cylinder).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_hrows(const HLeafs L, const HBlk* __restrict__ blk,
                                               const z32* __restrict__ dval, const z32* __restrict__ uv,
                                               const int64_t* __restrict__ toff, const z64* __restrict__ T,
                                               const z64* __restrict__ X, int nr, z64* __restrict__ Y) {
    const int leaf = L.order[blockIdx.x];
    __shared__ HStage<kRowsKC> S;
    switch ((L.m[leaf] + 15) >> 4) {
    case 1: hrows_leaf<1>(L, leaf, blk, dval, uv, toff, T, X, nr, Y, S); break;
    case 2: hrows_leaf<2>(L, leaf, blk, dval, uv, toff, T, X, nr, Y, S); break;
    case 3: hrows_leaf<3>(L, leaf, blk, dval, uv, toff, T, X, nr, Y, S); break;
    default: hrows_leaf<4>(L, leaf, blk, dval, uv, toff, T, X, nr, Y, S); break;
    }
}

void launch_hproj(hipStream_t s, int ntiles, const HTile* tiles, const int64_t* vrow, const z32* uv, const z64* X,
                  int nr, z64* T, int nsplit, const HSplit* splits, z64* P) {
    if (ntiles > 0) hipLaunchKernelGGL(k_hproj, dim3(ntiles), dim3(256), 0, s, tiles, vrow, uv, X, nr, T, P);
    if (nsplit > 0) hipLaunchKernelGGL(k_hproj_sum, dim3(nsplit), dim3(256), 0, s, splits, P, nr, T);
}
void launch_hrows(hipStream_t s, const HLeafs& L, const HBlk* blk, const z32* dval, const z32* uv,
                  const int64_t* toff, const z64* T, const z64* X, int nr, z64* Y) {
    hipLaunchKernelGGL(k_hrows, dim3(L.count), dim3(256), 0, s, L, blk, dval, uv, toff, T, X, nr, Y);
}

// ---------------------------------------------------------------------------
// Block-Jacobi preconditioner: inverses of the diagonal leaf blocks (FP64,
// Gauss-Jordan with row pivoting; a singular block keeps the identity and
// raises status[0]).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_hjacobi_inv(const HLeafs L, const HBlk* __restrict__ blk,
                                                     const z32* __restrict__ dval, z64* __restrict__ minv,
                                                     int* __restrict__ status) {
    const int leaf = blockIdx.x;
    const int m = L.m[leaf];
    const HBlk b = blk[L.diag[leaf]];
    const int tid = threadIdx.x;
    __shared__ z64 A[64][65];
    __shared__ int piv[64];
    __shared__ double sv[4];
    __shared__ int s_bad;
    for (int q = tid; q < 64 * 64; q += 256) {
        const int i = q >> 6, j = q & 63;
        z64 v{i == j ? 1.0 : 0.0, 0.0};
        if (i < m && j < m) {
            const z32 d = dval[b.off + i * m + j];
            v = z64{d.x, d.y};
        }
        A[i][j] = v;
    }
    if (tid == 0) s_bad = 0;
    __syncthreads();
    for (int k = 0; k < m; ++k) {
        // pivot row: largest |A[i][k]|, i >= k (wave 0)
        if (tid < 64) {
            double v = -1.0;
            int idx = k;
            if (tid >= k && tid < m) {
                v = cabs1(A[tid][k]);
                idx = tid;
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                const double ov = __shfl_xor(v, off, 64);
                const int oi = __shfl_xor(idx, off, 64);
                if (ov > v || (ov == v && oi < idx)) {
                    v = ov;
                    idx = oi;
                }
            }
            if (tid == 0) {
                piv[k] = idx;
                sv[0] = v;
            }
        }
        __syncthreads();
        const int p = piv[k];
        if (!(sv[0] > 0.0)) {
            if (tid == 0) s_bad = 1;
            break;
        }
        if (p != k)
            for (int j = tid; j < 64; j += 256) {
                const z64 t = A[k][j];
                A[k][j] = A[p][j];
                A[p][j] = t;
            }
        __syncthreads();
        const z64 ip = cinv(A[k][k]);
        __syncthreads();
        if (tid == 0) A[k][k] = z64{1.0, 0.0};
        __syncthreads();
        for (int j = tid; j < 64; j += 256) A[k][j] = cmul(A[k][j], ip);
        __syncthreads();
        // eliminate column k from every other row
        for (int q = tid; q < 64 * 64; q += 256) {
            const int i = q >> 6, j = q & 63;
            if (i != k && j != k) A[i][j] = csub(A[i][j], cmul(A[i][k], A[k][j]));
        }
        __syncthreads();
        for (int i = tid; i < 64; i += 256)
            if (i != k) A[i][k] = cmul(z64{-A[i][k].x, -A[i][k].y}, A[k][k]);
        __syncthreads();
    }
    if (s_bad) {
        if (tid == 0) atomicAdd(status, 1);
        for (int q = tid; q < 64 * 64; q += 256) {
            const int i = q >> 6, j = q & 63;
            minv[(int64_t)leaf * 4096 + q] = z64{i == j ? 1.0 : 0.0, 0.0};
        }
        return;
    }
    // undo the row interchanges on the columns, last first
    for (int k = m - 1; k >= 0; --k) {
        const int p = piv[k];
        if (p != k)
            for (int i = tid; i < 64; i += 256) {
                const z64 t = A[i][k];
                A[i][k] = A[i][p];
                A[i][p] = t;
            }
        __syncthreads();
    }
    for (int q = tid; q < 64 * 64; q += 256) minv[(int64_t)leaf * 4096 + q] = A[q >> 6][q & 63];
}

// Y[leaf rows] = Minv_leaf X[leaf rows]
__global__ __launch_bounds__(256) void k_hjacobi_apply(const HLeafs L, const z64* __restrict__ minv,
                                                       const z64* __restrict__ X, int nr, z64* __restrict__ Y) {
    const int leaf = blockIdx.x;
    const int lo = L.lo[leaf], m = L.m[leaf];
    const int tid = threadIdx.x, c = tid & 63, rg = tid >> 6;
    __shared__ z64 M[64][64];
    for (int q = tid; q < 64 * 64; q += 256) M[q >> 6][q & 63] = minv[(int64_t)leaf * 4096 + q];
    __syncthreads();
    if (c >= nr) return;
    z64 acc[16] = {};
    for (int j = 0; j < m; ++j) {
        const z64 x = X[(int64_t)(lo + j) * nr + c];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const z64 a = M[rg * 16 + q][j];
            acc[q].x += a.x * x.x - a.y * x.y;
            acc[q].y += a.x * x.y + a.y * x.x;
        }
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int i = rg * 16 + q;
        if (i < m) Y[(int64_t)(lo + i) * nr + c] = acc[q];
    }
}

void launch_hjacobi_inv(hipStream_t s, const HLeafs& L, const HBlk* blk, const z32* dval, z64* minv, int* status) {
    hipLaunchKernelGGL(k_hjacobi_inv, dim3(L.count), dim3(256), 0, s, L, blk, dval, minv, status);
}
void launch_hjacobi_apply(hipStream_t s, const HLeafs& L, const z64* minv, const z64* X, int nr, z64* Y) {
    hipLaunchKernelGGL(k_hjacobi_apply, dim3(L.count), dim3(256), 0, s, L, minv, X, nr, Y);
}

// ---------------------------------------------------------------------------
// Batched GMRES vector kernels.  Basis vector i at V + i * vstride.
// ---------------------------------------------------------------------------
// partial[blk][i][c] = sum over this workgroup's positions of conj(V_i) W
__global__ __launch_bounds__(256) void k_gm_dots(int64_t n, int nr, int nb, const z64* __restrict__ V, int64_t vstride,
                                                 const z64* __restrict__ W, z64* __restrict__ partial) {
    const int tid = threadIdx.x, c = tid & 63, g = tid >> 6;
    __shared__ z64 red[4][8][64];
    for (int i0 = 0; i0 < nb; i0 += 8) {
        z64 acc[8] = {};
        if (c < nr)
            for (int64_t p = (int64_t)blockIdx.x * 4 + g; p < n; p += (int64_t)gridDim.x * 4) {
                const z64 w = W[p * nr + c];
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if (i0 + q < nb) {
                        const z64 v = V[(int64_t)(i0 + q) * vstride + p * nr + c];
                        acc[q].x += v.x * w.x + v.y * w.y;
                        acc[q].y += v.x * w.y - v.y * w.x;
                    }
            }
#pragma unroll
        for (int q = 0; q < 8; ++q) red[g][q][c] = acc[q];
        __syncthreads();
        for (int e = tid; e < 8 * 64; e += 256) {
            const int q = e >> 6, cc = e & 63;
            if (i0 + q < nb && cc < nr) {
                z64 sum = red[0][q][cc];
                for (int gg = 1; gg < 4; ++gg) {
                    sum.x += red[gg][q][cc].x;
                    sum.y += red[gg][q][cc].y;
                }
                partial[((int64_t)blockIdx.x * nb + i0 + q) * nr + cc] = sum;
            }
        }
        __syncthreads();
    }
}

// out[i][c] = sum over workgroups of partial (fixed order)
__global__ void k_gm_reduce(int nblk, int nb, int nr, const z64* __restrict__ partial, z64* __restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nb * nr) return;
    double sx = 0, sy = 0;
    for (int b = 0; b < nblk; ++b) {
        sx += partial[(int64_t)b * nb * nr + e].x;
        sy += partial[(int64_t)b * nb * nr + e].y;
    }
    out[e] = z64{sx, sy};
}

constexpr int kGmBlocks = 512;

void launch_gm_dots(hipStream_t s, int64_t n, int nr, int nb, const z64* V, int64_t vstride, const z64* W,
                    z64* partial, z64* out) {
    hipLaunchKernelGGL(k_gm_dots, dim3(kGmBlocks), dim3(256), 0, s, n, nr, nb, V, vstride, W, partial);
    hipLaunchKernelGGL(k_gm_reduce, dim3((nb * nr + 255) / 256), dim3(256), 0, s, kGmBlocks, nb, nr, partial, out);
}

// Shared-basis products for the coarse space and the recycled solutions: Q
// holds nb columns of n entries (column-major), one set for every
// right-hand side.  out[i][c] = sum_p conj(Q[i][p]) X[p][c] (fixed-order
// two-pass reduction), and Y[p][c] += sum_i Q[i][p] coef[i][c].
__global__ __launch_bounds__(256) void k_basis_dots(int64_t n, int nr, int nb, const z64* __restrict__ Q,
                                                    const z64* __restrict__ X, z64* __restrict__ partial) {
    const int tid = threadIdx.x, c = tid & 63, g = tid >> 6;
    __shared__ z64 red[4][8][64];
    for (int i0 = 0; i0 < nb; i0 += 8) {
        z64 acc[8] = {};
        if (c < nr)
            for (int64_t p = (int64_t)blockIdx.x * 4 + g; p < n; p += (int64_t)gridDim.x * 4) {
                const z64 x = X[p * nr + c];
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if (i0 + q < nb) {
                        const z64 v = Q[(int64_t)(i0 + q) * n + p];
                        acc[q].x += v.x * x.x + v.y * x.y;
                        acc[q].y += v.x * x.y - v.y * x.x;
                    }
            }
#pragma unroll
        for (int q = 0; q < 8; ++q) red[g][q][c] = acc[q];
        __syncthreads();
        for (int e = tid; e < 8 * 64; e += 256) {
            const int q = e >> 6, cc = e & 63;
            if (i0 + q < nb && cc < nr) {
                z64 sum = red[0][q][cc];
                for (int gg = 1; gg < 4; ++gg) {
                    sum.x += red[gg][q][cc].x;
                    sum.y += red[gg][q][cc].y;
                }
                partial[((int64_t)blockIdx.x * nb + i0 + q) * nr + cc] = sum;
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_basis_combine(int64_t n, int nr, int nb, const z64* __restrict__ Q,
                                                       const z64* __restrict__ coef, z64* __restrict__ Y) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * nr) return;
    const int64_t p = e / nr;
    const int c = (int)(e % nr);
    double sx = 0, sy = 0;
    for (int i = 0; i < nb; ++i) {
        const z64 v = Q[(int64_t)i * n + p], h = coef[i * nr + c];
        sx += v.x * h.x - v.y * h.y;
        sy += v.x * h.y + v.y * h.x;
    }
    Y[e] = z64{Y[e].x + sx, Y[e].y + sy};
}

// out[i][c] = sum_j M[i][j] w[j][c] for a small nb x nb matrix M (row-major)
__global__ void k_small_apply(int nb, int nr, const z64* __restrict__ M, const z64* __restrict__ w,
                              z64* __restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nb * nr) return;
    const int i = e / nr, c = e % nr;
    double sx = 0, sy = 0;
    for (int j = 0; j < nb; ++j) {
        const z64 a = M[i * nb + j], b = w[j * nr + c];
        sx += a.x * b.x - a.y * b.y;
        sy += a.x * b.y + a.y * b.x;
    }
    out[e] = z64{sx, sy};
}

// out[p] = X[p * nr + c]: one right-hand side's column, contiguous
__global__ void k_gather_col(int64_t n, int nr, int c, const z64* __restrict__ X, z64* __restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p < n) out[p] = X[p * nr + c];
}
void launch_gather_col(hipStream_t s, int64_t n, int nr, int c, const z64* X, z64* out) {
    hipLaunchKernelGGL(k_gather_col, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, nr, c, X, out);
}

// out = a X + b Y + c Z (complex vectors, real scalars; Z may be null)
__global__ void k_lincomb3(int64_t tot, double a, const z64* __restrict__ X, double b, const z64* __restrict__ Y,
                           double c, const z64* __restrict__ Z, z64* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= tot) return;
    z64 v{a * X[e].x + b * Y[e].x, a * X[e].y + b * Y[e].y};
    if (Z) {
        v.x += c * Z[e].x;
        v.y += c * Z[e].y;
    }
    out[e] = v;
}
void launch_lincomb3(hipStream_t s, int64_t tot, double a, const z64* X, double b, const z64* Y, double c,
                     const z64* Z, z64* out) {
    hipLaunchKernelGGL(k_lincomb3, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, tot, a, X, b, Y, c, Z, out);
}

// Local 2m x 2m block inverses on [p; q] (rows 0..N-1 and N..2N-1 of X):
// [y_p; y_q] of leaf l = Binv_l [x_p; x_q] with Binv_l at lblk + l * 128 * 128
// (row-major, leading dimension 128).  One workgroup per leaf, a thread per
// (output row, right-hand side).
__global__ __launch_bounds__(256) void k_hblock2_apply(const HLeafs L, int64_t nt, const z64* __restrict__ lblk,
                                                       const z64* __restrict__ X, int nr, z64* __restrict__ Y) {
    const int leaf = blockIdx.x;
    const int lo = L.lo[leaf], m = L.m[leaf];
    const z64* Bl = lblk + (int64_t)leaf * 128 * 128;
    for (int e = threadIdx.x; e < 2 * m * nr; e += 256) {
        const int i = e / nr, c = e % nr;
        double sx = 0, sy = 0;
        for (int j = 0; j < 2 * m; ++j) {
            const int64_t row = j < m ? lo + j : nt + lo + (j - m);
            const z64 a = Bl[i * 128 + j], x = X[row * nr + c];
            sx += a.x * x.x - a.y * x.y;
            sy += a.x * x.y + a.y * x.x;
        }
        const int64_t orow = i < m ? lo + i : nt + lo + (i - m);
        Y[orow * nr + c] = z64{sx, sy};
    }
}
void launch_hblock2_apply(hipStream_t s, const HLeafs& L, int64_t nt, const z64* lblk, const z64* X, int nr,
                          z64* Y) {
    hipLaunchKernelGGL(k_hblock2_apply, dim3(L.count), dim3(256), 0, s, L, nt, lblk, X, nr, Y);
}

void launch_basis_dots(hipStream_t s, int64_t n, int nr, int nb, const z64* Q, const z64* X, z64* partial,
                       z64* out) {
    hipLaunchKernelGGL(k_basis_dots, dim3(kGmBlocks), dim3(256), 0, s, n, nr, nb, Q, X, partial);
    hipLaunchKernelGGL(k_gm_reduce, dim3((nb * nr + 255) / 256), dim3(256), 0, s, kGmBlocks, nb, nr, partial, out);
}
void launch_basis_combine(hipStream_t s, int64_t n, int nr, int nb, const z64* Q, const z64* coef, z64* Y) {
    const int64_t tot = n * nr;
    hipLaunchKernelGGL(k_basis_combine, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, nr, nb, Q, coef,
                       Y);
}
void launch_small_apply(hipStream_t s, int nb, int nr, const z64* M, const z64* w, z64* out) {
    hipLaunchKernelGGL(k_small_apply, dim3((nb * nr + 63) / 64), dim3(64), 0, s, nb, nr, M, w, out);
}

// W[p][c] += sign * sum_i V_i[p][c] coef[i][c]
__global__ __launch_bounds__(256) void k_gm_combine(int64_t n, int nr, int nb, const z64* __restrict__ V,
                                                    int64_t vstride, const z64* __restrict__ coef, double sign,
                                                    z64* __restrict__ W) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * nr) return;
    const int c = (int)(e % nr);
    double sx = 0, sy = 0;
    for (int i = 0; i < nb; ++i) {
        const z64 v = V[(int64_t)i * vstride + e], h = coef[i * nr + c];
        sx += v.x * h.x - v.y * h.y;
        sy += v.x * h.y + v.y * h.x;
    }
    W[e] = z64{W[e].x + sign * sx, W[e].y + sign * sy};
}

void launch_gm_combine(hipStream_t s, int64_t n, int nr, int nb, const z64* V, int64_t vstride, const z64* coef,
                       double sign, z64* W) {
    const int64_t tot = n * nr;
    hipLaunchKernelGGL(k_gm_combine, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, nr, nb, V, vstride, coef,
                       sign, W);
}

// out[p][c] = W[p][c] * scale[c]
__global__ void k_gm_scale(int64_t n, int nr, const z64* __restrict__ W, const double* __restrict__ scale,
                           z64* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * nr) return;
    const double f = scale[e % nr];
    out[e] = z64{W[e].x * f, W[e].y * f};
}

void launch_gm_scale(hipStream_t s, int64_t n, int nr, const z64* W, const double* scale, z64* out) {
    const int64_t tot = n * nr;
    hipLaunchKernelGGL(k_gm_scale, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, nr, W, scale, out);
}

// out[p][c] = a[p][c] - b[p][c]
__global__ void k_gm_sub(int64_t tot, const z64* __restrict__ a, const z64* __restrict__ b, z64* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= tot) return;
    out[e] = z64{a[e].x - b[e].x, a[e].y - b[e].y};
}

void launch_gm_sub(hipStream_t s, int64_t tot, const z64* a, const z64* b, z64* out) {
    hipLaunchKernelGGL(k_gm_sub, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, tot, a, b, out);
}

// ---------------------------------------------------------------------------
// Incident traces (reference src/core/incident.jl:64-76): q_inc in tree
// order for the right-hand side b = B q_inc of the scattered-field equation
// (B = S + (i/k)(1/2 I + K') or S, full.jl:144-148), and the scattered traces
// p_scat = x, q_scat = -q_inc of a rigid surface (full.jl:160-161).
// ---------------------------------------------------------------------------
__device__ inline void inc_at(const DevMesh& m, double k, const double* d, int i, z64& p, z64& q) {
    const int nt = m.nt;
    const double ph = k * (d[0] * m.cen[i] + d[1] * m.cen[nt + i] + d[2] * m.cen[2 * nt + i]);
    double sn, cs;
    sincos(ph, &sn, &cs);
    const double dn = k * (d[0] * m.nrm[i] + d[1] * m.nrm[nt + i] + d[2] * m.nrm[2 * nt + i]);
    p = z64{cs, sn};
    q = z64{-dn * sn, dn * cs};
}

__global__ void k_hqinc(DevMesh m, double k, int nr, const double* __restrict__ dirs, const int* __restrict__ perm,
                        z64* __restrict__ Q) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)m.nt * nr) return;
    const int pos = (int)(e / nr), c = (int)(e % nr);
    z64 p, q;
    inc_at(m, k, dirs + 3 * c, perm[pos], p, q);
    Q[e] = q;
}

__global__ void k_htraces(DevMesh m, double k, int nr, const double* __restrict__ dirs, const int* __restrict__ perm,
                          const z64* __restrict__ X, z64* __restrict__ P, z64* __restrict__ Q) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)m.nt * nr) return;
    const int pos = (int)(e / nr), c = (int)(e % nr);
    const int i = perm[pos];
    z64 p, q;
    inc_at(m, k, dirs + 3 * c, i, p, q);
    P[(int64_t)c * m.nt + i] = X[e];
    Q[(int64_t)c * m.nt + i] = z64{-q.x, -q.y};
}

// Mueller right-hand side [p_inc; q_inc] in tree order (unknown I < nt: the
// pressure of panel I; I >= nt: the normal derivative of panel I - nt).
__global__ void k_hfluid_rhs(DevMesh m, double k, int nr, const double* __restrict__ dirs, const int* __restrict__ perm,
                             z64* __restrict__ B) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= 2 * (int64_t)m.nt * nr) return;
    const int pos = (int)(e / nr), c = (int)(e % nr);
    const int I = perm[pos];
    z64 p, q;
    inc_at(m, k, dirs + 3 * c, I < m.nt ? I : I - m.nt, p, q);
    B[e] = I < m.nt ? p : q;
}
// scattered traces from the Mueller solution (total fields): x - incident
__global__ void k_hfluid_traces(DevMesh m, double k, int nr, const double* __restrict__ dirs,
                                const int* __restrict__ perm, const z64* __restrict__ X, z64* __restrict__ P,
                                z64* __restrict__ Q) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= 2 * (int64_t)m.nt * nr) return;
    const int pos = (int)(e / nr), c = (int)(e % nr);
    const int I = perm[pos];
    const int i = I < m.nt ? I : I - m.nt;
    z64 p, q;
    inc_at(m, k, dirs + 3 * c, i, p, q);
    const z64 x = X[e];
    if (I < m.nt)
        P[(int64_t)c * m.nt + i] = z64{x.x - p.x, x.y - p.y};
    else
        Q[(int64_t)c * m.nt + i] = z64{x.x - q.x, x.y - q.y};
}
void launch_hfluid_rhs(hipStream_t s, const DevMesh& m, double k, int nr, const double* dirs, const int* perm,
                       z64* B) {
    const int64_t tot = 2 * (int64_t)m.nt * nr;
    hipLaunchKernelGGL(k_hfluid_rhs, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, m, k, nr, dirs, perm, B);
}
void launch_hfluid_traces(hipStream_t s, const DevMesh& m, double k, int nr, const double* dirs, const int* perm,
                          const z64* X, z64* P, z64* Q) {
    const int64_t tot = 2 * (int64_t)m.nt * nr;
    hipLaunchKernelGGL(k_hfluid_traces, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, m, k, nr, dirs, perm,
                       X, P, Q);
}

void launch_hqinc(hipStream_t s, const DevMesh& m, double k, int nr, const double* dirs, const int* perm, z64* Q) {
    const int64_t tot = (int64_t)m.nt * nr;
    hipLaunchKernelGGL(k_hqinc, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, m, k, nr, dirs, perm, Q);
}
void launch_htraces(hipStream_t s, const DevMesh& m, double k, int nr, const double* dirs, const int* perm,
                    const z64* X, z64* P, z64* Q) {
    const int64_t tot = (int64_t)m.nt * nr;
    hipLaunchKernelGGL(k_htraces, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, m, k, nr, dirs, perm, X, P,
                       Q);
}

// Undo the ACA scaling of Mueller blocks in the compacted factors: rows of
// normal-derivative unknowns of U times sig, of V divided by sig.
__global__ void k_hunscale(int nblk, const HBlk* __restrict__ blk, const int* __restrict__ list,
                           const int* __restrict__ perm, int nt, z32* __restrict__ uv) {
    const HBlk b = blk[list[blockIdx.x]];
    if (b.rank == 0 || b.sig == 1.f) return;
    z32* U = uv + b.off;
    z32* V = uv + b.off + (int64_t)b.rank * b.m;
    const float is = 1.f / b.sig;
    for (int64_t e = threadIdx.x; e < (int64_t)b.rank * b.m; e += 256)
        if (perm[b.row_lo + (int)(e % b.m)] >= nt) {
            U[e].x *= b.sig;
            U[e].y *= b.sig;
        }
    for (int64_t e = threadIdx.x; e < (int64_t)b.rank * b.n; e += 256)
        if (perm[b.col_lo + (int)(e % b.n)] >= nt) {
            V[e].x *= is;
            V[e].y *= is;
        }
}
void launch_hunscale(hipStream_t s, int nblk, const HBlk* blk, const int* list, const int* perm, int nt, z32* uv) {
    if (nblk > 0) hipLaunchKernelGGL(k_hunscale, dim3(nblk), dim3(256), 0, s, nblk, blk, list, perm, nt, uv);
}

// compact copy of low-rank factors: dst[i] for i < len from src
__global__ void k_hcompact(int nblk, const HBlk* __restrict__ oldb, const HBlk* __restrict__ newb,
                           const z32* __restrict__ src, z32* __restrict__ dst) {
    const HBlk a = oldb[blockIdx.x], b = newb[blockIdx.x];
    const int64_t nu = (int64_t)b.rank * b.m, nv = (int64_t)b.rank * b.n;
    for (int64_t e = threadIdx.x; e < nu; e += 256) dst[b.off + e] = src[a.off + e];
    for (int64_t e = threadIdx.x; e < nv; e += 256) dst[b.off + nu + e] = src[a.off + (int64_t)a.rmax * a.m + e];
}

void launch_hcompact(hipStream_t s, int nblk, const HBlk* oldb, const HBlk* newb, const z32* src, z32* dst) {
    if (nblk > 0) hipLaunchKernelGGL(k_hcompact, dim3(nblk), dim3(256), 0, s, nblk, oldb, newb, src, dst);
}

}  // namespace bem
