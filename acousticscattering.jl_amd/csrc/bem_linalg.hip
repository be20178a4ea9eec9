// Dense complex linear algebra for the BEM systems (CDNA4 / gfx950).
//
// Matrices are row-major complex (C2<R>), R = double (plain FP64 path) or
// float (mixed path).  Row-major keeps the partial-pivoting row interchanges
// and the panel's per-row reads contiguous.
//
// LU: right-looking, block width 64, partial pivoting by max |re|+|im|
// (the LAPACK zgetrf/cgetrf rule used by the reference,
// src/solvers/bem/full.jl:404, src/solvers/bem/fluid_system.jl:217).
//   panel   one cooperative launch factors the n-k0 x 64 panel; each row
//           lives in the registers of a lane pair; one grid barrier per column
//   swap    row interchanges outside the panel (coalesced row segments)
//   trtri   inverses of the 64x64 unit-lower L11 and upper U11 blocks
//   trsm    U12 = inv(L11) A12
//   gemm    A22 -= L21 U12 on MFMA (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32),
//           complex as four real products, 128x128 tiles staged in LDS
// Solves use the stored block inverses: one launch per 64-row block.
#include <climits>
#include <cmath>
#include <vector>

#include "bem_device.h"

namespace bem {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ inline double wave_add(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
template <typename R>
__device__ inline double cabs_d(C2<R> z) { return hypot((double)z.x, (double)z.y); }

__global__ void k_fill_int(int* p, int n, int v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
void launch_fill_int(hipStream_t s, int* p, int n, int v) {
    hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
}
__global__ void k_fill_double(double* p, int64_t n, double v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
void launch_fill_double(hipStream_t s, double* p, int64_t n, double v) {
    hipLaunchKernelGGL(k_fill_double, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

// dst = scale * src - sub   (sub may be null)
__global__ void k_copy_z(int n, const z64* src, z64* dst, double scale, const z64* sub) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    z64 v = src[i];
    v.x *= scale; v.y *= scale;
    if (sub) { v.x -= sub[i].x; v.y -= sub[i].y; }
    dst[i] = v;
}
void launch_copy_z(hipStream_t s, int n, const z64* src, z64* dst, double scale, const z64* sub) {
    hipLaunchKernelGGL(k_copy_z, dim3((n + 255) / 256), dim3(256), 0, s, n, src, dst, scale, sub);
}

// ---------------------------------------------------------------------------
// equilibration (reference src/solvers/bem/full.jl:388-394,
// fluid_system.jl:196-216): row max, then column max of the row-scaled matrix
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_row_stats(const C2<R>* __restrict__ A, int64_t lda, int n,
                                                   double* rows, double* rowsum) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    double mx = 0, sm = 0;
    const C2<R>* row = A + (int64_t)i * lda;
    for (int j = lane; j < n; j += 64) {
        double a = cabs_d(row[j]);
        mx = fmax(mx, a);
        sm += a;
    }
    mx = wave_max(mx);
    sm = wave_add(sm);
    if (lane == 0) {
        rows[i] = mx;
        if (rowsum) rowsum[i] = sm;
    }
}
template <typename R>
void launch_row_stats(hipStream_t s, const C2<R>* A, int64_t lda, int n, double* rows, double* rowsum) {
    hipLaunchKernelGGL((k_row_stats<R>), dim3((n + 3) / 4), dim3(256), 0, s, A, lda, n, rows, rowsum);
}

constexpr int kColChunk = 512;
template <typename R>
__global__ __launch_bounds__(256) void k_col_max(const C2<R>* __restrict__ A, int64_t lda, int n,
                                                 const double* __restrict__ rows, double* cols) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int i0 = blockIdx.y * kColChunk;
    if (j >= n) return;
    const int i1 = min(n, i0 + kColChunk);
    double mx = 0;
    for (int i = i0; i < i1; ++i) {
        double r = rows[i];
        if (r == 0) r = 1;
        mx = fmax(mx, cabs_d(A[(int64_t)i * lda + j]) / r);
    }
    // non-negative doubles order like their bit patterns
    atomicMax((unsigned long long*)&cols[j], (unsigned long long)__double_as_longlong(mx));
}
template <typename R>
void launch_col_max(hipStream_t s, const C2<R>* A, int64_t lda, int n, const double* rows, double* cols) {
    launch_fill_double(s, cols, n, 0.0);
    hipLaunchKernelGGL((k_col_max<R>), dim3((n + 255) / 256, (n + kColChunk - 1) / kColChunk), dim3(256), 0, s,
                       A, lda, n, rows, cols);
}

__global__ void k_fix_scale(int n, double* v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && v[i] == 0) v[i] = 1.0;
}
void launch_fix_scale(hipStream_t s, int n, double* v) {
    hipLaunchKernelGGL(k_fix_scale, dim3((n + 255) / 256), dim3(256), 0, s, n, v);
}

template <typename Ri, typename Ro>
__global__ __launch_bounds__(256) void k_scale_copy(const C2<Ri>* __restrict__ A, int64_t lda, int n,
                                                    const double* __restrict__ rows,
                                                    const double* __restrict__ cols, C2<Ro>* __restrict__ out,
                                                    int64_t ldo) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    for (int i = blockIdx.y; i < n; i += gridDim.y) {
        const C2<Ri> a = A[(int64_t)i * lda + j];
        double sx = a.x, sy = a.y;
        if (rows) {
            const double r = rows[i], c = cols[j];
            sx = sx / r / c;
            sy = sy / r / c;
        }
        out[(int64_t)i * ldo + j] = C2<Ro>{(Ro)sx, (Ro)sy};
    }
}
template <typename Ri, typename Ro>
void launch_scale_copy(hipStream_t s, const C2<Ri>* A, int64_t lda, int n, const double* rows,
                       const double* cols, C2<Ro>* out, int64_t ldo) {
    const int gy = n < 8192 ? n : 8192;
    hipLaunchKernelGGL((k_scale_copy<Ri, Ro>), dim3((n + 255) / 256, gy), dim3(256), 0, s, A, lda, n, rows, cols,
                       out, ldo);
}

// ---------------------------------------------------------------------------
// y = A x in FP64 arithmetic (A in R storage).  With b: r = b - A x and
// den = |b| + |A||x| (componentwise backward error, fluid_system.jl:241-262).
// One wave per row, fixed-order reduction.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_matvec(const C2<R>* __restrict__ A, int64_t lda, int rows, int n,
                                                const z64* __restrict__ x, z64* __restrict__ y,
                                                const z64* __restrict__ b, double* __restrict__ den) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= rows) return;
    const C2<R>* row = A + (int64_t)i * lda;
    double sx = 0, sy = 0, sa = 0;
    for (int j = lane; j < n; j += 64) {
        const C2<R> a = row[j];
        const z64 v = x[j];
        const double ax = a.x, ay = a.y;
        sx += ax * v.x - ay * v.y;
        sy += ax * v.y + ay * v.x;
        if (den) sa += hypot(ax, ay) * hypot(v.x, v.y);
    }
    sx = wave_add(sx);
    sy = wave_add(sy);
    if (den) sa = wave_add(sa);
    if (lane == 0) {
        if (b) {
            const z64 bi = b[i];
            y[i] = z64{bi.x - sx, bi.y - sy};
            if (den) den[i] = hypot(bi.x, bi.y) + sa;
        } else {
            y[i] = z64{sx, sy};
        }
    }
}
template <typename R>
void launch_matvec(hipStream_t s, const C2<R>* A, int64_t lda, int rows, int cols, const z64* x, z64* y,
                   const z64* b, double* den) {
    hipLaunchKernelGGL((k_matvec<R>), dim3((rows + 3) / 4), dim3(256), 0, s, A, lda, rows, cols, x, y, b, den);
}

// ---------------------------------------------------------------------------
// refinement helpers
// ---------------------------------------------------------------------------
template <typename R>
__global__ void k_prep_rhs(int n, const z64* b, const double* rows, double inv_mag, C2<R>* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double r = rows ? rows[i] : 1.0;
    out[i] = C2<R>{(R)(b[i].x / r * inv_mag), (R)(b[i].y / r * inv_mag)};
}
template <typename R>
void launch_prep_rhs(hipStream_t s, int n, const z64* b, const double* rows, double inv_mag, C2<R>* out) {
    hipLaunchKernelGGL((k_prep_rhs<R>), dim3((n + 255) / 256), dim3(256), 0, s, n, b, rows, inv_mag, out);
}

template <typename R>
__global__ void k_finish(int n, const C2<R>* v, double mag, const double* cols, z64* x, int acc, int* changed) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double c = cols ? cols[i] : 1.0;
    double dx = (double)v[i].x * mag / c, dy = (double)v[i].y * mag / c;
    if (acc) {
        z64 o = x[i];
        z64 nw{o.x + dx, o.y + dy};
        if ((nw.x != o.x || nw.y != o.y) && changed) *changed = 1;
        x[i] = nw;
    } else {
        x[i] = z64{dx, dy};
    }
}
template <typename R>
void launch_finish(hipStream_t s, int n, const C2<R>* v, double mag, const double* cols, z64* x, int acc,
                   int* changed) {
    hipLaunchKernelGGL((k_finish<R>), dim3((n + 255) / 256), dim3(256), 0, s, n, v, mag, cols, x, acc, changed);
}

__global__ __launch_bounds__(256) void k_bwd_stats(int n, const z64* r, const double* den, const z64* x,
                                                   const z64* b, double* out) {
    double mr = 0, mq = 0, mx = 0, mb = 0, bad = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double ar = hypot(r[i].x, r[i].y), d = den[i];
        double ax = hypot(x[i].x, x[i].y), ab = hypot(b[i].x, b[i].y);
        if (!isfinite(ar) || !isfinite(d)) bad = 1;
        mr = fmax(mr, ar);
        double q = (d == 0) ? (ar == 0 ? 0.0 : INFINITY) : ar / d;
        mq = fmax(mq, q);
        mx = fmax(mx, ax);
        mb = fmax(mb, ab);
    }
    __shared__ double sh[5][4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    mr = wave_max(mr); mq = wave_max(mq); mx = wave_max(mx); mb = wave_max(mb); bad = wave_max(bad);
    if (lane == 0) { sh[0][wv] = mr; sh[1][wv] = mq; sh[2][wv] = mx; sh[3][wv] = mb; sh[4][wv] = bad; }
    __syncthreads();
    if (threadIdx.x < 5) {
        double v = sh[threadIdx.x][0];
        for (int q = 1; q < 4; ++q) v = fmax(v, sh[threadIdx.x][q]);
        out[blockIdx.x * 5 + threadIdx.x] = v;
    }
}
void launch_bwd_stats(hipStream_t s, int n, const z64* r, const double* den, const z64* x, const z64* b,
                      double* out) {
    hipLaunchKernelGGL(k_bwd_stats, dim3(kRedBlocks), dim3(256), 0, s, n, r, den, x, b, out);
}

__global__ __launch_bounds__(256) void k_max_abs_scaled(int n, const z64* b, const double* rows, double* out) {
    double m = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double r = rows ? rows[i] : 1.0;
        double v = hypot(b[i].x / r, b[i].y / r);
        if (!isfinite(v)) v = INFINITY;
        m = fmax(m, v);
    }
    __shared__ double sh[4];
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}
void launch_max_abs_scaled(hipStream_t s, int n, const z64* b, const double* rows, double* out) {
    hipLaunchKernelGGL(k_max_abs_scaled, dim3(kRedBlocks), dim3(256), 0, s, n, b, rows, out);
}

__global__ __launch_bounds__(256) void k_norm_stats(int n, const z64* r, const z64* b, const double* rows,
                                                    double* out) {
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double rr = r[i].x * r[i].x + r[i].y * r[i].y, bb = b[i].x * b[i].x + b[i].y * b[i].y;
        double w = rows ? rows[i] : 1.0;
        s0 += rr; s1 += bb; s2 += rr / (w * w); s3 += bb / (w * w);
    }
    __shared__ double sh[4][4];
    s0 = wave_add(s0); s1 = wave_add(s1); s2 = wave_add(s2); s3 = wave_add(s3);
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sh[0][wv] = s0; sh[1][wv] = s1; sh[2][wv] = s2; sh[3][wv] = s3; }
    __syncthreads();
    if (threadIdx.x < 4) {
        double v = 0;
        for (int q = 0; q < 4; ++q) v += sh[threadIdx.x][q];
        out[blockIdx.x * 4 + threadIdx.x] = v;
    }
}
void launch_norm_stats(hipStream_t s, int n, const z64* r, const z64* b, const double* rows, double* out) {
    hipLaunchKernelGGL(k_norm_stats, dim3(kRedBlocks), dim3(256), 0, s, n, r, b, rows, out);
}

// ---------------------------------------------------------------------------
// LU panel: cooperative kernel, 128 rows per workgroup, each row held by a
// lane pair (columns 0-31 and 32-63 of the panel) in registers.
// ---------------------------------------------------------------------------
constexpr int kNB = 64;

// Cross-workgroup traffic of the cooperative panel uses relaxed agent-scope
// atomic loads and stores only.  On gfx950 these are plain global loads and
// stores with the sc1 bit, coherent across XCDs without the L2-wide
// writeback/invalidate (buffer_wbl2 / buffer_inv) that agent-scope
// release/acquire fences emit.  Ordering comes from s_waitcnt: a wave's
// communication stores have completed before its workgroup arrives.
template <typename T>
__device__ __forceinline__ void st_ag(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T ld_ag(const T* p) {
    return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename R>
__device__ __forceinline__ void st_ag(C2<R>* p, C2<R> v) {
    st_ag(&p->x, v.x);
    st_ag(&p->y, v.y);
}
template <typename R>
__device__ __forceinline__ C2<R> ld_ag(const C2<R>* p) {
    return C2<R>{ld_ag(&p->x), ld_ag(&p->y)};
}
__device__ __forceinline__ void wait_mem() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Generation barrier over all workgroups of a cooperative launch.  The spin
// is bounded: a timeout sets the sticky flag status[1], after which every
// later barrier returns at once, so a failed launch drains quickly and the
// host reports the error.
__device__ inline void grid_barrier(unsigned* bar, unsigned G, int* status) {
    wait_mem();
    __syncthreads();
    if (threadIdx.x == 0 && ld_ag(&status[1]) == 0) {
        const unsigned gen = ld_ag(&bar[1]);
        wait_mem();
        const unsigned arrived = __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == G - 1) {
            st_ag(&bar[0], 0u);
            wait_mem();
            __hip_atomic_fetch_add(&bar[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            unsigned spins = 0;
            while (ld_ag(&bar[1]) == gen) {
                __builtin_amdgcn_s_sleep(1);
                if ((++spins & 1023u) == 0 && ld_ag(&status[1]) != 0) break;
                if (spins > (1u << 23)) {
                    st_ag(&status[1], 1);
                    break;
                }
            }
        }
        wait_mem();
    }
    __syncthreads();
}

// argmax of (value, index): larger value wins, ties go to the smaller index.
__device__ inline void wave_argmax(double& v, int& idx, int& blk) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        double ov = __shfl_xor(v, off, 64);
        int oi = __shfl_xor(idx, off, 64);
        int ob = __shfl_xor(blk, off, 64);
        if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; blk = ob; }
    }
}

// Panel layout: lane c of wave g owns column c of kPanelRPT consecutive rows,
// so a workgroup holds 4 * kPanelRPT rows x 64 columns in registers and the
// loads of a row are coalesced across the wave.
constexpr int kPanelRPT = 32;
constexpr int kPanelRows = 4 * kPanelRPT;

template <typename R>
__device__ __forceinline__ C2<R> pick(const C2<R> (&a)[kPanelRPT], int r) {
    C2<R> v{0, 0};
#pragma unroll
    for (int q = 0; q < kPanelRPT; ++q)
        if (q == r) v = a[q];
    return v;
}
template <typename R>
__device__ __forceinline__ void place(C2<R> (&a)[kPanelRPT], int r, C2<R> v) {
#pragma unroll
    for (int q = 0; q < kPanelRPT; ++q)
        if (q == r) a[q] = v;
}

// Per-column exchange: every workgroup publishes its best candidate (value,
// row index, the row's 64 values) in slot [j & 1][b] and then stamps it with
// seq = epoch + j + 1; wave 0 of every workgroup polls the stamps, reduces the
// candidates, and all lanes read the winning row and the diagonal row.  The
// stamps are unique per launch and column, so no counter needs resetting; the
// two slot parities keep a slow reader's column-j data intact while a fast
// workgroup publishes column j + 1.
template <typename R>
__global__ __launch_bounds__(256) void k_lu_panel(C2<R>* __restrict__ A, int64_t lda, int n, int k0, int w,
                                                  int* __restrict__ ipiv, PanelWS ws, unsigned epoch) {
    constexpr int RPT = kPanelRPT;
    const int G = gridDim.x, b = blockIdx.x;
    const int tid = threadIdx.x, c = tid & 63, g = tid >> 6;
    const int wg0 = k0 + b * kPanelRows;  // first row of this workgroup
    const int base = wg0 + g * RPT;       // first row of this lane
    C2<R>* cand_row = (C2<R>*)ws.cand_row;
    C2<R>* diag_row = (C2<R>*)ws.diag_row;
    unsigned* seqs = ws.bar;              // [2][G] stamps
    __shared__ double s_v[4];
    __shared__ int s_i[4];
    __shared__ int s_win, s_blk;
    __shared__ C2<R> s_prow[kNB], s_diag[kNB];
    __shared__ C2<R> s_l[kPanelRows];

    C2<R> a[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int row = base + r;
        a[r] = (row < n && c < w) ? A[(int64_t)row * lda + k0 + c] : C2<R>{0, 0};
    }
#pragma unroll 1
    for (int j = 0; j < w; ++j) {
        const int par = j & 1;
        const unsigned stamp = epoch + (unsigned)j + 1u;
        // -- candidates in column j among rows >= k0 + j (first maximum wins)
        if (c == j) {
            double bv = -1.0;
            int bi = INT_MAX;
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const int row = base + r;
                if (row < n && row >= k0 + j) {
                    const double v = (double)cabs1(a[r]);
                    if (v > bv) { bv = v; bi = row; }
                }
            }
            s_v[g] = bv;
            s_i[g] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            double bv = s_v[0];
            int bi = s_i[0];
            for (int q = 1; q < 4; ++q)
                if (s_v[q] > bv || (s_v[q] == bv && s_i[q] < bi)) { bv = s_v[q]; bi = s_i[q]; }
            st_ag(&ws.cand_val[par * G + b], bv);
            st_ag(&ws.cand_idx[par * G + b], bi);
            s_win = bi;
        }
        __syncthreads();
        {
            const int wrow = s_win;
            if (wrow != INT_MAX) {
                const int rel = wrow - wg0;
                if (g == rel / RPT) st_ag(&cand_row[(size_t)(par * G + b) * kNB + c], pick(a, rel % RPT));
            }
        }
        if (b == 0 && g == j / RPT) st_ag(&diag_row[par * kNB + c], pick(a, j % RPT));
        wait_mem();
        __syncthreads();
        if (tid == 0) st_ag(&seqs[par * G + b], stamp);
        // -- global pivot: poll every stamp, then reduce the candidates
        if (g == 0) {
            double bv = -2.0;
            int bi = INT_MAX, bb = 0;
            for (int q = c; q < G; q += 64) {
                unsigned spins = 0;
                while (ld_ag(&seqs[par * G + q]) != stamp) {
                    __builtin_amdgcn_s_sleep(1);
                    if (++spins > (1u << 23)) {
                        st_ag(&ws.status[1], 1);
                        break;
                    }
                }
                const double vv = ld_ag(&ws.cand_val[par * G + q]);
                const int ii = ld_ag(&ws.cand_idx[par * G + q]);
                if (vv > bv || (vv == bv && ii < bi)) { bv = vv; bi = ii; bb = q; }
            }
            wave_argmax(bv, bi, bb);
            if (c == 0) { s_win = bi; s_blk = bb; }
        }
        __syncthreads();
        const int p = s_win, pb = s_blk;
        if (tid < kNB) {
            s_prow[tid] = ld_ag(&cand_row[(size_t)(par * G + pb) * kNB + tid]);
            s_diag[tid] = ld_ag(&diag_row[par * kNB + tid]);
        }
        __syncthreads();
        const C2<R> pv = s_prow[j];
        const C2<R> prc = s_prow[c];
        // -- interchange rows k0 + j and p
        if (p != k0 + j) {
            if (b == 0 && g == j / RPT) place(a, j % RPT, prc);
            const int rel = p - wg0;
            if (rel >= 0 && rel < kPanelRows && g == rel / RPT) place(a, rel % RPT, s_diag[c]);
        }
        if (b == 0 && tid == 0) {
            ipiv[k0 + j] = p;
            if (pv.x == 0 && pv.y == 0) atomicMin(&ws.status[0], k0 + j);
        }
        // -- multipliers (lanes of column j), then the rank-1 update
        if (c == j) {
            const bool nz = pv.x != 0 || pv.y != 0;
            const C2<R> ip = nz ? cinv(pv) : C2<R>{0, 0};
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const int row = base + r;
                C2<R> l{0, 0};
                if (row < n && row > k0 + j) {
                    l = cmul(a[r], ip);
                    a[r] = l;
                }
                s_l[g * RPT + r] = l;
            }
        }
        __syncthreads();
        if (c > j && c < w) {
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const int row = base + r;
                if (row < n && row > k0 + j) a[r] = csub(a[r], cmul(s_l[g * RPT + r], prc));
            }
        }
    }
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int row = base + r;
        if (row < n && c < w) A[(int64_t)row * lda + k0 + c] = a[r];
    }
}

// Row interchanges of pivots ipiv[k0 .. k0+w) applied, in order, to the
// columns [cb0, ce0) and [cb1, ce1); each thread owns one column, so every
// interchange touches two contiguous row segments across the grid.
template <typename R>
__global__ __launch_bounds__(256) void k_laswp(C2<R>* __restrict__ A, int64_t lda, int k0, int w,
                                               const int* __restrict__ ipiv, int cb0, int ce0, int cb1, int ce1) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int n0 = ce0 - cb0;
    int col;
    if (t < n0) col = cb0 + t;
    else if (t - n0 < ce1 - cb1) col = cb1 + (t - n0);
    else return;
    for (int s = 0; s < w; ++s) {
        const int r = k0 + s, p = ipiv[r];
        if (p != r) {
            C2<R> x = A[(int64_t)r * lda + col];
            A[(int64_t)r * lda + col] = A[(int64_t)p * lda + col];
            A[(int64_t)p * lda + col] = x;
        }
    }
}

// Inverses of the diagonal block's unit-lower L and upper U (bs <= 64),
// stored row-major as inv[(2 kb) * 4096] (L) and inv[(2 kb + 1) * 4096] (U).
// Four lanes own one column and split each substitution dot product by the
// summation index mod 4 (combined with two xor shuffles), so no block-wide
// barrier is needed inside a substitution.
template <typename R>
__global__ __launch_bounds__(256) void k_trtri(const C2<R>* __restrict__ A, int64_t lda, int k0, int bs,
                                               C2<R>* __restrict__ inv) {
    __shared__ C2<R> D[kNB][kNB + 1];
    __shared__ C2<R> X[kNB][kNB + 1];
    const int tid = threadIdx.x;
    for (int e = tid; e < kNB * kNB; e += 256) {
        const int r = e / kNB, c = e % kNB;
        D[r][c] = (r < bs && c < bs) ? A[(int64_t)(k0 + r) * lda + k0 + c] : C2<R>{0, 0};
    }
    __syncthreads();
    const int c = tid >> 2;  // column
    const int h = tid & 3;   // summation index mod 4
    C2<R>* out = inv + (size_t)(2 * (k0 / kNB)) * kNB * kNB;
    // inv(L): x_r = delta_rc - sum_{c <= s < r} L_rs x_s
    for (int r = 0; r < kNB; ++r) {
        double sx = 0, sy = 0;
        if (r > c && r < bs) {
            for (int s2 = c + h; s2 < r; s2 += 4) {
                const C2<R> d = D[r][s2], x = X[s2][c];
                sx += (double)d.x * x.x - (double)d.y * x.y;
                sy += (double)d.x * x.y + (double)d.y * x.x;
            }
        }
        sx += __shfl_xor(sx, 1, 64);
        sy += __shfl_xor(sy, 1, 64);
        sx += __shfl_xor(sx, 2, 64);
        sy += __shfl_xor(sy, 2, 64);
        if (h == 0) {
            C2<R> v{0, 0};
            if (r < bs && r >= c) v = (r == c) ? C2<R>{1, 0} : C2<R>{(R)-sx, (R)-sy};
            X[r][c] = v;
        }
    }
    __syncthreads();
    for (int e = tid; e < kNB * kNB; e += 256) out[e] = X[e / kNB][e % kNB];
    __syncthreads();
    // inv(U): x_r = (delta_rc - sum_{r < s <= c} U_rs x_s) / U_rr
    for (int r = kNB - 1; r >= 0; --r) {
        double sx = 0, sy = 0;
        if (r < c && c < bs) {
            for (int s2 = r + 1 + h; s2 <= c; s2 += 4) {
                const C2<R> d = D[r][s2], x = X[s2][c];
                sx += (double)d.x * x.x - (double)d.y * x.y;
                sy += (double)d.x * x.y + (double)d.y * x.x;
            }
        }
        sx += __shfl_xor(sx, 1, 64);
        sy += __shfl_xor(sy, 1, 64);
        sx += __shfl_xor(sx, 2, 64);
        sy += __shfl_xor(sy, 2, 64);
        if (h == 0) {
            C2<R> v{0, 0};
            if (r <= c && c < bs) {
                const C2<R> num = (r == c) ? C2<R>{1, 0} : C2<R>{(R)-sx, (R)-sy};
                v = cmul(num, cinv(D[r][r]));
            }
            X[r][c] = v;
        }
    }
    __syncthreads();
    for (int e = tid; e < kNB * kNB; e += 256) out[kNB * kNB + e] = X[e / kNB][e % kNB];
}

// Rows k0 .. k0+bs of columns [cb, ce): X = inv(L_kb) X on 64-column strips,
// in place (L_kb is the unit-lower diagonal block of block k0 / 64).
template <typename R>
__global__ __launch_bounds__(256) void k_trsm(C2<R>* __restrict__ A, int64_t lda, int k0, int bs, int cb, int ce,
                                              const C2<R>* __restrict__ inv) {
    __shared__ C2<R> Li[kNB][kNB];
    __shared__ C2<R> S[kNB][kNB];
    const int c0 = cb + blockIdx.x * kNB;
    const int tid = threadIdx.x;
    const C2<R>* Lg = inv + (size_t)(2 * (k0 / kNB)) * kNB * kNB;
    for (int e = tid; e < kNB * kNB; e += 256) {
        const int r = e / kNB, c = e % kNB;
        Li[r][c] = Lg[e];
        S[r][c] = (r < bs && c0 + c < ce) ? A[(int64_t)(k0 + r) * lda + c0 + c] : C2<R>{0, 0};
    }
    __syncthreads();
    const int c = tid & 63, rg = tid >> 6;
    for (int rr = 0; rr < 16; ++rr) {
        const int r = rg * 16 + rr;
        if (r >= bs || c0 + c >= ce) continue;
        C2<R> v{0, 0};
        for (int s2 = 0; s2 <= r; ++s2) v = cadd(v, cmul(Li[r][s2], S[s2][c]));
        A[(int64_t)(k0 + r) * lda + c0 + c] = v;
    }
}

// C -= A B, complex, row-major; MFMA 16x16x4 with four real products.
template <typename R>
struct Mfma;
// Operand layout (both): A[i][k] from lane i + 16k, B[k][j] from lane j + 16k.
// Result slot r of lane l holds D[row(l, r)][l % 16]; the row rule differs
// between the two instructions on gfx950 (measured with a one-wave probe):
//   f64 16x16x4: row = l/16 + 4r        f32 16x16x4: row = 4(l/16) + r
template <>
struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    __device__ static inline v4 op(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    __device__ static inline v4 op(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int kGM = 128, kGN = 128, kGK = 16;

// 512 threads: 8 waves as 2 (rows) x 4 (columns), each wave a 64 x 32 tile of
// 4 x 2 MFMA blocks (lane layouts in Mfma<R>).
template <typename R>
__global__ __launch_bounds__(512, 1) void k_gemm_sub(const C2<R>* __restrict__ A, int64_t lda,
                                                     const C2<R>* __restrict__ B, int64_t ldb,
                                                     C2<R>* __restrict__ C, int64_t ldc, int M, int N, int K) {
    typedef typename Mfma<R>::v4 v4;
    __shared__ R sA[2][kGK][kGM];
    __shared__ R sB[2][kGK][kGN];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = (wv >> 2) * 64, wn = (wv & 3) * 32;
    const int m0 = blockIdx.y * kGM, n0 = blockIdx.x * kGN;
    v4 accR[4][2], accI[4][2];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            accR[x][y] = v4{0, 0, 0, 0};
            accI[x][y] = v4{0, 0, 0, 0};
        }
    const int lm = tid & 127, lk = (tid >> 7) * 4;
    C2<R> ra[4], rb[4];
    auto gload = [&](int kc) {
        const int gm = m0 + lm, gn = n0 + lm;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int gk = kc + lk + t;
            ra[t] = (gm < M && gk < K) ? A[(int64_t)gm * lda + gk] : C2<R>{0, 0};
            rb[t] = (gn < N && gk < K) ? B[(int64_t)gk * ldb + gn] : C2<R>{0, 0};
        }
    };
    gload(0);
    for (int kc = 0; kc < K; kc += kGK) {
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            sA[0][lk + t][lm] = ra[t].x;
            sA[1][lk + t][lm] = ra[t].y;
            sB[0][lk + t][lm] = rb[t].x;
            sB[1][lk + t][lm] = rb[t].y;
        }
        __syncthreads();
        if (kc + kGK < K) gload(kc + kGK);
#pragma unroll
        for (int kk = 0; kk < kGK; kk += 4) {
            const int kr = kk + (lane >> 4), cl = lane & 15;
            R ar[4], ai[4], na[4], br[2], bi[2];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                ar[q] = sA[0][kr][wm + q * 16 + cl];
                ai[q] = sA[1][kr][wm + q * 16 + cl];
                na[q] = -ai[q];
            }
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                br[q] = sB[0][kr][wn + q * 16 + cl];
                bi[q] = sB[1][kr][wn + q * 16 + cl];
            }
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) accR[x][y] = Mfma<R>::op(ar[x], br[y], accR[x][y]);
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) accI[x][y] = Mfma<R>::op(ar[x], bi[y], accI[x][y]);
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) accR[x][y] = Mfma<R>::op(na[x], bi[y], accR[x][y]);
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) accI[x][y] = Mfma<R>::op(ai[x], br[y], accI[x][y]);
        }
    }
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm + x * 16 + Mfma<R>::row(lane, r);
                const int gn = n0 + wn + y * 16 + (lane & 15);
                if (gm < M && gn < N) {
                    C2<R> c = C[(int64_t)gm * ldc + gn];
                    c.x -= accR[x][y][r];
                    c.y -= accI[x][y][r];
                    C[(int64_t)gm * ldc + gn] = c;
                }
            }
}

// ---------------------------------------------------------------------------
// drivers
// ---------------------------------------------------------------------------
struct ProfHook;  // defined in the API unit; null when profiling is off
void prof_begin(void* prof, hipStream_t s, int cls);
void prof_end(void* prof, hipStream_t s, int cls, double flops, double bytes);
enum { CLS_PANEL = 2, CLS_GEMM = 3, CLS_TRSM = 4, CLS_SWAP = 5, CLS_SOLVE = 6 };

int panel_capacity(int cu_count) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_lu_panel<double>), 256, 0) != hipSuccess)
        return 0;
    int per_cu32 = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu32, reinterpret_cast<const void*>(&k_lu_panel<float>), 256, 0) != hipSuccess)
        return 0;
    return (per_cu < per_cu32 ? per_cu : per_cu32) * cu_count;
}

template <typename R>
static void launch_laswp(hipStream_t s, C2<R>* A, int64_t lda, int k0, int w, const int* ipiv, int cb0, int ce0,
                         int cb1, int ce1, void* prof) {
    const int cols = (ce0 - cb0) + (ce1 - cb1);
    if (cols <= 0) return;
    prof_begin(prof, s, CLS_SWAP);
    hipLaunchKernelGGL((k_laswp<R>), dim3((cols + 255) / 256), dim3(256), 0, s, A, lda, k0, w, ipiv, cb0, ce0, cb1, ce1);
    prof_end(prof, s, CLS_SWAP, 0.0, 4.0 * (double)cols * w * sizeof(C2<R>));
}

template <typename R>
static void launch_trsm(hipStream_t s, C2<R>* A, int64_t lda, int k0, int bs, int cb, int ce, const C2<R>* inv,
                        void* prof) {
    if (ce <= cb) return;
    prof_begin(prof, s, CLS_TRSM);
    hipLaunchKernelGGL((k_trsm<R>), dim3((ce - cb + kNB - 1) / kNB), dim3(256), 0, s, A, lda, k0, bs, cb, ce, inv);
    prof_end(prof, s, CLS_TRSM, 8.0 * (double)bs * bs * (ce - cb) / 2.0, 2.0 * (double)bs * (ce - cb) * sizeof(C2<R>));
}

// C[M x N] -= A[M x K] B[K x N] with C at (r0, c0), A at (r0, ka), B at (ka, c0).
template <typename R>
static void launch_update(hipStream_t s, C2<R>* A, int64_t lda, int r0, int c0, int ka, int M, int N, int K,
                          void* prof) {
    if (M <= 0 || N <= 0 || K <= 0) return;
    prof_begin(prof, s, CLS_GEMM);
    dim3 grid((N + kGN - 1) / kGN, (M + kGM - 1) / kGM);
    hipLaunchKernelGGL((k_gemm_sub<R>), grid, dim3(512), 0, s, A + (int64_t)r0 * lda + ka, lda,
                       A + (int64_t)ka * lda + c0, lda, A + (int64_t)r0 * lda + c0, lda, M, N, K);
    prof_end(prof, s, CLS_GEMM, 8.0 * (double)M * N * K, (2.0 * (double)M * N + (double)(M + N) * K) * sizeof(C2<R>));
}

// Two-level right-looking LU: 64-wide panels inside 256-wide outer blocks, so
// the trailing update of each outer block is one K = 256 MFMA GEMM.
constexpr int kOuter = 256;

template <typename R>
int lu_factor(hipStream_t s, C2<R>* A, int64_t lda, int n, int* ipiv, C2<R>* inv, PanelWS& ws, void* prof) {
    for (int K0 = 0; K0 < n; K0 += kOuter) {
        const int KW = (n - K0) < kOuter ? (n - K0) : kOuter;
        const int Kend = K0 + KW;
        for (int k0 = K0; k0 < Kend; k0 += kNB) {
            const int w = (Kend - k0) < kNB ? (Kend - k0) : kNB;
            const int m = n - k0;
            const int G = (m + kPanelRows - 1) / kPanelRows;
            if (G > ws.max_blocks) return 2;
            int k0v = k0, wv = w, nv = n;
            int64_t ldv = lda;
            C2<R>* Av = A;
            int* ipv = ipiv;
            PanelWS wsv = ws;
            unsigned epoch = ws.epoch;
            ws.epoch += kNB;
            void* args[] = {&Av, &ldv, &nv, &k0v, &wv, &ipv, &wsv, &epoch};
            prof_begin(prof, s, CLS_PANEL);
            if (hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_lu_panel<R>), dim3(G), dim3(256), args, 0,
                                           s) != hipSuccess)
                return 3;
            prof_end(prof, s, CLS_PANEL, 8.0 * (double)m * w * w / 2.0, 2.0 * (double)m * w * sizeof(C2<R>));
            // this panel's interchanges inside the outer block
            launch_laswp<R>(s, A, lda, k0, w, ipiv, K0, k0, k0 + w, Kend, prof);
            hipLaunchKernelGGL((k_trtri<R>), dim3(1), dim3(256), 0, s, A, lda, k0, w, inv);
            if (k0 + w < Kend) {
                launch_trsm<R>(s, A, lda, k0, w, k0 + w, Kend, inv, prof);
                launch_update<R>(s, A, lda, k0 + w, k0 + w, k0, n - k0 - w, Kend - k0 - w, w, prof);
            }
        }
        // the outer block's interchanges on every other column
        launch_laswp<R>(s, A, lda, K0, KW, ipiv, 0, K0, Kend, n, prof);
        if (Kend < n) {
            // U12 = inv(L11) A12 by 64-row blocks, then the K = KW trailing update
            for (int i0 = K0; i0 < Kend; i0 += kNB) {
                const int wi = (Kend - i0) < kNB ? (Kend - i0) : kNB;
                launch_trsm<R>(s, A, lda, i0, wi, Kend, n, inv, prof);
                launch_update<R>(s, A, lda, i0 + wi, Kend, i0, Kend - i0 - wi, n - Kend, wi, prof);
            }
            launch_update<R>(s, A, lda, Kend, Kend, K0, n - Kend, n - Kend, KW, prof);
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// triangular solves with the stored block inverses
// ---------------------------------------------------------------------------
template <typename R>
__global__ void k_gather(int n, const int* perm, const C2<R>* b, C2<R>* y) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = b[perm[i]];
}

// Block kb of a forward (lower = 1) or backward (lower = 0) substitution:
//   z = inv(D_kb) y[block]; out[block] = z;
//   y[r] -= sum_c T[r][r0 + c] z[c] for rows r after (forward) / before (backward).
template <typename R, int LOWER>
__global__ __launch_bounds__(256) void k_trsv_block(const C2<R>* __restrict__ T, int64_t lda, int n, int kb,
                                                    const C2<R>* __restrict__ inv, C2<R>* __restrict__ y,
                                                    C2<R>* __restrict__ out) {
    __shared__ C2<R> z[kNB];
    const int r0 = kb * kNB;
    const int bs = (n - r0) < kNB ? (n - r0) : kNB;
    const int tid = threadIdx.x;
    const C2<R>* Di = inv + (size_t)(2 * kb + (LOWER ? 0 : 1)) * kNB * kNB;
    if (tid < kNB) {
        C2<R> v{0, 0};
        if (tid < bs)
            for (int c = 0; c < bs; ++c) v = cadd(v, cmul(Di[tid * kNB + c], y[r0 + c]));
        z[tid] = v;
        if (blockIdx.x == 0 && tid < bs) out[r0 + tid] = v;
    }
    __syncthreads();
    const int lane = tid & 63, wv = tid >> 6;
    const int sub = lane >> 4, l16 = lane & 15;
    const int rbeg = LOWER ? r0 + bs : 0;
    const int rend = LOWER ? n : r0;
    for (int base = rbeg + (blockIdx.x * 4 + wv) * 4; base < rend; base += gridDim.x * 16) {
        const int r = base + sub;
        double sx = 0, sy = 0;
        if (r < rend) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int c = l16 * 4 + q;
                if (c < bs) {
                    const C2<R> t = T[(int64_t)r * lda + r0 + c];
                    const C2<R> zc = z[c];
                    sx += (double)t.x * zc.x - (double)t.y * zc.y;
                    sy += (double)t.x * zc.y + (double)t.y * zc.x;
                }
            }
        }
#pragma unroll
        for (int off = 8; off >= 1; off >>= 1) {
            sx += __shfl_xor(sx, off, 64);
            sy += __shfl_xor(sy, off, 64);
        }
        if (r < rend && l16 == 0) {
            C2<R> v = y[r];
            v.x -= (R)sx;
            v.y -= (R)sy;
            y[r] = v;
        }
    }
}

template <typename R>
void lu_solve(hipStream_t s, const C2<R>* LU, int64_t lda, int n, const int* perm, const C2<R>* inv,
              const C2<R>* b, C2<R>* work, C2<R>* x, void* prof) {
    const int nblk = (n + kNB - 1) / kNB;
    prof_begin(prof, s, CLS_SOLVE);
    hipLaunchKernelGGL((k_gather<R>), dim3((n + 255) / 256), dim3(256), 0, s, n, perm, b, work);
    for (int kb = 0; kb < nblk; ++kb) {
        const int rows = n - (kb + 1) * kNB;
        int grid = rows > 0 ? (rows + 15) / 16 : 1;
        if (grid > 1024) grid = 1024;
        hipLaunchKernelGGL((k_trsv_block<R, 1>), dim3(grid), dim3(256), 0, s, LU, lda, n, kb, inv, work, x);
    }
    // backward: x holds L^{-1} P b; reuse `work` as the running right-hand side
    (void)hipMemcpyAsync(work, x, sizeof(C2<R>) * n, hipMemcpyDeviceToDevice, s);
    for (int kb = nblk - 1; kb >= 0; --kb) {
        const int rows = kb * kNB;
        int grid = rows > 0 ? (rows + 15) / 16 : 1;
        if (grid > 1024) grid = 1024;
        hipLaunchKernelGGL((k_trsv_block<R, 0>), dim3(grid), dim3(256), 0, s, LU, lda, n, kb, inv, work, x);
    }
    prof_end(prof, s, CLS_SOLVE, 8.0 * (double)n * n, (double)n * n * sizeof(C2<R>));
}

// explicit instantiations
template void launch_row_stats<double>(hipStream_t, const z64*, int64_t, int, double*, double*);
template void launch_row_stats<float>(hipStream_t, const z32*, int64_t, int, double*, double*);
template void launch_col_max<double>(hipStream_t, const z64*, int64_t, int, const double*, double*);
template void launch_col_max<float>(hipStream_t, const z32*, int64_t, int, const double*, double*);
template void launch_scale_copy<double, double>(hipStream_t, const z64*, int64_t, int, const double*, const double*, z64*, int64_t);
template void launch_scale_copy<double, float>(hipStream_t, const z64*, int64_t, int, const double*, const double*, z32*, int64_t);
template void launch_scale_copy<float, float>(hipStream_t, const z32*, int64_t, int, const double*, const double*, z32*, int64_t);
template void launch_scale_copy<float, double>(hipStream_t, const z32*, int64_t, int, const double*, const double*, z64*, int64_t);
template void launch_matvec<double>(hipStream_t, const z64*, int64_t, int, int, const z64*, z64*, const z64*, double*);
template void launch_matvec<float>(hipStream_t, const z32*, int64_t, int, int, const z64*, z64*, const z64*, double*);
template void launch_prep_rhs<double>(hipStream_t, int, const z64*, const double*, double, z64*);
template void launch_prep_rhs<float>(hipStream_t, int, const z64*, const double*, double, z32*);
template void launch_finish<double>(hipStream_t, int, const z64*, double, const double*, z64*, int, int*);
template void launch_finish<float>(hipStream_t, int, const z32*, double, const double*, z64*, int, int*);
template int lu_factor<double>(hipStream_t, z64*, int64_t, int, int*, z64*, PanelWS&, void*);
template int lu_factor<float>(hipStream_t, z32*, int64_t, int, int*, z32*, PanelWS&, void*);
template void lu_solve<double>(hipStream_t, const z64*, int64_t, int, const int*, const z64*, const z64*, z64*, z64*, void*);
template void lu_solve<float>(hipStream_t, const z32*, int64_t, int, const int*, const z32*, const z32*, z32*, z32*, void*);

}  // namespace bem
