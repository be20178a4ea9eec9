// Dense complex linear algebra for the BEM systems (CDNA4 / gfx950).
//
// Matrices are row-major complex (C2<R>), R = double (plain FP64 path) or
// float (mixed path).  Row-major keeps the partial-pivoting row interchanges
// and the panel's per-row reads contiguous.
//
// LU: right-looking, 64-wide panels inside 512-wide outer blocks.  The
// reference factors with LAPACK zgetrf/cgetrf (partial pivoting,
// src/solvers/bem/full.jl:404, src/solvers/bem/fluid_system.jl:217); the
// panels here pivot by tournament (communication-avoiding LU), which has the
// same backward stability in practice and needs no grid-wide barrier.
//   panel   tournament pivoting (TSLU) over 128-row leaves and a log-depth
//           merge tree, then LU of the 64x64 block, its L/U inverses and
//           L21 = A21 inv(U11) on MFMA
//   swap    row interchanges outside the panel (coalesced row segments)
//   trsm    U12 = inv(L11) A12
//   gemm    A22 -= L21 U12 on MFMA: v_mfma_f32_16x16x4f32 with three real
//           products per complex product (FP32), v_mfma_f64_16x16x4f64 with
//           four (FP64); 128x128 tiles staged in LDS
// Solves use the stored block inverses: one launch per solve, forward and
// backward phases together.
#include <algorithm>
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
// LU: 64-column panels (kNB), tournament-pivoted panel factorization, MFMA
// trailing updates.
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

// ---------------------------------------------------------------------------
// Panel factorization by tournament pivoting (CALU / TSLU: Grigori, Demmel and
// Xiang, SIAM J. Matrix Anal. Appl. 32 (2011)).  The m x w panel is split into
// 256-row leaves.  Each leaf runs partial pivoting on its rows (max |re|+|im|,
// ties to the lowest current position, as LAPACK's i?amax) and nominates w
// rows; nominations are merged four at a time up the tree, each merge again
// running partial pivoting on the (up to 4w) candidate rows, re-read from A by
// row index (A is not written while the tree runs).  The root's selection
// order is the panel's pivot order.  Every level costs w sequential
// elimination steps whatever its width, so the four-way tree (half the levels
// of a binary one) halves the panel's latency.  No grid-wide barrier: a merge
// is done by the last workgroup to arrive at a tree node (atomic ticket), the
// others exit.
//
// One 256-row elimination (512 threads, 8 waves of 32 rows): lanes r and
// r + 32 of a wave hold columns 0-31 and 32-63 of row r's active window.
// Rows never move: a selected row retires and its LAPACK position is swapped
// in bookkeeping only.  After each step the window shifts left by one column,
// so the pivot column is always register 0 and every register index is
// static.  One block barrier per step: each wave stages its best row in LDS
// (double-buffered by step parity) before it, every wave picks the global
// winner after it.
// ---------------------------------------------------------------------------
constexpr int kTsRows = 256;     // rows per leaf / merge
constexpr int kTsWaves = 8;      // 32 rows per wave, two lanes per row
constexpr int kTsThreads = 64 * kTsWaves;
constexpr int kTsArity = 4;      // children per merge
constexpr int kTsPS = 34;        // LDS stride of a half-row: 16-byte aligned, off the other half's banks

template <typename R>
struct TsCand {
    R v;
    int pos;
    C2<R> ip;  // inverse of the row's pivot value (zero for a zero pivot)
};

template <typename R>
struct TsLds {
    TsCand<R> cand[2][kTsWaves];
    __attribute__((aligned(16))) C2<R> prow[2][kTsWaves][2 * kTsPS];
    int sel[kNB];
    int ticket;
};

// Pivot key of a live row: |re| + |im|; NaN sorts below every number (it is
// picked only when nothing comparable is left); retired rows never win.
template <typename R>
__device__ __forceinline__ R ts_key(C2<R> z, bool live) {
    R v = cabs1(z);
    if (v != v) v = R(-0.5);
    return live ? v : R(-1);
}

// Cross-lane moves without LDS: DPP within 16-lane rows, readlane across rows,
// and the gfx950 half-wave swap (lane l <-> lane l ^ 32).
template <int CTRL>
__device__ __forceinline__ int dpp_i(int x) {
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_r(float x) {
    return __int_as_float(dpp_i<CTRL>(__float_as_int(x)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_r(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = dpp_i<CTRL>((int)(b & 0xffffffffll)), hi = dpp_i<CTRL>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float rdl(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double rdl(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ int rdl(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
// value of lane l ^ 32
__device__ __forceinline__ int xhalf_i(int v, bool upper) {
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return upper ? (int)r[0] : (int)r[1];
}
__device__ __forceinline__ float xhalf(float v, bool upper) { return __int_as_float(xhalf_i(__float_as_int(v), upper)); }
__device__ __forceinline__ double xhalf(double v, bool upper) {
    const long long b = __double_as_longlong(v);
    const int lo = xhalf_i((int)(b & 0xffffffffll), upper), hi = xhalf_i((int)(b >> 32), upper);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Wave-uniform max / min: butterflies inside each 16-lane row (quad swaps,
// half-row and row mirrors), then the four row results by readlane.
template <typename T>
__device__ __forceinline__ T ts_max(T a, T b) { return a > b ? a : b; }
template <typename T>
__device__ __forceinline__ T wave_max_u(T x) {
    x = ts_max(x, dpp_r<0xB1>(x));
    x = ts_max(x, dpp_r<0x4E>(x));
    x = ts_max(x, dpp_r<0x141>(x));
    x = ts_max(x, dpp_r<0x140>(x));
    return ts_max(ts_max(rdl(x, 0), rdl(x, 16)), ts_max(rdl(x, 32), rdl(x, 48)));
}
__device__ __forceinline__ int wave_min_u(int x) {
    x = min(x, dpp_i<0xB1>(x));
    x = min(x, dpp_i<0x4E>(x));
    x = min(x, dpp_i<0x141>(x));
    x = min(x, dpp_i<0x140>(x));
    return min(min(rdl(x, 0), rdl(x, 16)), min(rdl(x, 32), rdl(x, 48)));
}

// 1 / z for z != 0, scaled by |re| + |im| against overflow; only steers the
// pivot choice (the factors are recomputed from original values).
__device__ __forceinline__ C2<float> ts_inv(C2<float> z) {
    const float is = __builtin_amdgcn_rcpf(fabsf(z.x) + fabsf(z.y));
    const float x = z.x * is, y = z.y * is;
    const float id = __builtin_amdgcn_rcpf(fmaf(x, x, y * y)) * is;
    return C2<float>{x * id, -y * id};
}
__device__ __forceinline__ C2<double> ts_inv(C2<double> z) {
    const double is = 1.0 / (fabs(z.x) + fabs(z.y));
    const double x = z.x * is, y = z.y * is;
    const double id = is / fma(x, x, y * y);
    return C2<double>{x * id, -y * id};
}
// a - l p with four fused multiply-adds
template <typename R>
__device__ __forceinline__ C2<R> cfms(C2<R> a, C2<R> l, C2<R> p) {
    return C2<R>{fma(-l.x, p.x, fma(l.y, p.y, a.x)), fma(-l.x, p.y, fma(-l.y, p.x, a.y))};
}

// Where the tournament's root records its elimination, which is the LU
// without pivoting of the selected rows in pivot order: the multiplier of
// every live row at every step, the pivot row's active window (row j of U),
// and the slot of step j's pivot.  k_tslu_apply assembles the panel's
// diagonal-block factors from it instead of factoring those rows again.
constexpr int kTsRecU = 2 * kNB;  // row pitch of the recorded U windows
template <typename R>
struct TsRec {
    C2<R>* l;  // [kNB][kTsRows]: step j, slot s
    C2<R>* u;  // [kNB][kTsRecU]: row j, column c >= j; columns >= kNB are padding
    int* slot; // [kNB]
};
template <typename R>
__host__ __device__ inline TsRec<R> ts_rec(void* base) {
    C2<R>* l = static_cast<C2<R>*>(base);
    C2<R>* u = l + kNB * kTsRows;
    return TsRec<R>{l, u, reinterpret_cast<int*>(u + kNB * kTsRecU)};
}
static_assert(sizeof(C2<double>) * (kNB * kTsRows + kNB * kTsRecU) + sizeof(int) * kNB == kPanelRecBytes,
              "PanelWS::rec holds the FP64 record");

// Block barrier for LDS traffic: waits for this wave's LDS operations only.
// __syncthreads() also releases global memory, which would make every
// elimination step wait for the root's record stores (nothing in the block
// reads them back; the kernel's end publishes them).
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Elimination steps [j0, j1) with a live window of W columns per lane (32
// while the upper half-row still holds live columns, 16 once fewer than 16
// remain).  The winner of step j writes its global row index to sh.sel[j].
// Each wave publishes its best live row, its key and the inverse of its
// pivot value before the barrier; after it, every lane loads the winning
// half-row in one batch of LDS reads (8-value groups for FP64, to bound the
// registers) and eliminates.
template <typename R, int W>
__device__ __forceinline__ void ts_steps(C2<R> (&a)[32], int j0, int j1, bool& done, int& mypos, int myrow,
                                         int slot, TsLds<R>& sh, const TsRec<R>& rec, bool record) {
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const bool upper = lane >= 32;
    constexpr int G = sizeof(R) == sizeof(float) ? W : 8;  // pivot values per LDS batch
#pragma unroll 1
    for (int j = j0; j < j1; ++j) {
        const int b = j & 1;
        // this wave's best live row: max key, then lowest position
        const R v = upper ? R(-1) : ts_key(a[0], !done);
        const R wv = wave_max_u(v);
        const int wp = wave_min_u((!upper && v == wv) ? mypos : INT_MAX);
        const bool best = !done && mypos == wp;  // both lanes of that row
        const unsigned long long bl = __ballot(best && !upper);
        C2<R> ipw{0, 0};
        if (bl != 0) {
            const int src = __ffsll(bl) - 1;
            const C2<R> pv{rdl(a[0].x, src), rdl(a[0].y, src)};
            if (pv.x != 0 || pv.y != 0) ipw = ts_inv(pv);
        }
        if (lane == 0) sh.cand[b][g] = TsCand<R>{wv, wp, ipw};
        if (best) {
#pragma unroll
            for (int c = 0; c < W; ++c) sh.prow[b][g][(upper ? kTsPS : 0) + c] = a[c];
        }
        lds_barrier();
        // the block's winner
        TsCand<R> cq[kTsWaves];
#pragma unroll
        for (int q = 0; q < kTsWaves; ++q) cq[q] = sh.cand[b][q];
        R gv = cq[0].v;
        int gp = cq[0].pos, gw = 0;
        C2<R> ip = cq[0].ip;
#pragma unroll
        for (int q = 1; q < kTsWaves; ++q) {
            const bool t = (cq[q].v > gv) | ((cq[q].v == gv) & (cq[q].pos < gp));
            gv = t ? cq[q].v : gv;
            gp = t ? cq[q].pos : gp;
            gw = t ? q : gw;
            ip.x = t ? cq[q].ip.x : ip.x;
            ip.y = t ? cq[q].ip.y : ip.y;
        }
        const C2<R>* pr = sh.prow[b][gw] + (upper ? kTsPS : 0);
        const C2<R> l0 = cmul(a[0], ip);
        if (record) {
            // lower lane: this row's multiplier.  Wave 0: the winner's window,
            // staged in LDS for the elimination, one value per lane (lower
            // half: columns j .. j + 31, upper: j + 32 .. j + 63; beyond 63 the
            // stores land in the row's padding, j + 63 < kTsRecU).
            if (!upper && !done) rec.l[j * kTsRows + slot] = l0;
            if (g == 0) rec.u[j * kTsRecU + j + lane] = sh.prow[b][gw][(upper ? kTsPS - 32 : 0) + lane];
            if (g == gw && !done && mypos == gp && !upper) rec.slot[j] = slot;
        }
        if (g == gw && !done && mypos == gp) {
            done = true;
            if (!upper) sh.sel[j] = myrow;
        } else if (mypos == j) {
            mypos = gp;  // the row at position j takes the winner's old position
        }
        // eliminate column j and shift the window left by one column; retired
        // rows compute values nobody reads
        C2<R> l = l0, o{0, 0};
        if (W == 32) {
            // one half-wave swap: the lower lane sends the multiplier, the
            // upper lane its window column 32 (which moves to the lower lane)
            const C2<R> s = upper ? a[0] : l0;
            o = C2<R>{xhalf(s.x, upper), xhalf(s.y, upper)};
            l = upper ? o : l0;
        }
        const C2<R> c32 = W == 32 ? sh.prow[b][gw][kTsPS] : C2<R>{0, 0};
#pragma unroll
        for (int c0 = 0; c0 < W - 1; c0 += G) {
            C2<R> p[G];
#pragma unroll
            for (int t = 0; t < G; ++t)
                if (c0 + t + 1 < W) p[t] = pr[c0 + t + 1];
#pragma unroll
            for (int t = 0; t < G; ++t)
                if (c0 + t < W - 1) a[c0 + t] = cfms(a[c0 + t + 1], l, p[t]);
        }
        if (W == 32) a[W - 1] = upper ? C2<R>{0, 0} : cfms(o, l, c32);
    }
}

// Partial pivoting over the block's nv rows (w columns); on return
// sh.sel[0 .. cnt) holds the selected global rows in pivot order.  The root
// records its factors in `rec` (record = true).
template <typename R>
__device__ int ts_eliminate(C2<R> (&a)[32], int nv, int w, int myrow, TsLds<R>& sh, TsRec<R> rec, bool record) {
    const int slot = (threadIdx.x >> 6) * 32 + (threadIdx.x & 31);
    const int cnt = nv < w ? nv : w;
    bool done = myrow < 0;
    int mypos = done ? 1024 + slot : slot;
    // the row loads have landed: otherwise the step loop would wait on its
    // first use of them in every iteration, and on gfx950 that count also
    // covers the root's record stores
    __builtin_amdgcn_s_waitcnt(0);
    ts_steps<R, 32>(a, 0, cnt < 48 ? cnt : 48, done, mypos, myrow, slot, sh, rec, record);
    ts_steps<R, 16>(a, 48, cnt, done, mypos, myrow, slot, sh, rec, record);
    __syncthreads();
    return cnt;
}

// Window columns [32 part, 32 part + 32) of `row` (zero beyond w or for row < 0).
template <typename R>
__device__ __forceinline__ void ts_load_row(C2<R> (&a)[32], const C2<R>* __restrict__ A, int64_t lda, int row,
                                            int col0, int w, int part) {
#pragma unroll
    for (int c = 0; c < 32; ++c) {
        const int cc = part * 32 + c;
        a[c] = (row >= 0 && cc < w) ? A[(int64_t)row * lda + col0 + cc] : C2<R>{0, 0};
    }
}

// Nodes of tree level L (leaves are level 0) and the index of its first node.
__device__ inline int ts_level_size(int G, int L) {
    int sz = G;
    for (int i = 0; i < L; ++i) sz = (sz + kTsArity - 1) / kTsArity;
    return sz;
}
__device__ inline int ts_level_off(int G, int L) {
    int off = 0, sz = G;
    for (int i = 0; i < L; ++i) {
        off += sz;
        sz = (sz + kTsArity - 1) / kTsArity;
    }
    return off;
}

template <typename R>
__global__ __launch_bounds__(kTsThreads) void k_tslu(const C2<R>* __restrict__ A, int64_t lda, int n, int k0, int w,
                                                     PanelWS ws) {
    __shared__ TsLds<R> sh;
    const int G = gridDim.x;
    const int tid = threadIdx.x, part = (tid & 63) >> 5;
    const int slot = (tid >> 6) * 32 + (tid & 31);
    const int m = n - k0;
    int level = 0, idx = blockIdx.x;
    int nv = min(kTsRows, m - idx * kTsRows);
    int myrow = slot < nv ? k0 + idx * kTsRows + slot : -1;
    if (tid < kNB) sh.sel[tid] = -1;
    C2<R> a[32];
    const TsRec<R> rec = ts_rec<R>(ws.rec);
    ts_load_row(a, A, lda, myrow, k0, w, part);
    int cnt = ts_eliminate(a, nv, w, myrow, sh, rec, G == 1);
    for (;;) {
        const int size = ts_level_size(G, level);
        if (size == 1) {
            // root: the panel's pivot rows in order
            if (tid < cnt) st_ag(&ws.root[1 + tid], sh.sel[tid]);
            if (tid == 0) st_ag(&ws.root[0], cnt);
            return;
        }
        const int nd = ts_level_off(G, level) + idx;
        if (tid < cnt) st_ag(&ws.node_rows[nd * kNB + tid], sh.sel[tid]);
        if (tid == 0) st_ag(&ws.node_cnt[nd], cnt);
        wait_mem();
        __syncthreads();
        const int parent = ts_level_off(G, level + 1) + idx / kTsArity;
        const int first = idx - idx % kTsArity;  // the parent's first child
        const int nch = min(kTsArity, size - first);
        if (nch > 1) {
            if (tid == 0) {
                const int t = __hip_atomic_fetch_add(&ws.tickets[parent], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (t == nch - 1) st_ag(&ws.tickets[parent], 0);
                sh.ticket = t;
            }
            __syncthreads();
            if (sh.ticket != nch - 1) return;  // a later sibling will merge
        }
        // merge the parent's children, their nominations concatenated in order
        const int c0 = ts_level_off(G, level) + first;
        int cum[kTsArity + 1];
        cum[0] = 0;
#pragma unroll
        for (int c = 0; c < kTsArity; ++c) cum[c + 1] = cum[c] + (c < nch ? ld_ag(&ws.node_cnt[c0 + c]) : 0);
        nv = cum[kTsArity];
        myrow = -1;
#pragma unroll
        for (int c = 0; c < kTsArity; ++c)
            if (slot >= cum[c] && slot < cum[c + 1]) myrow = ld_ag(&ws.node_rows[(c0 + c) * kNB + slot - cum[c]]);
        if (slot < nv && (myrow < k0 || myrow >= n)) {  // never expected: flag it, keep every access in bounds
            st_ag(&ws.status[1], 1);
            myrow = -1;
        }
        ts_load_row(a, A, lda, myrow, k0, w, part);
        cnt = ts_eliminate(a, nv, w, myrow, sh, rec, ts_level_size(G, level + 1) == 1);
        level += 1;
        idx /= kTsArity;
    }
}

// Invert the 64 x 64 triangular matrix T (LDS, leading dimension kNB + 1) in
// place: unit lower (upper = 0) or upper with its diagonal (upper = 1).
// Recursive doubling: 8 x 8 diagonal blocks by substitution, then
// X21 = -inv(T22) T21 inv(T11) (lower) or X12 = -inv(T11) T12 inv(T22)
// (upper) at block sizes 8, 16, 32.  W is LDS scratch of the same shape.
template <typename R>
__device__ void tri_inverse(C2<R> (*T)[kNB + 1], C2<R> (*W)[kNB + 1], int bs, int upper) {
    const int tid = threadIdx.x;  // 256 threads
    // pad beyond bs with the identity
    for (int e = tid; e < kNB * kNB; e += 256) {
        const int r = e / kNB, c = e % kNB;
        if (r >= bs || c >= bs) T[r][c] = (r == c) ? C2<R>{1, 0} : C2<R>{0, 0};
    }
    __syncthreads();
    // 8 x 8 diagonal blocks: thread (block, column) substitutes one column
    C2<R> x[8];
    const int blk = (tid & 63) >> 3, cc8 = tid & 7, o8 = blk * 8;
    if (tid < 64) {
        if (!upper) {
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                C2<R> v = (r == cc8) ? C2<R>{1, 0} : C2<R>{0, 0};
#pragma unroll
                for (int s2 = 0; s2 < 8; ++s2)
                    if (s2 < r) v = csub(v, cmul(T[o8 + r][o8 + s2], x[s2]));
                x[r] = (r >= cc8) ? v : C2<R>{0, 0};
            }
        } else {
#pragma unroll
            for (int r = 7; r >= 0; --r) {
                C2<R> v = (r == cc8) ? C2<R>{1, 0} : C2<R>{0, 0};
#pragma unroll
                for (int s2 = 0; s2 < 8; ++s2)
                    if (s2 > r) v = csub(v, cmul(T[o8 + r][o8 + s2], x[s2]));
                x[r] = (r <= cc8) ? cmul(v, cinv(T[o8 + r][o8 + r])) : C2<R>{0, 0};
            }
        }
    }
    __syncthreads();
    if (tid < 64) {
#pragma unroll
        for (int r = 0; r < 8; ++r) T[o8 + r][o8 + cc8] = x[r];
    }
    __syncthreads();
    for (int h = 8; h < kNB; h *= 2) {
        // pairs of h x h inverse blocks at offsets o, o + h (o = 0, 2h, ...)
        const int pairs = kNB / (2 * h);
        // step 1: W = T_off * inv(first)  (lower: T21 inv(T11); upper: T12 inv(T22))
        for (int e = tid; e < pairs * h * h; e += 256) {
            const int pr = e / (h * h), r = (e / h) % h, cc = e % h, o = pr * 2 * h;
            C2<R> v{0, 0};
            if (!upper) {
                for (int k = 0; k < h; ++k) v = cadd(v, cmul(T[o + h + r][o + k], T[o + k][o + cc]));
            } else {
                for (int k = 0; k < h; ++k) v = cadd(v, cmul(T[o + r][o + h + k], T[o + h + k][o + h + cc]));
            }
            W[o + r][cc] = v;
        }
        __syncthreads();
        // step 2: off-diagonal block = -inv(other) * W
        for (int e = tid; e < pairs * h * h; e += 256) {
            const int pr = e / (h * h), r = (e / h) % h, cc = e % h, o = pr * 2 * h;
            C2<R> v{0, 0};
            if (!upper) {
                for (int k = 0; k < h; ++k) v = cadd(v, cmul(T[o + h + r][o + h + k], W[o + k][cc]));
                T[o + h + r][o + cc] = C2<R>{-v.x, -v.y};
            } else {
                for (int k = 0; k < h; ++k) v = cadd(v, cmul(T[o + r][o + k], W[o + k][cc]));
                T[o + r][o + h + cc] = C2<R>{-v.x, -v.y};
            }
        }
        __syncthreads();
    }
}

// From the root's pivot rows: LAPACK-style ipiv, the row permutation of the
// panel columns, LU without pivoting of the 64 x 64 diagonal block, and the
// inverses of its L and U (inv[2 kb], inv[2 kb + 1]).
template <typename R>
__global__ __launch_bounds__(256) void k_tslu_apply(C2<R>* __restrict__ A, int64_t lda, int n, int k0, int w,
                                                    PanelWS ws, int* __restrict__ ipiv, C2<R>* __restrict__ inv) {
    __shared__ C2<R> D[kNB][kNB + 1];
    __shared__ C2<R> B[kNB][kNB + 1];
    __shared__ int s_pos[kNB], s_row[kNB], s_cnt, s_root[kNB];
    __shared__ int s_dst[kNB], s_src[kNB], s_nd, s_slot[kNB];
    const int tid = threadIdx.x, c = tid & 63, g = tid >> 6;
    if (tid <= kNB) {
        int v = ld_ag(&ws.root[tid]);
        // the root's rows must be w distinct rows of [k0, n); out-of-range
        // values are flagged and replaced so that every access stays in bounds
        if (tid == 0 && v != w) {
            st_ag(&ws.status[1], 1);
            v = w;
        }
        if (tid > 0 && tid <= w && (v < k0 || v >= n)) {
            st_ag(&ws.status[1], 1);
            v = k0 + tid - 1 < n ? k0 + tid - 1 : k0;
        }
        if (tid == 0) s_cnt = v;
        else s_root[tid - 1] = v;
    }
    __syncthreads();
    const int cnt = s_cnt;
    if (tid == 0) {
        // Sequential interchanges (the LAPACK ipiv convention).  Rows waiting
        // at unprocessed top-block positions are always original top-block
        // rows, and rows from below only move when selected, so the state is
        // two arrays over top-block offsets: s_row (row at position k0 + o)
        // and s_pos (position of original row k0 + o, possibly below the block).
        for (int o = 0; o < w; ++o) {
            s_row[o] = k0 + o;
            s_pos[o] = k0 + o;
        }
        for (int s2 = 0; s2 < cnt; ++s2) {
            const int t = s_root[s2];
            const bool top = t >= k0 && t < k0 + w;
            const int pt = top ? s_pos[t - k0] : t;
            ipiv[k0 + s2] = pt;
            if (pt != k0 + s2) {
                const int u = s_row[s2];
                s_row[s2] = t;
                s_pos[u - k0] = pt;
                if (pt < k0 + w) s_row[pt - k0] = u;
                if (top) s_pos[t - k0] = k0 + s2;
            }
        }
        // top-block rows that end below the block
        int nd = 0;
        for (int o = 0; o < w; ++o)
            if (s_pos[o] >= k0 + w) {
                s_dst[nd] = s_pos[o];
                s_src[nd] = k0 + o;
                ++nd;
            }
        s_nd = nd;
    }
    __syncthreads();
    const int nd = s_nd;
    // The diagonal block's LU without pivoting, as the tournament's root
    // recorded it: the root's elimination ran on exactly these rows, in this
    // order, with these pivots.  Recorded slots are range-checked like the
    // row indices above.
    const TsRec<R> rec = ts_rec<R>(ws.rec);
    if (tid < kNB) {
        int sl = tid < cnt ? rec.slot[tid] : 0;
        if (sl < 0 || sl >= kTsRows) {
            st_ag(&ws.status[1], 1);
            sl = 0;
        }
        s_slot[tid] = sl;
    }
    // the rows leaving the top block, read before any is written
    for (int t = g; t < nd; t += 4) B[t][c] = (c < w) ? A[(int64_t)s_src[t] * lda + k0 + c] : C2<R>{0, 0};
    __syncthreads();
    for (int t = g; t < nd; t += 4)
        if (c < w) A[(int64_t)s_dst[t] * lda + k0 + c] = B[t][c];
    for (int r = g; r < kNB; r += 4)
        D[r][c] = (r < cnt && c < w) ? (r > c ? rec.l[c * kTsRows + s_slot[r]] : rec.u[r * kTsRecU + c]) : C2<R>{0, 0};
    __syncthreads();
    if (tid < w && D[tid][tid].x == 0 && D[tid][tid].y == 0) atomicMin(&ws.status[0], k0 + tid);
    for (int r = g; r < w; r += 4)
        if (c < w) A[(int64_t)(k0 + r) * lda + k0 + c] = D[r][c];
    __syncthreads();
    // inv(L): copy the strict lower part with a unit diagonal
    for (int e = tid; e < kNB * kNB; e += 256) {
        const int r = e / kNB, cc = e % kNB;
        B[r][cc] = (r == cc) ? C2<R>{1, 0} : (r > cc ? D[r][cc] : C2<R>{0, 0});
    }
    __syncthreads();
    tri_inverse<R>(B, D, w, 0);  // D is scratch now; reload U afterwards
    C2<R>* out = inv + (size_t)(2 * (k0 / kNB)) * kNB * kNB;
    for (int e = tid; e < kNB * kNB; e += 256) out[e] = B[e / kNB][e % kNB];
    __syncthreads();
    for (int e = tid; e < kNB * kNB; e += 256) {
        const int r = e / kNB, cc = e % kNB;
        B[r][cc] = (r <= cc && r < w && cc < w) ? A[(int64_t)(k0 + r) * lda + k0 + cc] : C2<R>{0, 0};
    }
    __syncthreads();
    tri_inverse<R>(B, D, w, 1);
    for (int e = tid; e < kNB * kNB; e += 256) out[kNB * kNB + e] = B[e / kNB][e % kNB];
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
// FULL: M and N multiples of the tile and K of the chunk; no edge tests.
template <typename R, int ASSIGN, bool FULL>
__global__ __launch_bounds__(512, 1) void k_gemm_sub(const C2<R>* __restrict__ A, int64_t lda,
                                                     const C2<R>* __restrict__ B, int64_t ldb,
                                                     C2<R>* __restrict__ C, int64_t ldc, int M, int N, int K) {
    typedef typename Mfma<R>::v4 v4;
    // double-buffered LDS (128 KB in FP64): one barrier per K chunk
    __shared__ R sA[2][2][kGK][kGM];
    __shared__ R sB[2][2][kGK][kGN];
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
            ra[t] = (FULL || (gm < M && gk < K)) ? A[(int64_t)gm * lda + gk] : C2<R>{0, 0};
            rb[t] = (FULL || (gn < N && gk < K)) ? B[(int64_t)gk * ldb + gn] : C2<R>{0, 0};
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            sA[buf][0][lk + t][lm] = ra[t].x;
            sA[buf][1][lk + t][lm] = ra[t].y;
            sB[buf][0][lk + t][lm] = rb[t].x;
            sB[buf][1][lk + t][lm] = rb[t].y;
        }
    };
    gload(0);
    lstore(0);
    __syncthreads();
    int buf = 0;
    for (int kc = 0; kc < K; kc += kGK) {
        const bool more = kc + kGK < K;
        if (more) gload(kc + kGK);
#pragma unroll
        for (int kk = 0; kk < kGK; kk += 4) {
            const int kr = kk + (lane >> 4), cl = lane & 15;
            R ar[4], ai[4], na[4], br[2], bi[2];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                ar[q] = sA[buf][0][kr][wm + q * 16 + cl];
                ai[q] = sA[buf][1][kr][wm + q * 16 + cl];
                na[q] = -ai[q];
            }
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                br[q] = sB[buf][0][kr][wn + q * 16 + cl];
                bi[q] = sB[buf][1][kr][wn + q * 16 + cl];
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
        if (more) lstore(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm + x * 16 + Mfma<R>::row(lane, r);
                const int gn = n0 + wn + y * 16 + (lane & 15);
                if (FULL || (gm < M && gn < N)) {
                    if (ASSIGN) {
                        C[(int64_t)gm * ldc + gn] = C2<R>{accR[x][y][r], accI[x][y][r]};
                    } else {
                        C2<R> c = C[(int64_t)gm * ldc + gn];
                        c.x -= accR[x][y][r];
                        c.y -= accI[x][y][r];
                        C[(int64_t)gm * ldc + gn] = c;
                    }
                }
            }
}

// C -= A B (or C = A B), complex FP32, row-major; 3M products on MFMA (Gauss:
// T1 = Ar Br, T2 = Ai Bi, T3 = (Ar + Ai)(Br + Bi); Re = T1 - T2,
// Im = T3 - T1 - T2), 25 % fewer matrix-core products than four real GEMMs.
// 512 threads = 8 waves as 2 (rows) x 4 (columns), each wave 64 x 32 (4 x 2
// MFMA tiles).  K runs in chunks of 32 through double-buffered LDS holding the
// re and im planes of both operands k-contiguous (A as [m][k], B as [n][k],
// rows padded to 36 floats so a wave's 16-byte reads hit distinct banks).  A
// lane reads four consecutive k of its row with one 16-byte LDS read; the
// four MFMA steps of a 16-wide k group take one of the four each (MFMA step t
// covers k = 4q + t, q = 0..3, for A and B alike), and Ar + Ai, Br + Bi are
// formed in registers.  One barrier per 32-wide chunk; the next chunk's
// global loads overlap the chunk's 192 MFMAs per wave.  Workgroups are mapped to tiles so that each XCD (workgroups are
// dealt to the 8 XCDs round-robin) owns a contiguous range of tiles, walked in
// groups of 8 tile rows, so the A and B panels its CUs share stay in its L2.
constexpr int kG3K = 32;       // K chunk
constexpr int kG3P = 36;       // LDS row pitch (floats)
constexpr int kG3Group = 8;    // tile rows per group (default; the tuner may pick 4 or 16)

__device__ __forceinline__ void g3_tile(int nx, int ny, int group, int& tm, int& tn) {
    const int L = blockIdx.y * nx + blockIdx.x;
    const int T = nx * ny;
    const int q = T >> 3, r = T & 7, xcd = L & 7, idx = L >> 3;
    const int lin = xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
    const int per = group * nx;
    const int g = lin / per, first = g * group;
    const int gs = (ny - first) < group ? (ny - first) : group;
    const int w = lin - g * per;
    tm = first + w % gs;
    tn = w / gs;
}

// FULL: M and N multiples of the tile and K of the chunk (every trailing
// update of a 128-aligned system), so no load, store or product needs an edge
// test; the edge handling cost 7 % when it was compiled into every launch.
template <int ASSIGN, bool FULL>
__global__ __launch_bounds__(512, 1) void k_gemm3m(const z32* __restrict__ A, int64_t lda, const z32* __restrict__ B,
                                                   int64_t ldb, z32* __restrict__ C, int64_t ldc, int M, int N,
                                                   int K, int group) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    constexpr int KV = kG3P / 4;  // v4 slots per LDS row
    __shared__ v4 sA[2][2][kGM][KV];
    __shared__ v4 sB[2][2][kGN][KV];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: tile offsets stay scalar
    const int wm = (wv >> 2) * 64, wn = (wv & 3) * 32;
    int tm, tn;
    g3_tile(gridDim.x, gridDim.y, group, tm, tn);
    const int m0 = tm * kGM, n0 = tn * kGN;
    const bool full = FULL || (m0 + kGM <= M && n0 + kGN <= N);
    const int li = lane & 15, lq = lane >> 4;
    v4 t1[4][2], t2[4][2], t3[4][2];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) t1[x][y] = t2[x][y] = t3[x][y] = v4{0, 0, 0, 0};
    // lane (li, lq) holds rows 4 lq + r and column li of each 16 x 16 block
    const z32* cl = C + (int64_t)(4 * lq) * ldc + li;
    if (!ASSIGN) {
        // C - A B: the accumulators start from C (t1 = -Re C, t3 = -Re C - Im C),
        // so C is read once, up front, and the epilogue only stores
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int rb = m0 + wm + x * 16 + r, cb = n0 + wn + y * 16;
                    z32 c{0.f, 0.f};
                    if (full || (rb + 4 * lq < M && cb + li < N)) c = cl[(int64_t)rb * ldc + cb];
                    t1[x][y][r] = -c.x;
                    t3[x][y][r] = -c.x - c.y;
                }
    }
    // staging: A row m0 + lm and B column n0 + lm, k = kc + lk .. kc + lk + 7.
    // Rows, columns and k beyond the matrix are clamped to its last one, so
    // every load is issued unconditionally into the same registers, and the
    // values that lie outside are zeroed when they are stored to LDS, after
    // the chunk's products: no branch between the loads and their first use,
    // whose merge would force a wait for the prefetch before the products.
    constexpr int PT = kG3K / 4;  // k values per thread and chunk
    const int lm = tid & 127, lk = (tid >> 7) * PT;
    const bool okm = m0 + lm < M, okn = n0 + lm < N;
    const z32* pa = A + (int64_t)(okm ? m0 + lm : M - 1) * lda + lk;
    const z32* pb = B + (int64_t)lk * ldb + (okn ? n0 + lm : N - 1);
    z32 ra[PT], rb[PT];
    auto gload = [&](int kc) {
        if (FULL || kc + kG3K <= K) {
#pragma unroll
            for (int t = 0; t < PT; ++t) {
                ra[t] = pa[kc + t];
                rb[t] = pb[(int64_t)(kc + t) * ldb];
            }
        } else {
            // k clamped to K - 1 (an offset relative to lk, so possibly negative)
#pragma unroll
            for (int t = 0; t < PT; ++t) {
                const int dk = min(kc + t, K - 1 - lk);
                ra[t] = pa[dk];
                rb[t] = pb[(int64_t)dk * ldb];
            }
        }
    };
    auto lstore = [&](int buf, int kc) {
        if (!FULL && (!full || kc + kG3K > K)) {
#pragma unroll
            for (int t = 0; t < PT; ++t) {
                const bool okk = kc + lk + t < K;
                if (!(okm && okk)) ra[t] = z32{0.f, 0.f};
                if (!(okn && okk)) rb[t] = z32{0.f, 0.f};
            }
        }
#pragma unroll
        for (int h = 0; h < PT / 4; ++h) {
            const int q = (lk >> 2) + h;
            sA[buf][0][lm][q] = v4{ra[4 * h].x, ra[4 * h + 1].x, ra[4 * h + 2].x, ra[4 * h + 3].x};
            sA[buf][1][lm][q] = v4{ra[4 * h].y, ra[4 * h + 1].y, ra[4 * h + 2].y, ra[4 * h + 3].y};
            sB[buf][0][lm][q] = v4{rb[4 * h].x, rb[4 * h + 1].x, rb[4 * h + 2].x, rb[4 * h + 3].x};
            sB[buf][1][lm][q] = v4{rb[4 * h].y, rb[4 * h + 1].y, rb[4 * h + 2].y, rb[4 * h + 3].y};
        }
    };
    gload(0);
    lstore(0, 0);
    __syncthreads();
    int buf = 0;
    for (int kc = 0; kc < K; kc += kG3K) {
        const bool more = kc + kG3K < K;
        if (more) gload(kc + kG3K);
#pragma unroll
        for (int h = 0; h < kG3K / 16; ++h) {
            // each 16-wide k group reads its operands once the previous
            // group's products have issued (hoisting all reads of the chunk
            // holds 48 more registers and spills the full-tile form)
            __builtin_amdgcn_sched_barrier(0);
            v4 ar[4], ai[4], br[2], bi[2];
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                ar[x] = sA[buf][0][wm + 16 * x + li][4 * h + lq];
                ai[x] = sA[buf][1][wm + 16 * x + li][4 * h + lq];
            }
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                br[y] = sB[buf][0][wn + 16 * y + li][4 * h + lq];
                bi[y] = sB[buf][1][wn + 16 * y + li][4 * h + lq];
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
#pragma unroll
                for (int x = 0; x < 4; ++x)
#pragma unroll
                    for (int y = 0; y < 2; ++y) t1[x][y] = Mfma<float>::op(ar[x][t], br[y][t], t1[x][y]);
#pragma unroll
                for (int x = 0; x < 4; ++x)
#pragma unroll
                    for (int y = 0; y < 2; ++y) t2[x][y] = Mfma<float>::op(ai[x][t], bi[y][t], t2[x][y]);
                float as[4], bs[2];
#pragma unroll
                for (int x = 0; x < 4; ++x) as[x] = ar[x][t] + ai[x][t];
#pragma unroll
                for (int y = 0; y < 2; ++y) bs[y] = br[y][t] + bi[y][t];
#pragma unroll
                for (int x = 0; x < 4; ++x)
#pragma unroll
                    for (int y = 0; y < 2; ++y) t3[x][y] = Mfma<float>::op(as[x], bs[y], t3[x][y]);
            }
        }
        if (more) lstore(buf ^ 1, kc + kG3K);
        __syncthreads();
        buf ^= 1;
    }
    z32* cs = C + (int64_t)(4 * lq) * ldc + li;
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rb = m0 + wm + x * 16 + r, cb = n0 + wn + y * 16;
                if (full || (rb + 4 * lq < M && cb + li < N)) {
                    const float re = t1[x][y][r] - t2[x][y][r];
                    const float im = t3[x][y][r] - t1[x][y][r] - t2[x][y][r];
                    cs[(int64_t)rb * ldc + cb] = ASSIGN ? z32{re, im} : z32{-re, -im};
                }
            }
}

// ---------------------------------------------------------------------------
// Outer trailing updates as bf16-split products.  Each FP32 operand value is
// split three ways, x = hi + mid + lo with every piece a bf16 (round to
// nearest even; the residuals x - hi and x - hi - mid are exact in FP32, so
// the three pieces carry all 24 significand bits), and a real product is the
// sum of the six piece products whose size is at least 2^-16 of the leading
// one (hi hi, hi mid, mid hi, mid mid, hi lo, lo hi; the three dropped ones
// are below 2^-24 relative, FP32's own rounding unit), each on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation (Henry, Tang and Heinecke,
// "Leveraging the bfloat16 artificial intelligence datatype for
// higher-precision computations", ARITH 2019).  With the 3M form's three real
// products the update runs 18 bf16 products per complex product: 36 MNK
// matrix-core flops where the FP32 3M kernel runs 6 MNK at one sixteenth the
// rate, 2.67 times the FP32 kernel's ceiling.
//
// The operands are split once per outer block into planes
// [stream * 3 + piece][k / 8][row][8] (streams re, im, re + im of the 3M form;
// rows of L21, columns of U12; rows padded by zeros to a multiple of 128), so
// the GEMM stages 16-byte runs of 8 k, the MFMA operand of one lane, with
// coalesced loads and no conversion work.
// ---------------------------------------------------------------------------
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

__device__ __forceinline__ void split3(float x, __bf16& h, __bf16& m, __bf16& l) {
    h = (__bf16)x;
    const float r1 = x - (float)h;
    m = (__bf16)r1;
    l = (__bf16)(r1 - (float)m);
}

// the nine planes of eight consecutive k of one row or column
__device__ __forceinline__ void split_store(const z32 (&v)[8], bf8v* __restrict__ out, int64_t plane, int64_t at) {
    bf8v o[9];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        __bf16 h, m, l;
        split3(v[j].x, h, m, l);
        o[0][j] = h;
        o[1][j] = m;
        o[2][j] = l;
        split3(v[j].y, h, m, l);
        o[3][j] = h;
        o[4][j] = m;
        o[5][j] = l;
        split3(v[j].x + v[j].y, h, m, l);
        o[6][j] = h;
        o[7][j] = m;
        o[8][j] = l;
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) out[q * plane + at] = o[q];
}

// L21 (M rows x K, row-major at A): thread (row m, k group kg), rows fastest.
__global__ __launch_bounds__(256) void k_split_rows(const z32* __restrict__ A, int64_t lda, int M, int Mpad, int K,
                                                    bf8v* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int KG = K / 8;
    if (e >= (int64_t)Mpad * KG) return;
    const int m = (int)(e % Mpad), kg = (int)(e / Mpad);
    z32 v[8];
    if (m < M) {
        const z32* p = A + (int64_t)m * lda + 8 * kg;
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = p[j];
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = z32{0.f, 0.f};
    }
    split_store(v, out, (int64_t)KG * Mpad, (int64_t)kg * Mpad + m);
}

// U12 (K x N, row-major at B): thread (column c, k group kg), columns fastest.
__global__ __launch_bounds__(256) void k_split_cols(const z32* __restrict__ B, int64_t ldb, int N, int Npad, int K,
                                                    bf8v* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int KG = K / 8;
    if (e >= (int64_t)Npad * KG) return;
    const int c = (int)(e % Npad), kg = (int)(e / Npad);
    z32 v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = c < N ? B[(int64_t)(8 * kg + j) * ldb + c] : z32{0.f, 0.f};
    split_store(v, out, (int64_t)KG * Npad, (int64_t)kg * Npad + c);
}

// C[M x N] -= A B from split planes SA (rows, Mpad) and SB (columns, Npad,
// starting at column boff, a multiple of 128); K a multiple of 32.  128 x 128
// tiles in the tile order of k_gemm3m, 4 waves of 64 x 64 (4 x 4 MFMA
// blocks; one wave per SIMD, its three 3M accumulator sets, 192 values, in
// the accumulation registers).  Each 32-wide k chunk runs as three phases,
// one per 3M stream, each the six piece products of the stream in the order
// mid mid, hi lo, lo hi, hi mid, mid hi, hi hi (small terms first).  The
// planes go straight from global memory into LDS (global_load_lds_dwordx4:
// no staging registers, no LDS store instructions).  LDS holds one buffer
// per stream, each the three pieces of both operands for one chunk as
// [piece][k / 8][row] in 16-byte units (the 16 lanes of a read pass take 256
// consecutive bytes, so reads are free of bank conflicts), 144 KB in all.
// The copy of stream s + 2 is in flight while stream s multiplies, two
// phases ahead; each wave waits for its own copies (vmcnt) before the
// barrier that publishes them.  The padded rows and columns are zero, so
// only the stores of C test the edges.  (An earlier form with 8 waves of
// 64 x 32, two per SIMD, and register-staged copies one phase ahead ran the
// updates of a 40960-unknown factorization at 6.02 ms per launch against
// 5.32.)
typedef __attribute__((address_space(3))) void lds_void_t;
struct BfBuf {
    bf8v a[3][4][kGM];
    bf8v b[3][4][kGN];
};
__device__ __forceinline__ void bf_copy(const bf8v* src, bf8v* lds_base) {
    __builtin_amdgcn_global_load_lds((const void*)src, (lds_void_t*)lds_base, 16, 0, 0);
}
// s_waitcnt immediates (vmcnt in bits 3:0 and 15:14, expcnt 6:4, lgkmcnt
// 11:8): vmcnt(12) lgkmcnt(0), and vmcnt(0) lgkmcnt(0)
constexpr int kWaitVm12 = 0x007C, kWaitVm0 = 0x0070;
__global__ __launch_bounds__(256, 1) void k_gemm3m_bf(const bf8v* __restrict__ SA, int Mpad,
                                                       const bf8v* __restrict__ SB, int Npad, int boff,
                                                       z32* __restrict__ C, int64_t ldc, int M, int N, int K,
                                                       int group) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    __shared__ BfBuf buf0, buf1, buf2;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = (wv >> 1) * 64, wn = (wv & 1) * 64;
    int tm, tn;
    g3_tile(gridDim.x, gridDim.y, group, tm, tn);
    const int m0 = tm * kGM, n0 = tn * kGN;
    const bool full = m0 + kGM <= M && n0 + kGN <= N;
    const int li = lane & 15, lq = lane >> 4;
    v4 t1[4][4], t2[4][4], t3[4][4];
    const z32* cl = C + (int64_t)(4 * lq) * ldc + li;
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            t2[x][y] = v4{0, 0, 0, 0};
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rb = m0 + wm + x * 16 + r, cb = n0 + wn + y * 16;
                z32 c{0.f, 0.f};
                if (full || (rb + 4 * lq < M && cb + li < N)) c = cl[(int64_t)rb * ldc + cb];
                t1[x][y][r] = -c.x;
                t3[x][y][r] = -c.x - c.y;
            }
        }
    // copy units of this wave: waves 0, 1 the A planes, waves 2, 3 the B
    // planes; unit u (0..11) of wave half h is piece p = (12 h + u) / 8,
    // k group kg = (12 h + u) % 8 / 2, rows 64 ((12 h + u) % 2) .. + 63
    const bool opA = wv < 2;
    const int h = wv & 1;
    const int64_t pitch = opA ? Mpad : Npad;
    const int64_t plane = (int64_t)(K / 8) * pitch;
    const bf8v* gbase = (opA ? SA + m0 : SB + boff + n0) + lane;
    auto copy = [&](int c, int st, BfBuf& B) {
        const bf8v* g = gbase + (int64_t)(st * 3) * plane + (int64_t)(4 * c) * pitch;
#pragma unroll
        for (int u = 0; u < 12; ++u) {
            const int v = 12 * h + u, p = v >> 3, kg = (v & 7) >> 1, half = v & 1;
            bf8v* dst = opA ? &B.a[p][kg][64 * half] : &B.b[p][kg][64 * half];
            bf_copy(g + p * plane + kg * pitch + 64 * half, dst);
        }
    };
    auto products = [&](const BfBuf& B, v4 (&acc)[4][4]) {
        bf8v a[3][4], b[3][4];
#pragma unroll
        for (int p = 0; p < 3; ++p) {
#pragma unroll
            for (int x = 0; x < 4; ++x) a[p][x] = B.a[p][lq][wm + 16 * x + li];
#pragma unroll
            for (int y = 0; y < 4; ++y) b[p][y] = B.b[p][lq][wn + 16 * y + li];
        }
#define BF_TERM(P, Q)                                                                                    \
    _Pragma("unroll") for (int x = 0; x < 4; ++x) _Pragma("unroll") for (int y = 0; y < 4; ++y) acc[x][y] = \
        __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[P][x], b[Q][y], acc[x][y], 0, 0, 0);
        BF_TERM(1, 1)
        BF_TERM(0, 2)
        BF_TERM(2, 0)
        BF_TERM(0, 1)
        BF_TERM(1, 0)
        BF_TERM(0, 0)
#undef BF_TERM
    };
    // publish: this wave's copies of the buffer about to be read are done
    // (at most the 12 copies issued after them remain), every wave's LDS
    // reads of the previous phase are done, then the barrier
    auto sync = [&](bool later_in_flight) {
        if (later_in_flight) __builtin_amdgcn_s_waitcnt(kWaitVm12);
        else __builtin_amdgcn_s_waitcnt(kWaitVm0);
        __asm__ volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        __asm__ volatile("" ::: "memory");
    };
    const int nc = K / 32;
    copy(0, 0, buf0);
    copy(0, 1, buf1);
    for (int c = 0; c < nc; ++c) {
        const bool more = c + 1 < nc;
        sync(true);
        copy(c, 2, buf2);
        products(buf0, t1);
        sync(true);
        if (more) copy(c + 1, 0, buf0);
        products(buf1, t2);
        sync(more);
        if (more) copy(c + 1, 1, buf1);
        products(buf2, t3);
    }
    z32* cs = C + (int64_t)(4 * lq) * ldc + li;
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rb = m0 + wm + x * 16 + r, cb = n0 + wn + y * 16;
                if (full || (rb + 4 * lq < M && cb + li < N)) {
                    const float re = t1[x][y][r] - t2[x][y][r];
                    const float im = t3[x][y][r] - t1[x][y][r] - t2[x][y][r];
                    cs[(int64_t)rb * ldc + cb] = z32{-re, -im};
                }
            }
}

// ---------------------------------------------------------------------------
// drivers
// ---------------------------------------------------------------------------
struct ProfHook;  // defined in the API unit; null when profiling is off
#define HIPCHECK_LU(x)                       \
    do {                                     \
        if ((x) != hipSuccess) return 3;     \
    } while (0)
void prof_begin(void* prof, hipStream_t s, int cls);
void prof_end(void* prof, hipStream_t s, int cls, double flops, double bytes);
enum { CLS_PANEL = 2, CLS_GEMM = 3, CLS_TRSM = 4, CLS_SWAP = 5, CLS_SOLVE = 6, CLS_GEMM64 = 12, CLS_SPLIT = 15,
       CLS_GEMMBF = 16 };


// Trailing-update GEMM: 3M products for complex FP32; four real products for
// FP64, whose three 3M accumulator sets would not fit two waves per SIMD.
// The 128-aligned block of C in the full-tile form and the right and bottom
// edge strips (fewer than 128 columns / rows) in the general form, each
// element of C written by exactly one launch: taken when C is aligned (one
// launch) or when the aligned block spans enough waves of the chip that the
// full form's saving outweighs the strips' two extra serial launches.
constexpr int64_t kSplitTiles = 2048;
static inline bool split_pays(int M, int N, int Ma, int Na) {
    if (Ma == 0 || Na == 0) return false;
    if (Ma == M && Na == N) return true;
    return (int64_t)(Ma / kGM) * (Na / kGN) >= kSplitTiles;
}

template <typename R, int ASSIGN>
static void launch_gemm(hipStream_t s, dim3 grid, const C2<R>* A, int64_t lda, const C2<R>* B, int64_t ldb, C2<R>* C,
                        int64_t ldc, int M, int N, int K, int group = kG3Group) {
    if constexpr (sizeof(R) == sizeof(float)) {
        const int Ma = M - M % kGM, Na = N - N % kGN;
        if (K % kG3K == 0 && split_pays(M, N, Ma, Na)) {
            hipLaunchKernelGGL((k_gemm3m<ASSIGN, true>), dim3(Na / kGN, Ma / kGM), dim3(512), 0, s, (const z32*)A,
                               lda, (const z32*)B, ldb, (z32*)C, ldc, Ma, Na, K, group);
            if (Na < N)
                hipLaunchKernelGGL((k_gemm3m<ASSIGN, false>), dim3(1, (M + kGM - 1) / kGM), dim3(512), 0, s,
                                   (const z32*)A, lda, (const z32*)(B + Na), ldb, (z32*)(C + Na), ldc, M, N - Na, K,
                                   group);
            if (Ma < M)
                hipLaunchKernelGGL((k_gemm3m<ASSIGN, false>), dim3(Na / kGN, 1), dim3(512), 0, s,
                                   (const z32*)(A + (int64_t)Ma * lda), lda, (const z32*)B, ldb,
                                   (z32*)(C + (int64_t)Ma * ldc), ldc, M - Ma, Na, K, group);
        } else {
            hipLaunchKernelGGL((k_gemm3m<ASSIGN, false>), grid, dim3(512), 0, s, (const z32*)A, lda, (const z32*)B,
                               ldb, (z32*)C, ldc, M, N, K, group);
        }
    } else {
        const int Ma = M - M % kGM, Na = N - N % kGN;
        if (K % kGK == 0 && split_pays(M, N, Ma, Na)) {
            hipLaunchKernelGGL((k_gemm_sub<R, ASSIGN, true>), dim3(Na / kGN, Ma / kGM), dim3(512), 0, s, A, lda, B,
                               ldb, C, ldc, Ma, Na, K);
            if (Na < N)
                hipLaunchKernelGGL((k_gemm_sub<R, ASSIGN, false>), dim3(1, (M + kGM - 1) / kGM), dim3(512), 0, s, A,
                                   lda, B + Na, ldb, C + Na, ldc, M, N - Na, K);
            if (Ma < M)
                hipLaunchKernelGGL((k_gemm_sub<R, ASSIGN, false>), dim3(Na / kGN, 1), dim3(512), 0, s,
                                   A + (int64_t)Ma * lda, lda, B, ldb, C + (int64_t)Ma * ldc, ldc, M - Ma, Na, K);
        } else {
            hipLaunchKernelGGL((k_gemm_sub<R, ASSIGN, false>), grid, dim3(512), 0, s, A, lda, B, ldb, C, ldc, M, N, K);
        }
    }
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
// Timed as lu_gemm (FP32, the 3M kernel) or lu_gemm_fp64.
template <typename R>
static void launch_update(hipStream_t s, C2<R>* A, int64_t lda, int r0, int c0, int ka, int M, int N, int K,
                          void* prof, int group = kG3Group) {
    if (M <= 0 || N <= 0 || K <= 0) return;
    const int cls = sizeof(R) == sizeof(float) ? CLS_GEMM : CLS_GEMM64;
    prof_begin(prof, s, cls);
    dim3 grid((N + kGN - 1) / kGN, (M + kGM - 1) / kGM);
    launch_gemm<R, 0>(s, grid, A + (int64_t)r0 * lda + ka, lda, A + (int64_t)ka * lda + c0, lda,
                      A + (int64_t)r0 * lda + c0, lda, M, N, K, group);
    prof_end(prof, s, cls, 8.0 * (double)M * N * K, (2.0 * (double)M * N + (double)(M + N) * K) * sizeof(C2<R>));
}

// The split planes of outer block [K0, K0 + K): L21 (rows r0 .. n) and U12
// (columns r0 .. n) of the square trailing matrix at r0 = K0 + K.
struct SplitPlanes {
    bf8v *rows = nullptr, *cols = nullptr;
    int M = 0, Mpad = 0, K = 0;
};
static SplitPlanes launch_split(hipStream_t s, z32* A, int64_t lda, int n, int K0, int K, void* buf, void* prof) {
    SplitPlanes P;
    P.M = n - K0 - K;
    P.Mpad = (P.M + 127) / 128 * 128;
    P.K = K;
    P.rows = (bf8v*)buf;
    P.cols = P.rows + (size_t)9 * (K / 8) * P.Mpad;
    const int r0 = K0 + K;
    const int64_t tot = (int64_t)P.Mpad * (K / 8);
    prof_begin(prof, s, CLS_SPLIT);
    hipLaunchKernelGGL(k_split_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                       (const z32*)(A + (int64_t)r0 * lda + K0), lda, P.M, P.Mpad, K, P.rows);
    hipLaunchKernelGGL(k_split_cols, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                       (const z32*)(A + (int64_t)K0 * lda + r0), lda, P.M, P.Mpad, K, P.cols);
    prof_end(prof, s, CLS_SPLIT, 0.0, 2.0 * (double)P.M * K * (sizeof(z32) + 9 * sizeof(__bf16)));
    return P;
}
// Columns [c0, c0 + N) of the trailing matrix (c0 relative to its first
// column, a multiple of 128) from the split planes: the bf16 GEMM, timed as
// lu_gemm_bf16 with the algorithmic 8 MNK flops of the complex update.
static void launch_update_split(hipStream_t s, z32* A, int64_t lda, int r0, const SplitPlanes& P, int c0, int N,
                                int group, void* prof) {
    if (P.M <= 0 || N <= 0) return;
    prof_begin(prof, s, CLS_GEMMBF);
    dim3 grid((N + kGN - 1) / kGN, P.Mpad / kGM);
    hipLaunchKernelGGL(k_gemm3m_bf, grid, dim3(256), 0, s, P.rows, P.Mpad, P.cols, P.Mpad, c0,
                       A + (int64_t)r0 * lda + r0 + c0, lda, P.M, N, P.K, group);
    prof_end(prof, s, CLS_GEMMBF, 8.0 * (double)P.M * N * P.K,
             (2.0 * (double)P.M * N * sizeof(z32) + 9.0 * (double)(P.M + N) * P.K * sizeof(__bf16)));
}

// Two-level right-looking LU: 64-wide panels inside 512-wide outer blocks, so
// the trailing update of each outer block is one K = 512 MFMA GEMM (deep
// enough that reading and writing C once per update costs little against the
// products).
constexpr int kOuter = 512;

// The panels of outer block [K0, Kend): tournament panel, pivot-block
// factors, L21, and the updates confined to the block's own columns.
template <typename R>
static int panel_phase(hipStream_t s, C2<R>* A, int64_t lda, int n, int K0, int Kend, int* ipiv, C2<R>* inv,
                       PanelWS& ws, void* prof, int group) {
    for (int k0 = K0; k0 < Kend; k0 += kNB) {
        const int w = (Kend - k0) < kNB ? (Kend - k0) : kNB;
        const int m = n - k0;
        const int G = (m + kTsRows - 1) / kTsRows;
        if (G > ws.max_blocks) return 2;
        prof_begin(prof, s, CLS_PANEL);
        hipLaunchKernelGGL((k_tslu<R>), dim3(G), dim3(kTsThreads), 0, s, A, lda, n, k0, w, ws);
        hipLaunchKernelGGL((k_tslu_apply<R>), dim3(1), dim3(256), 0, s, A, lda, n, k0, w, ws, ipiv, inv);
        if (m > w) {
            // L21 = A21 inv(U11), in place
            const C2<R>* Ui = inv + (size_t)(2 * (k0 / kNB) + 1) * kNB * kNB;
            dim3 grid(1, (m - w + kGM - 1) / kGM);
            launch_gemm<R, 1>(s, grid, A + (int64_t)(k0 + w) * lda + k0, lda, Ui, (int64_t)kNB,
                              A + (int64_t)(k0 + w) * lda + k0, lda, m - w, w, w);
        }
        prof_end(prof, s, CLS_PANEL, 8.0 * (double)m * w * w / 2.0, 2.0 * (double)m * w * sizeof(C2<R>));
        // this panel's interchanges inside the outer block
        launch_laswp<R>(s, A, lda, k0, w, ipiv, K0, k0, k0 + w, Kend, prof);
        if (k0 + w < Kend) {
            launch_trsm<R>(s, A, lda, k0, w, k0 + w, Kend, inv, prof);
            launch_update<R>(s, A, lda, k0 + w, k0 + w, k0, n - k0 - w, Kend - k0 - w, w, prof, group);
        }
    }
    return 0;
}

// Two-level right-looking LU: 64-wide panels inside outer blocks of
// tune.outer columns (512 by default), so the trailing update of each outer
// block is one K = tune.outer MFMA GEMM.  With a side stream
// (LuStreams::side, tune.ahead) the panels of outer block j + 1 run there,
// overlapped with block j's trailing update: the update of block j + 1's
// columns goes first (event `ready`), block j + 1's panel phase follows on the
// side stream (event `panel`), and the main stream updates the remaining
// columns meanwhile, then waits for `panel` before applying block j + 1's
// interchanges elsewhere.  Columns touched concurrently are disjoint.  With
// tune.split (complex FP32) the outer trailing updates run from bf16 split
// planes of L21 and U12 (k_gemm3m_bf), written once per outer block into
// `split`.  The launch sequence depends only on n and the tune, never on the
// data, so it can be captured once and replayed.
template <typename R>
int lu_factor(const LuStreams& st, C2<R>* A, int64_t lda, int n, int* ipiv, C2<R>* inv, PanelWS& ws, void* prof,
              const LuTune& tune, void* split) {
    hipStream_t s = st.main;
    const bool ahead = st.side != nullptr && tune.ahead;
    const int outer = tune.outer == 256 ? 256 : kOuter;
    const int group = tune.group;
    const bool use_split = sizeof(R) == sizeof(float) && tune.split && split != nullptr;
    int rc = panel_phase<R>(s, A, lda, n, 0, std::min(n, outer), ipiv, inv, ws, prof, group);
    if (rc) return rc;
    for (int K0 = 0; K0 < n; K0 += outer) {
        const int KW = (n - K0) < outer ? (n - K0) : outer;
        const int Kend = K0 + KW;
        if (K0 > 0) {
            if (ahead) HIPCHECK_LU(hipStreamWaitEvent(s, st.panel, 0));
            else if ((rc = panel_phase<R>(s, A, lda, n, K0, Kend, ipiv, inv, ws, prof, group))) return rc;
        }
        // the outer block's interchanges on every other column
        launch_laswp<R>(s, A, lda, K0, KW, ipiv, 0, K0, Kend, n, prof);
        if (Kend >= n) break;
        // U12 = inv(L11) A12 by 64-row blocks, then the K = KW trailing update
        for (int i0 = K0; i0 < Kend; i0 += kNB) {
            const int wi = (Kend - i0) < kNB ? (Kend - i0) : kNB;
            launch_trsm<R>(s, A, lda, i0, wi, Kend, n, inv, prof);
            launch_update<R>(s, A, lda, i0 + wi, Kend, i0, Kend - i0 - wi, n - Kend, wi, prof, group);
        }
        const int Kend2 = std::min(n, Kend + outer);
        // update of columns [Kend + c0, Kend + c0 + N)
        SplitPlanes P;
        if constexpr (sizeof(R) == sizeof(float)) {
            if (use_split) P = launch_split(s, (z32*)A, lda, n, K0, KW, split, prof);
        }
        auto update = [&](int c0, int N) {
            if constexpr (sizeof(R) == sizeof(float)) {
                if (use_split) {
                    launch_update_split(s, (z32*)A, lda, Kend, P, c0, N, group, prof);
                    return;
                }
            }
            launch_update<R>(s, A, lda, Kend, Kend + c0, K0, n - Kend, N, KW, prof, group);
        };
        if (ahead) {
            update(0, Kend2 - Kend);
            HIPCHECK_LU(hipEventRecord(st.ready, s));
            HIPCHECK_LU(hipStreamWaitEvent(st.side, st.ready, 0));
            if ((rc = panel_phase<R>(st.side, A, lda, n, Kend, Kend2, ipiv, inv, ws, prof, group))) return rc;
            HIPCHECK_LU(hipEventRecord(st.panel, st.side));
            update(Kend2 - Kend, n - Kend2);
        } else {
            update(0, n - Kend);
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Triangular solves with the stored diagonal-block inverses: one launch per
// solve, forward and backward phases together.
//
// Workgroup t owns one 64-row block: t < nblk is forward block t, t >= nblk
// is backward block 2 nblk - 1 - t.  The index comes from a ticket counter
// taken on entry, so a workgroup only ever waits on workgroups that started
// before it, and the lowest unfinished one always has its inputs: progress
// does not depend on the dispatch order.  A block depends on every earlier
// block of its phase, so the blocks of a phase finish in order and one word
// per phase, (epoch << 32) | blocks finished, tells a waiter which of its
// inputs are ready.  The epoch counts the solves run on the workspace, so
// words left by earlier solves read as no progress and nothing is reset
// between solves.  A wait lasts at most `limit` wall-clock ticks; the first
// to expire raises the abort word, which releases every other waiter, so
// every workgroup exits and the host reports the failure.  The last
// workgroup to finish zeroes the ticket counters for the next launch.
//
// Solution blocks and progress words move through relaxed agent-scope
// atomic loads and stores (coherent across XCDs); a producer's solution
// stores have completed (s_waitcnt) before its wave stores the progress word.
// ---------------------------------------------------------------------------
struct TriSync {
    unsigned* ctr;             // [0] ticket, [1] finished workgroups
    unsigned long long* prog;  // [0] forward, [1] backward progress
    int* abort;                // nonzero once a wait has expired
    unsigned epoch;
    long long limit;           // wall-clock ticks one wait may last
};

// Tile geometry: a 64 x 64 tile of complex R read with 16-byte loads.
template <typename R>
struct TriGeom {
    static constexpr int V = 16 / (int)sizeof(C2<R>);  // complex values per load
    static constexpr int TPR = kNB / V;                // threads per tile row
    static constexpr int RPP = 256 / TPR;              // tile rows per pass
    static constexpr int P = kNB / RPP;                // passes per tile
};
template <typename R>
struct alignas(16) Vec16 {
    C2<R> e[TriGeom<R>::V];
};

// A 64 x 64 tile held in registers: thread tid holds the memory elements
// (mr + RPP p, V qg + v), mr = tid / TPR, qg = tid % TPR.  TR = 0 multiplies
// by the tile (partial sums per tile row), TR = 1 by its conjugate transpose
// (partial sums per tile column).  Products accumulate in FP64.
template <typename R, int TR>
struct TriTile {
    using G = TriGeom<R>;
    static constexpr int NX = TR ? G::P : G::V;  // vector entries a thread multiplies
    static constexpr int NA = TR ? G::V : G::P;  // partial sums a thread holds
    C2<R> t[G::P][G::V];

    __device__ static int xidx(int tid, int k) {
        return TR ? tid / G::TPR + G::RPP * k : (tid % G::TPR) * G::V + k;
    }
    // rows x cols: valid extent of the tile in memory (zeros outside).
    __device__ void load(const C2<R>* base, int64_t ld, int rows, int cols, bool vec, int tid) {
        const int mr = tid / G::TPR, q0 = (tid % G::TPR) * G::V;
        if (vec && rows == kNB && cols == kNB) {
#pragma unroll
            for (int p = 0; p < G::P; ++p) {
                const Vec16<R> w = *(const Vec16<R>*)(base + (int64_t)(mr + G::RPP * p) * ld + q0);
#pragma unroll
                for (int v = 0; v < G::V; ++v) t[p][v] = w.e[v];
            }
        } else {
#pragma unroll
            for (int p = 0; p < G::P; ++p)
#pragma unroll
                for (int v = 0; v < G::V; ++v) {
                    const int m = mr + G::RPP * p, q = q0 + v;
                    t[p][v] = (m < rows && q < cols) ? base[(int64_t)m * ld + q] : C2<R>{0, 0};
                }
        }
    }
    __device__ void fma(const z64 (&xv)[NX], z64 (&acc)[NA]) const {
#pragma unroll
        for (int p = 0; p < G::P; ++p)
#pragma unroll
            for (int v = 0; v < G::V; ++v) {
                const double tr = t[p][v].x, ti = t[p][v].y;
                if (TR == 0) {
                    acc[p].x += tr * xv[v].x - ti * xv[v].y;
                    acc[p].y += tr * xv[v].y + ti * xv[v].x;
                } else {
                    acc[v].x += tr * xv[p].x + ti * xv[p].y;
                    acc[v].y += tr * xv[p].y - ti * xv[p].x;
                }
            }
    }
};

// Sums the partial sums of all threads into out[0..64); ends with a barrier.
template <typename R, int TR, int NA>
__device__ __forceinline__ void tri_reduce(z64 (&acc)[NA], z64* out, z64* part, int tid) {
    using G = TriGeom<R>;
    if (TR == 0) {
#pragma unroll
        for (int p = 0; p < G::P; ++p)
#pragma unroll
            for (int off = G::TPR / 2; off >= 1; off >>= 1) {
                acc[p].x += __shfl_xor(acc[p].x, off, 64);
                acc[p].y += __shfl_xor(acc[p].y, off, 64);
            }
        if (tid % G::TPR == 0)
#pragma unroll
            for (int p = 0; p < G::P; ++p) out[tid / G::TPR + G::RPP * p] = acc[p];
        __syncthreads();
    } else {
        const int mr = tid / G::TPR, q0 = (tid % G::TPR) * G::V;
#pragma unroll
        for (int v = 0; v < G::V; ++v) part[mr * kNB + q0 + v] = acc[v];
        __syncthreads();
        if (tid < kNB) {
            z64 a{0, 0};
#pragma unroll
            for (int r = 0; r < G::RPP; ++r) {
                a.x += part[r * kNB + tid].x;
                a.y += part[r * kNB + tid].y;
            }
            out[tid] = a;
        }
        __syncthreads();
    }
}

// Wait until the phase word shows at least `need` finished blocks of this
// epoch; returns the count seen, or -1 on abort or expiry (wave-uniform).
__device__ __forceinline__ int tri_wait(const unsigned long long* w, unsigned epoch, int need, int* abortw,
                                     long long limit) {
    long long t0 = 0;
    for (int it = 0;; ++it) {
        const unsigned long long v = ld_ag(w);
        const int cnt = (unsigned)(v >> 32) == epoch ? (int)(unsigned)(v & 0xffffffffu) : 0;
        if (cnt >= need) return cnt;
        if (ld_ag(abortw) != 0) return -1;
        const long long now = wall_clock64();
        if (it == 0) {
            t0 = now;
        } else if (now - t0 > limit) {
            st_ag(abortw, 1);
            return -1;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

// ADJ = 0: x = inv(A) b for P A = L U (y = inv(L) P b, x = inv(U) y).
// ADJ = 1: x = inv(A)^H b = P^T inv(L^H) inv(U^H) b.
// Forward solutions go to `work`; backward ones to x (ADJ = 0) or to `work`
// in place and x[perm[r]] (ADJ = 1).  b, work and x must not overlap.
template <typename R, int ADJ>
__global__ __launch_bounds__(256) void k_trsolve(const C2<R>* __restrict__ LU, int64_t lda, int n, int nblk,
                                                 const int* __restrict__ perm, const C2<R>* __restrict__ inv,
                                                 const C2<R>* __restrict__ b, C2<R>* work, C2<R>* x, TriSync sy,
                                                 int vec) {
    using G = TriGeom<R>;
    using Tile = TriTile<R, ADJ>;
    __shared__ z64 s_acc[kNB];
    __shared__ z64 s_y[kNB];
    __shared__ z64 s_part[ADJ ? G::RPP * kNB : 1];
    __shared__ unsigned s_t;
    const int tid = threadIdx.x;
    if (tid == 0) s_t = __hip_atomic_fetch_add(&sy.ctr[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const unsigned t = s_t;
    if (t >= gridDim.x) {
        if (tid == 0) st_ag(sy.abort, 2);
    } else {
        const bool fwd = (int)t < nblk;
        const int blk = fwd ? (int)t : 2 * nblk - 1 - (int)t;
        const int r0 = blk * kNB;
        const int bs = (n - r0) < kNB ? (n - r0) : kNB;
        const int cnt = fwd ? blk : nblk - 1 - blk;
        const C2<R>* xs = (ADJ || fwd) ? work : x;
        const unsigned long long* pw = &sy.prog[fwd ? 0 : 1];
        Tile D;
        D.load(inv + (size_t)(2 * blk + (fwd != (ADJ != 0) ? 0 : 1)) * kNB * kNB, kNB, bs, bs, true, tid);
        z64 acc[Tile::NA];
#pragma unroll
        for (int k = 0; k < Tile::NA; ++k) acc[k] = z64{0, 0};
        int ready = 0;
        bool ab = false;
        for (int jj = 0; jj < cnt; ++jj) {
            const int j = fwd ? jj : nblk - 1 - jj;
            const int c0 = j * kNB;
            const int cs = (n - c0) < kNB ? (n - c0) : kNB;
            Tile T;
            if (ADJ)
                T.load(LU + (int64_t)c0 * lda + r0, lda, cs, bs, vec != 0, tid);
            else
                T.load(LU + (int64_t)r0 * lda + c0, lda, bs, cs, vec != 0, tid);
            if (jj >= ready) {
                ready = tri_wait(pw, sy.epoch, jj + 1, sy.abort, sy.limit);
                if (ready < 0) {
                    ab = true;
                    break;
                }
            }
            z64 xv[Tile::NX];
#pragma unroll
            for (int k = 0; k < Tile::NX; ++k) {
                const int c = Tile::xidx(tid, k);
                xv[k] = z64{0, 0};
                if (c < cs) {
                    const C2<R> v = ld_ag(xs + c0 + c);
                    xv[k] = z64{(double)v.x, (double)v.y};
                }
            }
            T.fma(xv, acc);
        }
        // the backward phase starts from this block's forward solution
        if (!ab && !fwd && tri_wait(&sy.prog[0], sy.epoch, blk + 1, sy.abort, sy.limit) < 0) ab = true;
        if (__syncthreads_or(ab)) {
            if (tid < bs) {
                const C2<R> nan{(R)NAN, (R)NAN};
                if (fwd)
                    st_ag(work + r0 + tid, nan);
                else if (ADJ)
                    x[perm[r0 + tid]] = nan;
                else
                    st_ag(x + r0 + tid, nan);
            }
        } else {
            tri_reduce<R, ADJ>(acc, s_acc, s_part, tid);
            if (tid < kNB) {
                z64 y{0, 0};
                if (tid < bs) {
                    C2<R> rv;
                    if (fwd)
                        rv = ADJ ? b[r0 + tid] : b[perm[r0 + tid]];
                    else
                        rv = ld_ag(work + r0 + tid);
                    y = z64{(double)rv.x - s_acc[tid].x, (double)rv.y - s_acc[tid].y};
                }
                s_y[tid] = y;
            }
            __syncthreads();
            z64 yv[Tile::NX];
#pragma unroll
            for (int k = 0; k < Tile::NX; ++k) yv[k] = s_y[Tile::xidx(tid, k)];
#pragma unroll
            for (int k = 0; k < Tile::NA; ++k) acc[k] = z64{0, 0};
            D.fma(yv, acc);
            tri_reduce<R, ADJ>(acc, s_acc, s_part, tid);
            if (tid < kNB) {
                if (tid < bs) {
                    const C2<R> z{(R)s_acc[tid].x, (R)s_acc[tid].y};
                    if (ADJ) {
                        st_ag(work + r0 + tid, z);
                        if (!fwd) x[perm[r0 + tid]] = z;
                    } else {
                        st_ag((fwd ? work : x) + r0 + tid, z);
                    }
                }
                wait_mem();
                if (tid == 0)
                    st_ag(&sy.prog[fwd ? 0 : 1],
                          ((unsigned long long)sy.epoch << 32) | (unsigned)(fwd ? blk + 1 : nblk - blk));
            }
        }
    }
    __syncthreads();
    if (tid == 0) {
        const unsigned d = __hip_atomic_fetch_add(&sy.ctr[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (d == gridDim.x - 1) {
            st_ag(&sy.ctr[0], 0u);
            st_ag(&sy.ctr[1], 0u);
        }
    }
}

template <typename R>
static void tri_launch(hipStream_t s, const C2<R>* LU, int64_t lda, int n, const int* perm, const C2<R>* inv,
                       const C2<R>* b, C2<R>* work, C2<R>* x, PanelWS& ws, int adj) {
    if (n <= 0) return;
    const int nblk = (n + kNB - 1) / kNB;
    if (++ws.tri_epoch == 0) ws.tri_epoch = 1;
    const TriSync sy{ws.tri_ctr, ws.tri_prog, ws.status + 2, ws.tri_epoch, ws.spin_ticks};
    const int vec = (sizeof(C2<R>) == 16 || (lda % 2 == 0 && ((uintptr_t)LU & 15) == 0)) ? 1 : 0;
    if (adj)
        hipLaunchKernelGGL((k_trsolve<R, 1>), dim3(2 * nblk), dim3(256), 0, s, LU, lda, n, nblk, perm, inv, b, work,
                           x, sy, vec);
    else
        hipLaunchKernelGGL((k_trsolve<R, 0>), dim3(2 * nblk), dim3(256), 0, s, LU, lda, n, nblk, perm, inv, b, work,
                           x, sy, vec);
}

// x = inv(A)^H b for P A = L U.
template <typename R>
void lu_solve_h(hipStream_t s, const C2<R>* LU, int64_t lda, int n, const int* perm, const C2<R>* inv,
                const C2<R>* b, C2<R>* work, C2<R>* x, PanelWS& ws, void* prof) {
    prof_begin(prof, s, CLS_SOLVE);
    tri_launch<R>(s, LU, lda, n, perm, inv, b, work, x, ws, 1);
    prof_end(prof, s, CLS_SOLVE, 8.0 * (double)n * n, (double)n * n * sizeof(C2<R>));
}

template <typename R>
void lu_solve(hipStream_t s, const C2<R>* LU, int64_t lda, int n, const int* perm, const C2<R>* inv,
              const C2<R>* b, C2<R>* work, C2<R>* x, PanelWS& ws, void* prof) {
    prof_begin(prof, s, CLS_SOLVE);
    tri_launch<R>(s, LU, lda, n, perm, inv, b, work, x, ws, 0);
    prof_end(prof, s, CLS_SOLVE, 8.0 * (double)n * n, (double)n * n * sizeof(C2<R>));
}

// ---------------------------------------------------------------------------
// Dense products and many-right-hand-side solves (low-frequency Calderon
// blocks of the Mueller system).
// ---------------------------------------------------------------------------
template <typename R>
void gemm_assign(hipStream_t s, const C2<R>* A, int64_t lda, const C2<R>* B, int64_t ldb, C2<R>* C, int64_t ldc,
                 int M, int N, int K) {
    if (M <= 0 || N <= 0 || K <= 0) return;
    dim3 grid((N + kGN - 1) / kGN, (M + kGM - 1) / kGM);
    launch_gemm<R, 1>(s, grid, A, lda, B, ldb, C, ldc, M, N, K);
}

// Row interchanges of ipiv, then block forward substitution with L and
// backward substitution with U: per 64-row block, one in-place product with
// the stored inverse of the diagonal block (M <= 64, so C may alias B) and
// one trailing product for the rows still to be solved.
template <typename R>
void lu_solve_many(hipStream_t s, const C2<R>* LU, int64_t lda, int n, const int* ipiv, const C2<R>* inv, C2<R>* X,
                   int64_t ldx, int ncols, void* prof) {
    if (n <= 0 || ncols <= 0) return;
    prof_begin(prof, s, CLS_SOLVE);
    for (int k0 = 0; k0 < n; k0 += kNB) launch_laswp<R>(s, X, ldx, k0, std::min(kNB, n - k0), ipiv, 0, ncols, 0, 0, nullptr);
    const int nblk = (n + kNB - 1) / kNB;
    dim3 g1((ncols + kGN - 1) / kGN, 1);
    for (int kb = 0; kb < nblk; ++kb) {
        const int r0 = kb * kNB, bs = std::min(kNB, n - r0);
        launch_gemm<R, 1>(s, g1, inv + (size_t)(2 * kb) * kNB * kNB, (int64_t)kNB, X + (int64_t)r0 * ldx, ldx,
                          X + (int64_t)r0 * ldx, ldx, bs, ncols, bs);
        const int rest = n - r0 - bs;
        if (rest > 0) {
            dim3 g((ncols + kGN - 1) / kGN, (rest + kGM - 1) / kGM);
            launch_gemm<R, 0>(s, g, LU + (int64_t)(r0 + bs) * lda + r0, lda, X + (int64_t)r0 * ldx, ldx,
                              X + (int64_t)(r0 + bs) * ldx, ldx, rest, ncols, bs);
        }
    }
    for (int kb = nblk - 1; kb >= 0; --kb) {
        const int r0 = kb * kNB, bs = std::min(kNB, n - r0);
        launch_gemm<R, 1>(s, g1, inv + (size_t)(2 * kb + 1) * kNB * kNB, (int64_t)kNB, X + (int64_t)r0 * ldx, ldx,
                          X + (int64_t)r0 * ldx, ldx, bs, ncols, bs);
        if (r0 > 0) {
            dim3 g((ncols + kGN - 1) / kGN, (r0 + kGM - 1) / kGM);
            launch_gemm<R, 0>(s, g, LU + r0, lda, X + (int64_t)r0 * ldx, ldx, X, ldx, r0, ncols, bs);
        }
    }
    prof_end(prof, s, CLS_SOLVE, 8.0 * (double)n * n * ncols, (double)n * n * sizeof(C2<R>));
}

__global__ void k_add_diag(int n, z64* X, int64_t ldx, double v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) X[(int64_t)i * ldx + i].x += v;
}
void launch_add_diag(hipStream_t s, int n, z64* X, int64_t ldx, double v) {
    hipLaunchKernelGGL(k_add_diag, dim3((n + 255) / 256), dim3(256), 0, s, n, X, ldx, v);
}

template <typename R>
__global__ void k_calderon_bottom(int nt, const z64* __restrict__ Me, const z64* __restrict__ Mi, double ig,
                                  C2<R>* __restrict__ A) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nt * nt) return;
    const int i = (int)(e / nt), j = (int)(e % nt);
    const int64_t ld = 2 * (int64_t)nt, src = (int64_t)i * ld;
    const z64 ke = Me[src + j], he = Me[src + nt + j], ki = Mi[src + j], hi = Mi[src + nt + j];
    A[(int64_t)(nt + i) * ld + j] = C2<R>{R(-he.x + hi.x * ig), R(-he.y + hi.y * ig)};
    A[(int64_t)(nt + i) * ld + nt + j] = C2<R>{R((i == j ? 1.0 : 0.0) + ke.x - ki.x), R(ke.y - ki.y)};
}
template <typename R>
void launch_calderon_bottom(hipStream_t s, int nt, const z64* Me, const z64* Mi, double g, C2<R>* A) {
    const int64_t tot = (int64_t)nt * nt;
    hipLaunchKernelGGL((k_calderon_bottom<R>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, nt, Me, Mi,
                       1.0 / g, A);
}

// Synthetic matrix for timing factorizations: entries in [-1/2, 1/2) from a
// hash of their index, the diagonal raised by 2 so no pivot vanishes.
template <typename R>
__global__ void k_fill_test(C2<R>* A, int n) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    uint64_t h = (uint64_t)e * 0x9E3779B97F4A7C15ull;
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 29;
    const double re = (double)(h & 0xFFFFFF) / 16777216.0 - 0.5, im = (double)((h >> 24) & 0xFFFFFF) / 16777216.0 - 0.5;
    const bool diag = e / n == e % n;
    A[e] = C2<R>{(R)(re + (diag ? 2.0 : 0.0)), (R)im};
}
template <typename R>
void launch_fill_test(hipStream_t s, C2<R>* A, int n) {
    const int64_t tot = (int64_t)n * n;
    hipLaunchKernelGGL((k_fill_test<R>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, A, n);
}

// explicit instantiations
template void launch_fill_test<double>(hipStream_t, z64*, int);
template void launch_fill_test<float>(hipStream_t, z32*, int);
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
template void gemm_assign<double>(hipStream_t, const z64*, int64_t, const z64*, int64_t, z64*, int64_t, int, int, int);
template void lu_solve_many<double>(hipStream_t, const z64*, int64_t, int, const int*, const z64*, z64*, int64_t, int, void*);
template void launch_calderon_bottom<double>(hipStream_t, int, const z64*, const z64*, double, z64*);
template void launch_calderon_bottom<float>(hipStream_t, int, const z64*, const z64*, double, z32*);
template int lu_factor<double>(const LuStreams&, z64*, int64_t, int, int*, z64*, PanelWS&, void*, const LuTune&,
                               void*);
template int lu_factor<float>(const LuStreams&, z32*, int64_t, int, int*, z32*, PanelWS&, void*, const LuTune&,
                              void*);
template void lu_solve<double>(hipStream_t, const z64*, int64_t, int, const int*, const z64*, const z64*, z64*, z64*, PanelWS&, void*);
template void lu_solve<float>(hipStream_t, const z32*, int64_t, int, const int*, const z32*, const z32*, z32*, z32*, PanelWS&, void*);
template void lu_solve_h<double>(hipStream_t, const z64*, int64_t, int, const int*, const z64*, const z64*, z64*, z64*, PanelWS&, void*);
template void lu_solve_h<float>(hipStream_t, const z32*, int64_t, int, const int*, const z32*, const z32*, z32*, z32*, PanelWS&, void*);

}  // namespace bem
