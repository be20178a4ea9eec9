// C ABI of the MI355X BEM library: contexts, meshes, operators (assembled +
// factored systems resident in HBM), solutions, refinement and sweeps.
//
// Reference behaviour mirrored here (paths relative to the reference repo):
//   solve_full_bem drivers           src/solvers/bem/full.jl:52-66,130-184,273-495
//   equilibrated LU + unscaling      src/solvers/bem/full.jl:364-440
//   MixedRefinement state machine    src/solvers/bem/fluid_system.jl:172-309
//   residual report                  src/numerics/diagnostics.jl:9-18
//   far field / target strength      src/solvers/bem/full.jl:509-531
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <functional>
#include <list>
#include <map>
#include <set>
#include <unordered_map>
#include <thread>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bemhip.h"
#include "bem_device.h"
#include "bem_hsolve.h"
#include "bem_internal.h"

using namespace bem;

namespace {

struct BemError : std::runtime_error {
    int code;
    BemError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHECK(expr)                                                                        \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess)                                                                 \
            throw BemError(BEM_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// "h_apply" is reported as the sum of its two kernels' classes, "h_proj" and
// "h_rows" (the event timers do not nest).
const char* kClassNames[] = {"assemble", "near",      "lu_panel", "lu_gemm", "lu_trsm", "lu_swap",
                             "solve",    "residual",  "far_field", "h_apply", "h_aca",  "h_ortho",
                             "lu_gemm_fp64", "h_proj", "h_rows", "lu_split", "lu_gemm_bf16", "lu_graph",
                             "nys_assemble", "nys_defect", "nys_fit"};
constexpr int kNumClasses = 21;
enum { C_ASSEMBLE = 0, C_NEAR = 1, C_RESIDUAL = 7, C_FAR = 8, C_HAPPLY = 9, C_HACA = 10, C_HORTHO = 11,
       C_HPROJ = 13, C_HROWS = 14, C_LUGRAPH = 17, C_NYS_ASM = 18, C_NYS_DEF = 19, C_NYS_FIT = 20 };

struct ProfRec {
    int cls;
    hipEvent_t a, b;
    double flops, bytes;
};

}  // namespace

namespace bem {
struct ProfLane;
}

namespace {

// Per-class totals of the context's kernel timing.  Each lane records into
// its own ProfLane (records and event pool under the lane's own lock), so
// lanes launching concurrently never wait on each other here; records are
// folded into these totals once their events have completed.
struct Prof {
    std::mutex mu;  // totals and the lane list
    std::atomic<bool> on{false};
    std::vector<bem::ProfLane*> lanes;
    int64_t launches[kNumClasses] = {};
    double ms[kNumClasses] = {}, flops[kNumClasses] = {}, bytes[kNumClasses] = {};
    void drain();                   // every lane's records, waiting for their events
    void attach(bem::ProfLane* pl);
    void detach(bem::ProfLane* pl);  // drains pl first
};

}  // namespace

namespace bem {
// Event timing of one launch (or launch group) on a lane's streams.
struct ProfLane {
    Prof* prof = nullptr;
    hipEvent_t cur = nullptr;  // begin event of the open record
    std::mutex mu;             // the lane's thread and statistics readers
    std::deque<ProfRec> recs;  // in issue order
    std::vector<hipEvent_t> pool;
    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        return hipEventCreate(&e) == hipSuccess ? e : nullptr;
    }
    // Fold records into the totals, oldest first: all of them (wait), or
    // those whose end event has completed.  Without waiting, nothing is
    // queried while an error is pending on this thread, and the not-ready
    // status a query leaves is cleared, so the caller's own launch check
    // still sees only its launches' errors.
    void harvest(bool wait) {
        int64_t n[kNumClasses] = {};
        double t[kNumClasses] = {}, f[kNumClasses] = {}, b[kNumClasses] = {};
        bool any = false;
        if (!wait && hipPeekAtLastError() != hipSuccess) return;
        while (!recs.empty()) {
            const ProfRec r = recs.front();
            if (wait) {
                (void)hipEventSynchronize(r.b);
            } else if (hipEventQuery(r.b) != hipSuccess) {
                (void)hipGetLastError();
                break;
            }
            float ms = 0;
            (void)hipEventElapsedTime(&ms, r.a, r.b);
            n[r.cls] += 1;
            t[r.cls] += ms;
            f[r.cls] += r.flops;
            b[r.cls] += r.bytes;
            pool.push_back(r.a);
            pool.push_back(r.b);
            recs.pop_front();
            any = true;
        }
        if (!any) return;
        std::lock_guard<std::mutex> g(prof->mu);
        for (int c = 0; c < kNumClasses; ++c) {
            prof->launches[c] += n[c];
            prof->ms[c] += t[c];
            prof->flops[c] += f[c];
            prof->bytes[c] += b[c];
        }
    }
    ~ProfLane() {
        for (auto& r : recs) {
            hipEventDestroy(r.a);
            hipEventDestroy(r.b);
        }
        for (auto e : pool) hipEventDestroy(e);
        if (cur) hipEventDestroy(cur);
    }
};
void prof_begin(void* prof, hipStream_t s, int cls) {
    ProfLane* pl = static_cast<ProfLane*>(prof);
    if (!pl || !pl->prof || !pl->prof->on.load()) return;
    (void)cls;
    std::lock_guard<std::mutex> g(pl->mu);
    if (!pl->cur) pl->cur = pl->get();
    if (pl->cur) hipEventRecord(pl->cur, s);
}
void prof_end(void* prof, hipStream_t s, int cls, double flops, double bytes) {
    ProfLane* pl = static_cast<ProfLane*>(prof);
    if (!pl || !pl->prof || !pl->prof->on.load()) return;
    std::lock_guard<std::mutex> g(pl->mu);
    if (!pl->cur) return;
    hipEvent_t e = pl->get();
    if (!e) return;
    hipEventRecord(e, s);
    pl->recs.push_back({cls, pl->cur, e, flops, bytes});
    pl->cur = nullptr;
    if (pl->recs.size() >= 64) pl->harvest(pl->recs.size() > 65536);
}
}  // namespace bem

namespace {
void Prof::drain() {
    std::vector<bem::ProfLane*> ls;
    {
        std::lock_guard<std::mutex> g(mu);
        ls = lanes;
    }
    for (auto* pl : ls) {
        std::lock_guard<std::mutex> g(pl->mu);
        pl->harvest(true);
    }
}
void Prof::attach(bem::ProfLane* pl) {
    pl->prof = this;
    std::lock_guard<std::mutex> g(mu);
    lanes.push_back(pl);
}
void Prof::detach(bem::ProfLane* pl) {
    {
        std::lock_guard<std::mutex> g(pl->mu);
        pl->harvest(true);
    }
    std::lock_guard<std::mutex> g(mu);
    lanes.erase(std::remove(lanes.begin(), lanes.end(), pl), lanes.end());
}
}  // namespace

// One execution lane: a HIP stream, the panel workspace of the LU on that
// stream, and (for batched sweeps) a bump-allocated device arena that holds a
// whole solve, so concurrent lanes never call hipMalloc/hipFree mid-sweep.
// A factorization captured as a HIP graph: replayed when a lane factors the
// same buffers with the same shape again (a sweep's arena hands each item of
// a size class the same addresses).
struct LuGraph {
    const void *lu, *ipiv, *inv, *split;
    int n, single;
    LuTune t;
    hipGraphExec_t exec;
};
constexpr size_t kLaneGraphs = 6;

struct Lane {
    hipStream_t s = nullptr;
    hipStream_t s2 = nullptr;          // LU panel look-ahead
    hipEvent_t ev_ready = nullptr, ev_panel = nullptr;
    std::vector<LuGraph> graphs;       // oldest first, at most kLaneGraphs
    PanelWS ws{};
    void* ws_mem = nullptr;
    void* rec_mem = nullptr;           // ws.rec
    char* arena = nullptr;
    size_t cap = 0, off = 0;
    ProfLane pl;
};

struct bem_ctx_s {
    int device = 0;
    int cu_count = 0;
    uint64_t budget = 0, used = 0;
    std::string err;
    std::recursive_mutex mu;
    std::mutex amu;  // allocator bookkeeping
    Prof prof;
    std::vector<Lane*> lanes;  // lanes[0] serves the single-solve entry points
    int max_lanes = 8;         // concurrent solves of bem_sweep
    int max_heavy = 1;         // of which heavy ones (kHeavyUnknowns)
    void* scratch = nullptr;   // this device's address of the fixed scratch array (bem_far_field_noalloc)
    // LU shapes: 0 fixed (FP32 3M updates), 1 fixed with bf16-split outer
    // updates, 2 tuned per size class (bem_ctx_autotune, or by bem_sweep
    // for the classes it meets)
    int lu_mode = 2;
    std::mutex tmu;
    std::map<std::pair<int, int>, bem_lu_tune> tunes;  // (unknowns, single) -> tuner's report

    Lane* lane0() { return lanes[0]; }
    // Device memory: a caching allocator.  Released blocks go to a free pool
    // instead of hipFree (which waits on the whole device and would serialize
    // concurrent lanes); an allocation reuses the smallest pooled block of at
    // least its size and at most a quarter larger.  `used` counts live bytes
    // against the budget; pooled bytes are freed when the budget needs them.
    // Callers release a block only once the work using it has completed.
    static size_t granule(size_t bytes) {
        const size_t g = bytes >= ((size_t)1 << 20) ? ((size_t)2 << 20) : 256;
        return (bytes + g - 1) / g * g;
    }
    void* alloc(size_t bytes) {
        if (bytes == 0) bytes = 16;
        bytes = granule(bytes);
        std::lock_guard<std::mutex> g(amu);
        auto it = pool.lower_bound(bytes);
        if (it != pool.end() && it->first <= bytes + bytes / 4) {
            void* p = it->second;
            const size_t b = it->first;
            pool.erase(it);
            pooled -= b;
            used += b;
            sizes[p] = b;
            return p;
        }
        if (used + bytes > budget)
            throw BemError(BEM_ERR_ADMISSION, "request of " + std::to_string(bytes) + " bytes exceeds the HBM budget (" +
                                                  std::to_string(used) + " of " + std::to_string(budget) + " in use)");
        while (used + pooled + bytes > budget && !pool.empty()) trim_one();
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            while (!pool.empty()) trim_one();
            e = hipMalloc(&p, bytes);
        }
        if (e != hipSuccess) throw BemError(BEM_ERR_ADMISSION, std::string("hipMalloc failed: ") + hipGetErrorString(e));
        used += bytes;
        sizes[p] = bytes;
        return p;
    }
    void release(void* p) {
        if (!p) return;
        std::lock_guard<std::mutex> g(amu);
        auto it = sizes.find(p);
        if (it == sizes.end()) return;
        used -= it->second;
        pool.emplace(it->second, p);
        pooled += it->second;
        sizes.erase(it);
    }
    // largest pooled block back to the device (amu held)
    void trim_one() {
        auto it = std::prev(pool.end());
        hipFree(it->second);
        pooled -= it->first;
        pool.erase(it);
    }
    void trim_all() {
        std::lock_guard<std::mutex> g(amu);
        while (!pool.empty()) trim_one();
    }
    std::map<void*, size_t> sizes;
    std::multimap<size_t, void*> pool;
    uint64_t pooled = 0;
};

struct bem_mesh_s {
    bem_ctx ctx = nullptr;
    HostMesh h;
    DevMesh d;
    double* sources = nullptr;  // Nystrom meshes: 3 x ns point sources of the interpolation basis
    std::vector<void*> bufs;
    std::atomic<int> refs{1};
};

struct bem_operator_s {
    bem_ctx ctx = nullptr;
    Lane* lane = nullptr;
    bool use_arena = false;
    size_t arena_mark = 0;
    std::vector<void*> heap;      // allocations outside the lane arena
    bem_mesh mesh = nullptr;
    bem_problem prob{};
    int n = 0, nt = 0;
    bool single = false;          // FP32 storage / factor
    void* A = nullptr;            // original system, row-major, R storage
    void* B = nullptr;            // RHS operator (rigid/soft)
    void* LU = nullptr;           // factor in R
    void* LU64 = nullptr;         // FP64 factor after promotion
    int* ipiv = nullptr;
    int* perm = nullptr;
    void* inv = nullptr;
    void* split = nullptr;        // bf16 planes of the factorization's outer updates
    int* ipiv64 = nullptr;
    bool calderon = false;        // low-frequency Mueller: K', H from the Calderon identities
    int* perm64 = nullptr;
    void* inv64 = nullptr;
    double* rows = nullptr;       // equilibration (null when off)
    double* cols = nullptr;
    double matrix_norm = 0;
    int precision = 1;            // 1 single factor active, 0 double
    int reason = BEM_REASON_NONE;
    double rcond = NAN;
    double assembly_ms = 0, factor_ms = 0;
    // scratch vectors
    z64 *vb = nullptr, *vx = nullptr, *vr = nullptr, *vt = nullptr;
    double* vden = nullptr;
    void *wv = nullptr, *wv2 = nullptr, *wv3 = nullptr;  // R vectors
    double* red = nullptr;
    int* flag = nullptr;
};

struct bem_solution_s {
    bem_ctx ctx = nullptr;
    bem_mesh mesh = nullptr;
    bem_problem prob{};  // boundary, contrasts and wavenumber of the solve
    double k = 0;
    int nrhs = 0, nt = 0;
    std::vector<double> angles;
    z64 *p = nullptr, *q = nullptr;  // device, nt x nrhs
    std::vector<std::complex<double>> hp, hq;
    std::vector<bem_diagnostics> diag;
};

namespace {

// A lane: its own stream and tournament-pivoting workspace (and no arena).
Lane* make_lane(bem_ctx c) {
    auto* L = new Lane();
    c->prof.attach(&L->pl);
    HIPCHECK(hipStreamCreateWithFlags(&L->s, hipStreamNonBlocking));
    HIPCHECK(hipStreamCreateWithFlags(&L->s2, hipStreamNonBlocking));
    HIPCHECK(hipEventCreateWithFlags(&L->ev_ready, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&L->ev_panel, hipEventDisableTiming));
    const int G = 2048;  // 524288 rows (256 per leaf)
    const int nodes = 2 * G + 16;
    L->ws.max_blocks = G;
    L->ws.max_nodes = nodes;
    const size_t bytes = 64 + sizeof(int) * nodes * (2 + 64) + sizeof(int) * 80 + 10 * 256;
    char* m = (char*)c->alloc(bytes);
    L->ws_mem = m;
    HIPCHECK(hipMemset(m, 0, bytes));
    size_t off = 0;
    auto take = [&](size_t b) { void* p = m + off; off += (b + 255) & ~(size_t)255; return p; };
    L->ws.status = (int*)take(64);
    L->ws.tickets = (int*)take(sizeof(int) * nodes);
    L->ws.node_cnt = (int*)take(sizeof(int) * nodes);
    L->ws.node_rows = (int*)take(sizeof(int) * nodes * 64);
    L->ws.root = (int*)take(sizeof(int) * 80);
    L->ws.tri_ctr = (unsigned*)take(2 * sizeof(unsigned));
    L->ws.tri_prog = (unsigned long long*)take(2 * sizeof(unsigned long long));
    if (off > bytes) throw BemError(BEM_ERR_DEVICE, "workspace layout");
    L->rec_mem = c->alloc(kPanelRecBytes);
    L->ws.rec = L->rec_mem;
    // a device-side wait of the triangular solves gives up after 10 s
    int khz = 0;
    HIPCHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
    L->ws.spin_ticks = (long long)(khz > 0 ? khz : 100000) * 10000;
    HIPCHECK(hipDeviceSynchronize());
    return L;
}

void free_lane(bem_ctx c, Lane* L) {
    if (!L) return;
    if (L->s) hipStreamSynchronize(L->s);
    if (L->s2) hipStreamSynchronize(L->s2);
    for (LuGraph& g : L->graphs) hipGraphExecDestroy(g.exec);
    L->graphs.clear();
    c->prof.detach(&L->pl);
    c->release(L->ws_mem);
    c->release(L->rec_mem);
    c->release(L->arena);
    if (L->s) hipStreamDestroy(L->s);
    if (L->s2) hipStreamDestroy(L->s2);
    if (L->ev_ready) hipEventDestroy(L->ev_ready);
    if (L->ev_panel) hipEventDestroy(L->ev_panel);
    delete L;
}

// After a triangular solve whose device-side wait expired (ws.status[2]):
// clear the flag and the ticket counters, then fail the request.
void tri_check(Lane* L, int expired) {
    if (!expired) return;
    launch_fill_int(L->s, L->ws.status + 2, 1, 0);
    launch_fill_int(L->s, (int*)L->ws.tri_ctr, 2, 0);
    (void)hipStreamSynchronize(L->s);
    throw BemError(BEM_ERR_DEVICE, "triangular solve: a device-side wait expired");
}

template <typename F>
int guarded(bem_ctx ctx, F&& f) {
    try {
        f();
        if (ctx) ctx->err.clear();
        return BEM_OK;
    } catch (const BemError& e) {
        if (ctx) ctx->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        if (ctx) ctx->err = e.what();
        return BEM_ERR_DEVICE;
    }
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void mesh_release(bem_mesh m) {
    if (!m) return;
    if (--m->refs > 0) return;
    for (void* p : m->bufs) m->ctx->release(p);
    delete m;
}

void validate_problem(const bem_problem& p) {
    if (!(std::isfinite(p.k) && p.k > 0)) throw BemError(BEM_ERR_ARGUMENT, "k must be finite and positive");
    if (p.boundary == BEM_FLUID) {
        if (p.formulation != BEM_MULLER && p.formulation != BEM_CBIE)
            throw BemError(BEM_ERR_ARGUMENT, "fluid formulation must be :muller or :cbie");
        const double g = p.density_contrast, h = p.soundspeed_contrast;
        if (!(std::isfinite(g) && g > 0 && std::isfinite(h) && h > 0))
            throw BemError(BEM_ERR_ARGUMENT, "fluid density and soundspeed contrasts must be finite and positive");
    } else if (p.boundary == BEM_RIGID || p.boundary == BEM_PRESSURE_RELEASE) {
        if (p.formulation != BEM_BURTON_MILLER && p.formulation != BEM_CBIE)
            throw BemError(BEM_ERR_ARGUMENT, "formulation must be :cbie or :burton_miller");
    } else {
        throw BemError(BEM_ERR_ARGUMENT, "unknown boundary condition");
    }
    if (p.precision != BEM_DOUBLE && p.precision != BEM_MIXED)
        throw BemError(BEM_ERR_ARGUMENT, "precision must be :double or :mixed");
    if (p.precision == BEM_MIXED) {
        if (!p.equilibrate) throw BemError(BEM_ERR_ARGUMENT, "mixed precision requires equilibrate=true");
        if (!(std::isfinite(p.refine_tolerance) && p.refine_tolerance > 0 && p.refine_tolerance < 1))
            throw BemError(BEM_ERR_ARGUMENT, "invalid refinement tolerance");
        if (p.refine_maxiter < 0) throw BemError(BEM_ERR_ARGUMENT, "maxiter must be a nonnegative integer");
    }
}

void direction(double beta, double alpha, double* d) {
    d[0] = std::cos(beta);
    d[1] = std::sin(beta) * std::cos(alpha);
    d[2] = std::sin(beta) * std::sin(alpha);
}

// Device memory for an operator: from its lane's arena when the lane has one
// (batched sweeps), otherwise from the context heap.
void* op_alloc(bem_operator op, size_t bytes) {
    Lane* L = op->lane;
    if (L && L->arena && op->use_arena) {
        const size_t a = (L->off + 255) & ~(size_t)255;
        if (a + bytes <= L->cap) {
            L->off = a + bytes;
            return L->arena + a;
        }
    }
    void* p = op->ctx->alloc(bytes);
    op->heap.push_back(p);
    return p;
}

// The shape of an n-unknown factorization (bem_ctx_set_lu_tuning): the
// tuner's choice for a tuned size class; otherwise the default shape, with
// bf16-split outer updates for complex FP32 from kSplitMin unknowns.
constexpr int kSplitMin = 4096;
LuTune tune_of(const bem_lu_tune& r) {
    LuTune t;
    t.outer = r.outer;
    t.group = r.group;
    t.split = r.split;
    t.ahead = r.ahead;
    t.graph = r.graph;
    return t;
}
LuTune lu_tune_for(bem_ctx c, int n, bool single) {
    std::lock_guard<std::mutex> g(c->tmu);
    if (c->lu_mode == 2) {
        auto it = c->tunes.find({n, single ? 1 : 0});
        if (it != c->tunes.end()) return tune_of(it->second);
    }
    LuTune t;
    t.split = single && (c->lu_mode == 1 || (c->lu_mode == 2 && n >= kSplitMin));
    return t;
}

// Launches the factorization of LUbuf on the operator's lane with shape t:
// directly, or (t.graph and `graphs`) by replaying the lane's captured graph
// for these buffers and shape, capturing it first if the lane has none.
template <typename R>
void lu_launch(bem_operator op, void* LUbuf, int* ipiv, void* inv, const LuTune& t, bool graphs) {
    Lane* L = op->lane;
    hipStream_t s = L->s;
    const int n = op->n;
    void* split = nullptr;
    if (sizeof(R) == sizeof(float) && t.split) {
        if (!op->split) op->split = op_alloc(op, lu_split_bytes(n));
        split = op->split;
    }
    LuStreams st;
    st.main = s;
    st.side = L->s2;
    st.ready = L->ev_ready;
    st.panel = L->ev_panel;
    auto fail = [&](int rc) {
        if (rc == 2)
            throw BemError(BEM_ERR_ADMISSION, "system larger than the panel workspace (" +
                                                  std::to_string(256 * (int64_t)L->ws.max_blocks) + " rows)");
        throw BemError(BEM_ERR_DEVICE, "LU panel launch failed");
    };
    if (!(t.graph && graphs)) {
        const int rc = lu_factor<R>(st, (C2<R>*)LUbuf, n, n, ipiv, (C2<R>*)inv, L->ws, &L->pl, t, split);
        if (rc) fail(rc);
        return;
    }
    const int single = sizeof(R) == sizeof(float);
    auto same = [&](const LuGraph& g) {
        return g.lu == LUbuf && g.ipiv == ipiv && g.inv == inv && g.split == split && g.n == n &&
               g.single == single && g.t.outer == t.outer && g.t.group == t.group && g.t.split == t.split &&
               g.t.ahead == t.ahead;
    };
    auto it = std::find_if(L->graphs.begin(), L->graphs.end(), same);
    if (it == L->graphs.end()) {
        HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
        const int rc = lu_factor<R>(st, (C2<R>*)LUbuf, n, n, ipiv, (C2<R>*)inv, L->ws, nullptr, t, split);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &graph);
        if (rc || e != hipSuccess || !graph) {
            if (graph) hipGraphDestroy(graph);
            if (rc) fail(rc);
            throw BemError(BEM_ERR_DEVICE, std::string("LU graph capture failed: ") + hipGetErrorString(e));
        }
        hipGraphExec_t exec = nullptr;
        const hipError_t e2 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e2 != hipSuccess)
            throw BemError(BEM_ERR_DEVICE, std::string("LU graph instantiation failed: ") + hipGetErrorString(e2));
        if (L->graphs.size() >= kLaneGraphs) {
            hipGraphExecDestroy(L->graphs.front().exec);
            L->graphs.erase(L->graphs.begin());
        }
        L->graphs.push_back(LuGraph{LUbuf, ipiv, inv, split, n, single, t, exec});
        it = L->graphs.end() - 1;
    }
    const double n3 = (double)n * n * n;
    prof_begin(&L->pl, s, C_LUGRAPH);
    HIPCHECK(hipGraphLaunch(it->exec, s));
    prof_end(&L->pl, s, C_LUGRAPH, 8.0 / 3.0 * n3, 2.0 * (double)n * n * sizeof(C2<R>));
}

// Drops the lane's captured factorizations of buffers about to be released.
void drop_graphs(Lane* L, const void* lu) {
    for (auto it = L->graphs.begin(); it != L->graphs.end();) {
        if (it->lu == lu) {
            hipGraphExecDestroy(it->exec);
            it = L->graphs.erase(it);
        } else {
            ++it;
        }
    }
}

template <typename R>
void factor_into(bem_operator op, void* LUbuf, int* ipiv, int* perm, void* inv, const LuTune* force = nullptr) {
    hipStream_t s = op->lane->s;
    const int n = op->n;
    const LuTune t = force ? *force : lu_tune_for(op->ctx, n, sizeof(R) == sizeof(float));
    launch_fill_int(s, op->lane->ws.status, 1, INT_MAX);
    launch_fill_int(s, op->lane->ws.status + 1, 1, 0);
    // graphs only where the buffers repeat: a sweep's arena, or when forced
    lu_launch<R>(op, LUbuf, ipiv, inv, t, op->use_arena || force != nullptr);
    HIPCHECK(hipGetLastError());
    int status[2];
    std::vector<int> hp(n);
    HIPCHECK(hipMemcpyAsync(status, op->lane->ws.status, sizeof status, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(hp.data(), ipiv, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    if (status[1] != 0) throw BemError(BEM_ERR_DEVICE, "LU panel pivot selection was inconsistent");
    if (status[0] != INT_MAX) throw BemError(BEM_ERR_NUMERIC, "singular:" + std::to_string(status[0]));
    std::vector<int> pm(n);
    for (int i = 0; i < n; ++i) pm[i] = i;
    for (int i = 0; i < n; ++i) {
        int p = hp[i];
        if (p < i || p >= n) throw BemError(BEM_ERR_DEVICE, "LU produced an invalid pivot");
        std::swap(pm[i], pm[p]);
    }
    HIPCHECK(hipMemcpyAsync(perm, pm.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    HIPCHECK(hipStreamSynchronize(s));
}

// The tuner: factorizations of a synthetic n x n matrix on lane 0, timed
// with HIP events, by coordinate descent from the default shape: bf16-split
// updates (complex FP32), outer width 256, L2 group 4 and 16 (complex FP32),
// no look-ahead, graph replay (up to 16384 unknowns).  A knob moves when the
// time drops by more than 1.5 %.  Each shape is timed 3 times (once from
// 16384 unknowns, where a factorization takes long enough to be steady)
// after one untimed run.
template <typename R>
bem_lu_tune tune_lu(bem_ctx c, int n) {
    constexpr bool single = sizeof(R) == sizeof(float);
    Lane* L = c->lane0();
    hipStream_t s = L->s;
    const size_t mat = sizeof(C2<R>) * (size_t)n * n;
    bem_operator_s op;
    op.ctx = c;
    op.lane = L;
    op.n = n;
    op.single = single;
    std::vector<void*> bufs;
    auto al = [&](size_t b) { void* p = c->alloc(b); bufs.push_back(p); return p; };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bem_lu_tune rep{};
    try {
        void* A0 = al(mat);
        op.LU = al(mat);
        op.ipiv = (int*)al(sizeof(int) * n);
        op.inv = al(sizeof(C2<R>) * 2 * 4096 * (size_t)((n + 63) / 64));
        launch_fill_test<R>(s, (C2<R>*)A0, n);
        HIPCHECK(hipEventCreate(&e0));
        HIPCHECK(hipEventCreate(&e1));
        auto once = [&](const LuTune& t) {
            HIPCHECK(hipMemcpyAsync(op.LU, A0, mat, hipMemcpyDeviceToDevice, s));
            launch_fill_int(s, L->ws.status, 1, INT_MAX);
            launch_fill_int(s, L->ws.status + 1, 1, 0);
            HIPCHECK(hipEventRecord(e0, s));
            lu_launch<R>(&op, op.LU, op.ipiv, op.inv, t, true);
            HIPCHECK(hipEventRecord(e1, s));
            int status[2];
            HIPCHECK(hipMemcpyAsync(status, L->ws.status, sizeof status, hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            if (status[0] != INT_MAX || status[1] != 0) throw BemError(BEM_ERR_DEVICE, "tuning factorization failed");
            float ms = 0;
            HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
            return (double)ms;
        };
        const int reps = n >= 16384 ? 1 : 3;
        auto measure = [&](const LuTune& t) {
            once(t);
            double best = 1e300;
            for (int r = 0; r < reps; ++r) best = std::min(best, once(t));
            return best;
        };
        LuTune best;
        double best_ms = measure(best);
        rep.default_ms = best_ms;
        int cand = 1;
        auto consider = [&](LuTune t) {
            const double ms = measure(t);
            ++cand;
            if (ms < 0.985 * best_ms) {
                best = t;
                best_ms = ms;
            }
        };
        if (single) {
            LuTune t = best;
            t.split = 1;
            consider(t);
        }
        {
            LuTune t = best;
            t.outer = 256;
            consider(t);
        }
        if (single) {
            const LuTune base = best;
            for (int gsz : {4, 16}) {
                LuTune t = base;
                t.group = gsz;
                consider(t);
            }
        }
        {
            LuTune t = best;
            t.ahead = 0;
            consider(t);
        }
        if (n <= 16384) {
            LuTune t = best;
            t.graph = 1;
            consider(t);
        }
        rep.unknowns = n;
        rep.precision = single ? BEM_MIXED : BEM_DOUBLE;
        rep.outer = best.outer;
        rep.group = best.group;
        rep.split = best.split;
        rep.ahead = best.ahead;
        rep.graph = best.graph;
        rep.tuned = 1;
        rep.candidates = cand;
        rep.best_ms = best_ms;
    } catch (...) {
        hipStreamSynchronize(s);
        drop_graphs(L, op.LU);
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
        for (void* p : bufs) c->release(p);
        for (void* p : op.heap) c->release(p);
        throw;
    }
    drop_graphs(L, op.LU);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    for (void* p : bufs) c->release(p);
    for (void* p : op.heap) c->release(p);
    std::lock_guard<std::mutex> g(c->tmu);
    c->tunes[{n, single ? 1 : 0}] = rep;
    return rep;
}

// ---------------------------------------------------------------------------
// curved-element Nystrom operators (density interpolation)
// ---------------------------------------------------------------------------
// Regular-wave interpolation (src/solvers/bem/full.jl:316-317,
// quadrature/fluid.jl:9): Mueller bodies whose larger wavenumber satisfies
// k * radius <= 1 and 2 k * rms <= 1 over the quadrature nodes; every other
// system uses the point-source basis.
bool nys_regular(const HostMesh& h, const bem_problem& p) {
    if (p.boundary != BEM_FLUID || p.formulation != BEM_MULLER) return false;
    const double kmax = std::max(p.k, p.k / p.soundspeed_contrast);
    return kmax * h.qradius <= 1.0 && 2.0 * kmax * h.qrms <= 1.0;
}

NysBasis nys_basis(const bem_mesh m, double k, bool regular) {
    NysBasis b;
    b.k = k;
    if (regular) {
        b.kind = 1;
        b.degree = (int)std::floor(std::sqrt(2.0 * m->h.nq)) - 1;  // quadrature/fluid.jl:112-113
        b.P = (b.degree + 1) * (b.degree + 1);
        for (int c = 0; c < 3; ++c) b.center[c] = m->h.qcenter[c];
        b.radius = m->h.qradius;
    } else {
        b.kind = 0;
        b.P = (int)(m->h.sources.size() / 3);
        b.zs = m->sources;
    }
    return b;
}

// Device bytes of the correction workspace of an assembly (both media, both
// operator families).
size_t nys_bytes(const bem_mesh m) {
    if (!m->h.nq) return 0;
    const size_t nt = (size_t)m->d.nt;
    const int P = std::max((int)(m->h.sources.size() / 3), 25);
    return sizeof(z64) * (4 * (size_t)P * nt + nys_defect_parts((int)nt, P) + 8 * nt * m->h.nq) + 8 * 256;
}

// Workspace of the density-interpolation corrections of one mesh: basis
// traces, defects of both operator families and their per-slice partial sums.
struct NysWork {
    int P = 0;
    z64 *g0 = nullptr, *g1 = nullptr, *th_sd = nullptr, *th_kh = nullptr, *part = nullptr;
};
template <typename A>
NysWork nys_work(A&& alloc, const bem_mesh m, bool regular, bool derivative) {
    NysWork w;
    const size_t nt = (size_t)m->d.nt;
    w.P = nys_basis(m, 1.0, regular).P;
    w.g0 = (z64*)alloc(sizeof(z64) * w.P * nt);
    w.g1 = (z64*)alloc(sizeof(z64) * w.P * nt);
    w.th_sd = (z64*)alloc(sizeof(z64) * w.P * nt);
    if (derivative) w.th_kh = (z64*)alloc(sizeof(z64) * w.P * nt);
    w.part = (z64*)alloc(sizeof(z64) * nys_defect_parts((int)nt, w.P));
    return w;
}
// The corrections of S and D (and of K' and H when `derivative`) at
// wavenumber k: corr_sd / corr_kh in the launch_nys_weights layout.
void nys_medium(hipStream_t s, ProfLane* pl, const bem_mesh m, const NysWork& w, double k, bool regular,
                bool derivative, z64* corr_sd, z64* corr_kh) {
    const DevMesh& d = m->d;
    const double pairs = (double)d.nt * d.nt;
    const NysBasis b = nys_basis(m, k, regular);
    prof_begin(pl, s, C_NYS_FIT);
    launch_nys_basis(s, d, b, w.g0, w.g1);
    prof_end(pl, s, C_NYS_FIT, 0.0, 0.0);
    // algorithmic flops of the defects: two complex multiply-adds (8 flops
    // each) per pair, basis function and operator family
    prof_begin(pl, s, C_NYS_DEF);
    launch_nys_defect(s, d, b.k, w.P, derivative ? 3 : 1, w.g0, w.g1, w.part, w.th_sd, w.th_kh);
    prof_end(pl, s, C_NYS_DEF, pairs * (derivative ? 2.0 : 1.0) * 16.0 * w.P, 0.0);
    prof_begin(pl, s, C_NYS_FIT);
    launch_nys_weights(s, d, w.P, w.g0, w.g1, w.th_sd, derivative ? w.th_kh : nullptr, corr_sd, corr_kh);
    prof_end(pl, s, C_NYS_FIT, 0.0, 0.0);
}

// FP64 operations of one point-kernel evaluation (kpoint, bem_kernels.h),
// an FMA counted as two and sincos and rsqrt as one each: separation and
// R^2 (8), 1/R, R, kR and sincos (4), the Green's function (6), D and K'
// (16 each: projection, factor, complex product into the sum), H (20) and
// n_x.n_y (5) for S, D, K' and H (Burton-Miller, Mueller); 34 for S and D.
constexpr double kNysEvalFlops4 = 75.0, kNysEvalFlops2 = 34.0;

// Naive Nystrom system of kind ap.kind into o0/o1, then the corrections of
// S and D (and of K' and H when `derivative`) of each medium added into the
// element-diagonal blocks.
template <typename R>
void nys_assemble(bem_operator op, const AsmParams& ap, C2<R>* o0, C2<R>* o1, bool derivative) {
    hipStream_t s = op->lane->s;
    const bem_mesh m = op->mesh;
    const DevMesh& d = m->d;
    const int nt = d.nt;
    const bool two = ap.kind == ASM_MULLER || ap.kind == ASM_FLUID_CBIE;
    const bool regular = nys_regular(m->h, op->prob);
    // flops: kNysEvalFlops per pair and medium (the point kernels of the
    // system's operators); bytes: the entries written
    const double pairs = (double)nt * nt;
    const int media = two ? 2 : 1;
    const bool all4 = ap.kind == ASM_BM || ap.kind == ASM_MULLER;
    prof_begin(&op->lane->pl, s, C_NYS_ASM);
    launch_nys_assemble<R>(s, d, ap, o0, o1);
    prof_end(&op->lane->pl, s, C_NYS_ASM, pairs * media * (all4 ? kNysEvalFlops4 : kNysEvalFlops2),
             pairs * sizeof(C2<R>) * (two ? 4.0 : 2.0));
    const NysWork w = nys_work([&](size_t b) { return op_alloc(op, b); }, m, regular, derivative);
    z64* corr[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    for (int med = 0; med < media; ++med) {
        corr[med][0] = (z64*)op_alloc(op, sizeof(z64) * 2 * (size_t)nt * d.nq);
        if (derivative) corr[med][1] = (z64*)op_alloc(op, sizeof(z64) * 2 * (size_t)nt * d.nq);
        nys_medium(s, &op->lane->pl, m, w, med == 0 ? ap.k : ap.kint, regular, derivative, corr[med][0],
                   corr[med][1]);
    }
    prof_begin(&op->lane->pl, s, C_NYS_FIT);
    launch_nys_apply<R>(s, d, ap, corr[0][0], corr[0][1], corr[1][0], corr[1][1], o0, o1);
    prof_end(&op->lane->pl, s, C_NYS_FIT, 0.0, 0.0);
    HIPCHECK(hipGetLastError());
}

void require_panels(const bem_mesh m, const char* what) {
    if (m->h.nq) throw BemError(BEM_ERR_ARGUMENT, std::string(what) + " requires a flat-panel mesh");
}

// Near-field pressure of a Nystrom solution (src/fields/pressure.jl:404-472):
// regions by winding number, then per region (exterior at k, interior at
// k_int) and per basis the layer potentials with the nearest element excluded
// and recovered by density interpolation (maxdist = 2 radius,
// pressure.jl:441-462).  Points with k |x - c| <= 2 use the regular-wave basis
// when k is in its range (pressure.jl:446-450).
void nys_pressure(bem_solution sol, int r, int64_t npts, const double* points, int field, double* out) {
    bem_ctx ctx = sol->ctx;
    const bem_mesh m = sol->mesh;
    const HostMesh& h = m->h;
    const DevMesh& d = m->d;
    hipStream_t s = ctx->lane0()->s;
    FieldParams fp{};
    fp.k = sol->k;
    fp.fluid = sol->prob.boundary == BEM_FLUID ? 1 : 0;
    fp.g = fp.fluid ? sol->prob.density_contrast : 1.0;
    fp.kint = fp.fluid ? sol->k / sol->prob.soundspeed_contrast : sol->k;
    fp.field = field;
    direction(sol->angles[2 * r], sol->angles[2 * r + 1], fp.d);
    const z64* P = sol->p + (size_t)r * sol->nt;
    const z64* Q = sol->q + (size_t)r * sol->nt;
    std::vector<void*> held;
    auto alloc = [&](size_t bytes) {
        void* p = ctx->alloc(bytes);
        held.push_back(p);
        return p;
    };
    try {
        const int64_t chunk = std::min<int64_t>(npts, 1 << 18);
        double* dpts = (double*)alloc(sizeof(double) * 3 * chunk);
        z64* dout = (z64*)alloc(sizeof(z64) * chunk);
        int* dreg = (int*)alloc(sizeof(int) * chunk);
        double* dwind = (double*)alloc(sizeof(double) * chunk);
        int* didx = (int*)alloc(sizeof(int) * chunk);
        // basis traces and element pseudo-inverses per (region, basis), made once
        struct Prep { z64 *g0 = nullptr, *g1 = nullptr, *pinv = nullptr; NysBasis b; };
        Prep prep[2][2];
        auto prepared = [&](int inside, int regular) -> Prep& {
            Prep& pr = prep[inside][regular];
            if (!pr.g0) {
                pr.b = nys_basis(m, inside ? fp.kint : fp.k, regular != 0);
                const size_t nb = (size_t)pr.b.P * d.nt;
                pr.g0 = (z64*)alloc(sizeof(z64) * nb);
                pr.g1 = (z64*)alloc(sizeof(z64) * nb);
                pr.pinv = (z64*)alloc(sizeof(z64) * (size_t)(d.nt / d.nq) * 2 * d.nq * pr.b.P);
                launch_nys_basis(s, d, pr.b, pr.g0, pr.g1);
                launch_nys_pinv(s, d, pr.b.P, pr.g0, pr.g1, pr.pinv);
            }
            return pr;
        };
        auto in_range = [&](double k) { return k * h.qradius <= 1.0 && 2.0 * k * h.qrms <= 1.0; };
        std::vector<int> reg(chunk), idx;
        for (int64_t i0 = 0; i0 < npts; i0 += chunk) {
            const int nc = (int)std::min<int64_t>(chunk, npts - i0);
            HIPCHECK(hipMemcpyAsync(dpts, points + 3 * i0, sizeof(double) * 3 * nc, hipMemcpyHostToDevice, s));
            launch_nys_winding(s, d, fp, nc, dpts, dreg, dwind);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipMemcpyAsync(reg.data(), dreg, sizeof(int) * nc, hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            for (int i = 0; i < nc; ++i)
                if (reg[i] == REGION_NONE)
                    throw BemError(BEM_ERR_ARGUMENT, "the selected field has no acoustic pressure at this point");
            for (int inside = 0; inside < 2; ++inside)
                for (int regular = 0; regular < 2; ++regular) {
                    const double k = inside ? fp.kint : fp.k;
                    const bool ok = in_range(k);
                    idx.clear();
                    for (int i = 0; i < nc; ++i) {
                        if (reg[i] != (inside ? REGION_INTERIOR : REGION_EXTERIOR)) continue;
                        const double* x = points + 3 * (i0 + i);
                        const double dx = x[0] - h.qcenter[0], dy = x[1] - h.qcenter[1], dz = x[2] - h.qcenter[2];
                        const bool near = ok && k * std::sqrt(dx * dx + dy * dy + dz * dz) <= 2.0;
                        if (near == (regular != 0)) idx.push_back(i);
                    }
                    if (idx.empty()) continue;
                    Prep& pr = prepared(inside, regular);
                    NysPointEval ev{};
                    ev.inside = inside;
                    ev.k = k;
                    ev.maxdist = 2.0 * h.qradius;
                    ev.basis = pr.b;
                    ev.g0 = pr.g0;
                    ev.g1 = pr.g1;
                    ev.pinv = pr.pinv;
                    ev.g = fp.g;
                    ev.k0 = fp.k;
                    for (int c = 0; c < 3; ++c) ev.d[c] = fp.d[c];
                    ev.add_incident = !inside && field == BEM_FIELD_TOTAL;
                    HIPCHECK(hipMemcpyAsync(didx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice, s));
                    launch_nys_points(s, d, ev, (int)idx.size(), didx, dpts, P, Q, dout);
                    HIPCHECK(hipGetLastError());
                    HIPCHECK(hipStreamSynchronize(s));
                }
            HIPCHECK(hipMemcpyAsync(out + 2 * i0, dout, sizeof(z64) * nc, hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
        }
    } catch (...) {
        hipStreamSynchronize(s);
        for (void* p : held) ctx->release(p);
        throw;
    }
    for (void* p : held) ctx->release(p);
}

// Low-frequency Mueller systems take their derivative blocks from the
// Calderon identities S K' = D S and S H = D D - I/4 (full.jl:187-207,
// 319-339): when max(k, k_int) * radius <= pi/2, with radius the largest
// distance of a collocation point from their mean (quadrature/fluid.jl:1-7;
// the collocation points here are the panel centroids).
double collocation_radius(const HostMesh& h) {
    double c[3] = {0, 0, 0};
    for (int64_t j = 0; j < h.nt; ++j)
        for (int d = 0; d < 3; ++d) c[d] += h.cen[3 * j + d];
    for (int d = 0; d < 3; ++d) c[d] /= (double)h.nt;
    double r = 0;
    for (int64_t j = 0; j < h.nt; ++j) {
        const double dx = h.cen[3 * j] - c[0], dy = h.cen[3 * j + 1] - c[1], dz = h.cen[3 * j + 2] - c[2];
        r = std::max(r, std::sqrt(dx * dx + dy * dy + dz * dz));
    }
    return r;
}
bool calderon_applies(const HostMesh& h, const bem_problem& p) {
    if (p.boundary != BEM_FLUID || p.formulation != BEM_MULLER) return false;
    const double kmax = std::max(p.k, p.k / p.soundspeed_contrast);
    return kmax * collocation_radius(h) <= M_PI / 2;
}
// Device bytes of the Calderon workspace: S, D and [K' | H] of both media in
// FP64, the pivots and the block inverses of S.
size_t calderon_bytes(int64_t nt) {
    return sizeof(z64) * 6 * (size_t)nt * nt + sizeof(int) * nt + sizeof(z64) * 2 * 4096 * (size_t)((nt + 63) / 64) +
           8 * 256;
}

// Overwrites the bottom block row of op->A (assembled with direct K' and H)
// with [-He + Hi / g, I + K'e - K'i] from the Calderon blocks of each medium:
// M = [D S | D D - I/4], then M = inv(S) M with the device LU of S.
template <typename R>
void calderon_blocks(bem_operator op) {
    Lane* L = op->lane;
    hipStream_t s = L->s;
    const int nt = op->nt;
    const bem_problem& pr = op->prob;
    const size_t sq = (size_t)nt * nt;
    z64* S = (z64*)op_alloc(op, sizeof(z64) * sq);
    z64* D = (z64*)op_alloc(op, sizeof(z64) * sq);
    z64* M[2] = {(z64*)op_alloc(op, sizeof(z64) * 2 * sq), (z64*)op_alloc(op, sizeof(z64) * 2 * sq)};
    int* ipiv = (int*)op_alloc(op, sizeof(int) * nt);
    z64* inv = (z64*)op_alloc(op, sizeof(z64) * 2 * 4096 * (size_t)((nt + 63) / 64));
    const int64_t ld2 = 2 * (int64_t)nt;
    for (int med = 0; med < 2; ++med) {
        AsmParams ap;
        ap.kind = ASM_SD;
        ap.k = med == 0 ? pr.k : pr.k / pr.soundspeed_contrast;
        ap.kint = ap.k;
        ap.g = 1.0;
        ap.ld = nt;
        if (op->mesh->d.nq) {
            nys_assemble<double>(op, ap, S, D, false);
        } else {
            launch_assemble<double>(s, op->mesh->d, ap, S, D);
            launch_near_self<double>(s, op->mesh->d, ap, S, D);
        }
        gemm_assign<double>(s, D, nt, S, nt, M[med], ld2, nt, nt, nt);
        gemm_assign<double>(s, D, nt, D, nt, M[med] + nt, ld2, nt, nt, nt);
        launch_add_diag(s, nt, M[med] + nt, ld2, -0.25);
        launch_fill_int(s, L->ws.status, 1, INT_MAX);
        launch_fill_int(s, L->ws.status + 1, 1, 0);
        LuStreams st;
        st.main = s;
        st.side = L->s2;
        st.ready = L->ev_ready;
        st.panel = L->ev_panel;
        const int rc = lu_factor<double>(st, S, nt, nt, ipiv, inv, L->ws, &L->pl);
        if (rc == 2)
            throw BemError(BEM_ERR_ADMISSION, "system larger than the panel workspace (" +
                                                  std::to_string(256 * (int64_t)L->ws.max_blocks) + " rows)");
        if (rc != 0) throw BemError(BEM_ERR_DEVICE, "LU panel launch failed");
        int status[2];
        HIPCHECK(hipMemcpyAsync(status, L->ws.status, sizeof status, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        if (status[1] != 0) throw BemError(BEM_ERR_DEVICE, "LU panel pivot selection was inconsistent");
        if (status[0] != INT_MAX)
            throw BemError(BEM_ERR_NUMERIC, "singular single-layer operator in the Calderon reconstruction:" +
                                                std::to_string(status[0]));
        lu_solve_many<double>(s, S, nt, nt, ipiv, inv, M[med], ld2, 2 * nt, &L->pl);
    }
    launch_calderon_bottom<R>(s, nt, M[0], M[1], pr.density_contrast, (C2<R>*)op->A);
    HIPCHECK(hipGetLastError());
}

// Reciprocal condition number in the infinity norm of the equilibrated
// single-precision matrix, as LAPACK gecon('I') (fluid_system.jl:218-222):
// rcond = 1 / (||Ahat||_inf * est), est ~ ||inv(Ahat)||_inf = ||inv(Ahat)^H||_1
// from the 1-norm estimator of Hager (1984) in Higham's form (ACM TOMS 14,
// 1988): products with B = inv(Ahat)^H and B^H = inv(Ahat) are LU solves.
double estimate_rcond(bem_operator op, double anorm) {
    const int n = op->n;
    hipStream_t s = op->lane->s;
    if (anorm == 0) return 0.0;
    typedef std::complex<double> cd;
    std::vector<cd> v(n);
    std::vector<z32> buf(n);
    auto apply = [&](bool adjoint_of_B) {
        for (int i = 0; i < n; ++i) buf[i] = z32{(float)v[i].real(), (float)v[i].imag()};
        HIPCHECK(hipMemcpyAsync(op->wv, buf.data(), sizeof(z32) * n, hipMemcpyHostToDevice, s));
        if (adjoint_of_B)
            lu_solve<float>(s, (const z32*)op->LU, n, n, op->perm, (const z32*)op->inv, (const z32*)op->wv,
                            (z32*)op->wv2, (z32*)op->wv3, op->lane->ws, &op->lane->pl);
        else
            lu_solve_h<float>(s, (const z32*)op->LU, n, n, op->perm, (const z32*)op->inv, (const z32*)op->wv,
                              (z32*)op->wv2, (z32*)op->wv3, op->lane->ws, &op->lane->pl);
        int expired = 0;
        HIPCHECK(hipMemcpyAsync(buf.data(), op->wv3, sizeof(z32) * n, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(&expired, op->lane->ws.status + 2, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        tri_check(op->lane, expired);
        for (int i = 0; i < n; ++i) v[i] = cd(buf[i].x, buf[i].y);
    };
    auto norm1 = [&]() {
        double t = 0;
        for (const cd& z : v) t += std::abs(z);
        return t;
    };
    auto sign = [&]() {
        for (cd& z : v) {
            const double a = std::abs(z);
            z = a > 0 ? z / a : cd(1, 0);
        }
    };
    auto argmax = [&]() {
        int j = 0;
        double m = -1;
        for (int i = 0; i < n; ++i)
            if (std::abs(v[i]) > m) { m = std::abs(v[i]); j = i; }
        return j;
    };
    for (int i = 0; i < n; ++i) v[i] = cd(1.0 / n, 0);
    apply(false);
    if (n == 1) return 1.0 / (anorm * std::abs(v[0]));
    double est = norm1();
    if (!std::isfinite(est)) return NAN;
    sign();
    apply(true);
    int j = argmax();
    for (int iter = 2; iter <= 5; ++iter) {
        std::fill(v.begin(), v.end(), cd(0, 0));
        v[j] = cd(1, 0);
        apply(false);
        const double old = est;
        est = norm1();
        if (!std::isfinite(est)) return NAN;
        if (est <= old) {
            est = old;
            break;
        }
        sign();
        apply(true);
        const int jlast = j;
        const double zlast = std::abs(v[jlast]);
        j = argmax();
        if (zlast == std::abs(v[j])) break;
    }
    // alternating-sign test vector (Higham's final safeguard)
    double alt = 1.0;
    for (int i = 0; i < n; ++i) {
        v[i] = cd(alt * (1.0 + (double)i / (n - 1)), 0);
        alt = -alt;
    }
    apply(false);
    const double temp = 2.0 * norm1() / (3.0 * n);
    if (temp > est) est = temp;
    if (!(est > 0)) return NAN;
    return (1.0 / est) / anorm;
}

void promote(bem_operator op, int reason) {
    bem_ctx ctx = op->ctx;
    hipStream_t s = op->lane->s;
    const int n = op->n;
    if (!op->LU64) op->LU64 = op_alloc(op, sizeof(z64) * (size_t)n * n);
    if (!op->ipiv64) op->ipiv64 = (int*)op_alloc(op, sizeof(int) * n);
    if (!op->perm64) op->perm64 = (int*)op_alloc(op, sizeof(int) * n);
    if (!op->inv64) op->inv64 = op_alloc(op, sizeof(z64) * 2 * 4096 * (size_t)((n + 63) / 64));
    if (op->single)
        launch_scale_copy<float, double>(s, (const z32*)op->A, n, n, op->rows, op->cols, (z64*)op->LU64, n);
    else
        launch_scale_copy<double, double>(s, (const z64*)op->A, n, n, op->rows, op->cols, (z64*)op->LU64, n);
    try {
        factor_into<double>(op, op->LU64, op->ipiv64, op->perm64, op->inv64);
    } catch (const BemError& e) {
        if (e.code == BEM_ERR_NUMERIC) throw BemError(BEM_ERR_NUMERIC, "double-precision factorization is singular");
        throw;
    }
    op->precision = 0;
    op->reason = reason;
}

// max over the kRedBlocks partial results in op->red (stride `stride`, slot `slot`)
double host_max(bem_operator op, int count, int stride, int slot, double* all = nullptr) {
    std::vector<double> h((size_t)count * stride);
    HIPCHECK(hipMemcpyAsync(h.data(), op->red, sizeof(double) * h.size(), hipMemcpyDeviceToHost, op->lane->s));
    HIPCHECK(hipStreamSynchronize(op->lane->s));
    if (all) {
        for (int q = 0; q < stride; ++q) {
            double m = 0;
            for (int b = 0; b < count; ++b) m = std::max(m, h[(size_t)b * stride + q]);
            all[q] = m;
        }
        return all[slot];
    }
    double m = 0;
    for (int b = 0; b < count; ++b) {
        double v = h[(size_t)b * stride + slot];
        if (!(v <= m)) m = v;  // propagates inf/nan
    }
    return m;
}

// delta = correction(F, rhs) (fluid_system.jl:229-239).  Returns false when
// the scaled right-hand side is non-finite (delta filled with NaN).
bool correction(bem_operator op, const z64* rhs, z64* out, int accumulate, int* changed) {
    bem_ctx ctx = op->ctx;
    hipStream_t s = op->lane->s;
    const int n = op->n;
    launch_max_abs_scaled(s, n, rhs, op->rows, op->red);
    const double mag = host_max(op, kRedBlocks, 1, 0);
    if (!std::isfinite(mag)) return false;
    if (mag == 0) {
        if (!accumulate) launch_fill_double(s, (double*)out, 2 * (int64_t)n, 0.0);
        return true;
    }
    if (op->precision == 1) {
        launch_prep_rhs<float>(s, n, rhs, op->rows, 1.0 / mag, (z32*)op->wv);
        lu_solve<float>(s, (const z32*)op->LU, n, n, op->perm, (const z32*)op->inv, (const z32*)op->wv,
                        (z32*)op->wv2, (z32*)op->wv3, op->lane->ws, &op->lane->pl);
        launch_finish<float>(s, n, (const z32*)op->wv3, mag, op->cols, out, accumulate, changed);
    } else {
        const z64* LU = (const z64*)(op->LU64 ? op->LU64 : op->LU);
        const int* perm = op->LU64 ? op->perm64 : op->perm;
        const z64* inv = (const z64*)(op->LU64 ? op->inv64 : op->inv);
        launch_prep_rhs<double>(s, n, rhs, op->rows, 1.0 / mag, (z64*)op->wv);
        lu_solve<double>(s, LU, n, n, perm, inv, (const z64*)op->wv, (z64*)op->wv2, (z64*)op->wv3, op->lane->ws,
                         &op->lane->pl);
        launch_finish<double>(s, n, (const z64*)op->wv3, mag, op->cols, out, accumulate, changed);
    }
    return true;
}

void residual(bem_operator op, const z64* x, const z64* b, z64* r, double* den) {
    bem_ctx ctx = op->ctx;
    prof_begin(&op->lane->pl, op->lane->s, C_RESIDUAL);
    if (op->single)
        launch_matvec<float>(op->lane->s, (const z32*)op->A, op->n, op->n, op->n, x, r, b, den);
    else
        launch_matvec<double>(op->lane->s, (const z64*)op->A, op->n, op->n, op->n, x, r, b, den);
    prof_end(&op->lane->pl, op->lane->s, C_RESIDUAL, 8.0 * (double)op->n * op->n,
             (double)op->n * op->n * (op->single ? 8.0 : 16.0));
}

struct BwdErr {
    double nw, cw;
};

BwdErr backward_errors(bem_operator op, const z64* x, const z64* b) {
    const int n = op->n;
    residual(op, x, b, op->vr, op->vden);
    launch_bwd_stats(op->lane->s, n, op->vr, op->vden, x, b, op->red);
    double all[5];
    host_max(op, kRedBlocks, 5, 0, all);
    if (all[4] != 0) return {INFINITY, INFINITY};
    const double rn = all[0], cw = all[1], xm = all[2], bm = all[3];
    const double scale = op->matrix_norm * xm + bm;
    double nw;
    if (!std::isfinite(scale)) nw = INFINITY;
    else if (scale == 0) nw = (rn == 0) ? 0.0 : INFINITY;
    else nw = rn / scale;
    return {nw, cw};
}

// solve_refined (fluid_system.jl:264-307); x in op->vx, b in op->vb
void solve_refined(bem_operator op, bem_diagnostics& dg) {
    bem_ctx ctx = op->ctx;
    hipStream_t s = op->lane->s;
    const int n = op->n;
    const double tol = op->prob.refine_tolerance;
    const int maxiter = op->prob.refine_maxiter;
    auto fresh = [&]() {
        if (!correction(op, op->vb, op->vx, 0, nullptr)) launch_fill_double(s, (double*)op->vx, 2 * (int64_t)n, NAN);
    };
    fresh();
    int corrections = 0, attempts = 0, stalls = 0;
    double previous = INFINITY;
    for (int guard = 0; guard < 1000; ++guard) {
        BwdErr e = backward_errors(op, op->vx, op->vb);
        const double metric = std::max(e.nw, e.cw);
        if (std::isfinite(metric) && metric <= tol) {
            dg.normwise_backward_error = e.nw;
            dg.componentwise_backward_error = e.cw;
            dg.refinement_iterations = corrections;
            dg.converged = 1;
            return;
        }
        stalls = (metric >= previous * 0.9) ? stalls + 1 : 0;
        int reason = BEM_REASON_NONE;
        if (!std::isfinite(metric)) reason = BEM_REASON_NONFINITE;
        else if (attempts >= maxiter) reason = BEM_REASON_ITERATION_LIMIT;
        else if (stalls >= 2) reason = BEM_REASON_STAGNATION;
        if (reason != BEM_REASON_NONE) {
            if (op->precision == 0)
                throw BemError(BEM_ERR_NUMERIC, "double-precision fallback failed backward-error acceptance");
            promote(op, reason);
            fresh();
            previous = INFINITY;
            attempts = stalls = 0;
            continue;
        }
        // delta = correction(F, residual); x += delta
        launch_fill_int(s, op->flag, 1, 0);
        if (!correction(op, op->vr, op->vx, 1, op->flag)) {
            launch_fill_double(s, (double*)op->vx, 2 * (int64_t)n, NAN);
        } else {
            int changed = 0;
            HIPCHECK(hipMemcpyAsync(&changed, op->flag, sizeof(int), hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            if (!changed) {
                if (op->precision == 0) throw BemError(BEM_ERR_NUMERIC, "double-precision correction stagnated");
                promote(op, BEM_REASON_STAGNATION);
                fresh();
                previous = INFINITY;
                attempts = stalls = 0;
                continue;
            }
        }
        previous = metric;
        ++corrections;
        ++attempts;
    }
    throw BemError(BEM_ERR_NUMERIC, "refinement did not terminate");
}

void op_free(bem_operator op) {
    bem_ctx c = op->ctx;
    if (op->lane) hipStreamSynchronize(op->lane->s);  // its buffers may return to the pool
    for (void* p : op->heap) c->release(p);
    if (op->use_arena) op->lane->off = op->arena_mark;
    mesh_release(op->mesh);
    delete op;
}

void operator_factor(bem_operator op);

// Assembles the system of `prob` on `mesh`; unless `defer_factor`, also
// equilibrates and factors it (operator_factor).
void operator_create(bem_ctx ctx, Lane* lane, bem_mesh mesh, const bem_problem* prob, bem_operator* out,
                     bool use_arena = false, bool defer_factor = false) {
    if (!mesh || !prob || !out) throw BemError(BEM_ERR_ARGUMENT, "null argument");
    validate_problem(*prob);
    auto* op = new bem_operator_s();
    op->ctx = ctx;
    op->lane = lane;
    op->use_arena = use_arena && lane->arena;
    op->arena_mark = lane->off;
    hipStream_t s = lane->s;
    op->mesh = mesh;
    mesh->refs++;
    op->prob = *prob;
    op->nt = mesh->d.nt;
    const bool fluid = prob->boundary == BEM_FLUID;
    op->n = fluid ? 2 * op->nt : op->nt;
    op->single = prob->precision == BEM_MIXED;
    op->precision = op->single ? 1 : 0;
    const int n = op->n;
    const size_t esz = op->single ? sizeof(z32) : sizeof(z64);
    try {
        const double t0 = now_ms();
        op->A = op_alloc(op, esz * (size_t)n * n);
        if (!fluid) op->B = op_alloc(op, esz * (size_t)n * n);
        AsmParams ap;
        ap.k = prob->k;
        ap.g = fluid ? prob->density_contrast : 1.0;
        ap.kint = fluid ? prob->k / prob->soundspeed_contrast : prob->k;
        ap.ld = n;
        void *o0 = op->A, *o1 = op->B;
        if (fluid) {
            ap.kind = prob->formulation == BEM_MULLER ? ASM_MULLER : ASM_FLUID_CBIE;
        } else {
            ap.kind = prob->formulation == BEM_BURTON_MILLER ? ASM_BM : ASM_CBIE_RS;
            // rigid solves out0 (1/2 I - D [- aH]); soft solves out1 (S [+ a(1/2 I + K')])
            if (prob->boundary == BEM_PRESSURE_RELEASE) std::swap(o0, o1);
        }
        op->calderon = calderon_applies(mesh->h, *prob);
        if (mesh->d.nq) {
            // the derivative operators are direct unless the Calderon blocks replace them
            const bool deriv = ap.kind == ASM_BM || (ap.kind == ASM_MULLER && !op->calderon);
            if (op->single) nys_assemble<float>(op, ap, (z32*)o0, (z32*)o1, deriv);
            else nys_assemble<double>(op, ap, (z64*)o0, (z64*)o1, deriv);
        } else {
            const double evals = (double)op->nt * op->nt * 7.0 * (fluid ? 2.0 : 1.0);
            prof_begin(&op->lane->pl, s, C_ASSEMBLE);
            if (op->single) launch_assemble<float>(s, mesh->d, ap, (z32*)o0, (z32*)o1);
            else launch_assemble<double>(s, mesh->d, ap, (z64*)o0, (z64*)o1);
            prof_end(&op->lane->pl, s, C_ASSEMBLE, evals, (double)n * n * esz * (fluid ? 1.0 : 2.0));
            prof_begin(&op->lane->pl, s, C_NEAR);
            if (op->single) launch_near_self<float>(s, mesh->d, ap, (z32*)o0, (z32*)o1);
            else launch_near_self<double>(s, mesh->d, ap, (z64*)o0, (z64*)o1);
            prof_end(&op->lane->pl, s, C_NEAR, (double)mesh->d.nnear * kNearPts * (fluid ? 2.0 : 1.0), 0.0);
        }
        if (op->calderon) {
            if (op->single) calderon_blocks<float>(op);
            else calderon_blocks<double>(op);
        }
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(s));
        const double t1 = now_ms();
        op->assembly_ms = t1 - t0;

        // scratch
        op->vb = (z64*)op_alloc(op, sizeof(z64) * n);
        op->vx = (z64*)op_alloc(op, sizeof(z64) * n);
        op->vr = (z64*)op_alloc(op, sizeof(z64) * n);
        op->vt = (z64*)op_alloc(op, sizeof(z64) * 2 * (size_t)n);
        op->vden = (double*)op_alloc(op, sizeof(double) * n);
        op->wv = op_alloc(op, sizeof(z64) * n);
        op->wv2 = op_alloc(op, sizeof(z64) * n);
        op->wv3 = op_alloc(op, sizeof(z64) * n);
        op->red = (double*)op_alloc(op, sizeof(double) * 5 * kRedBlocks);
        op->flag = (int*)op_alloc(op, sizeof(int) * 4);
        if (!defer_factor) operator_factor(op);
    } catch (...) {
        op_free(op);
        throw;
    }
    *out = op;
}

// Equilibration, factorization and the single-precision guards of an
// assembled operator (full.jl:364-409; fluid_system.jl:196-227).  On failure
// the caller frees the operator.
void operator_factor(bem_operator op) {
    hipStream_t s = op->lane->s;
    const bem_problem* prob = &op->prob;
    const int n = op->n;
    const size_t esz = op->single ? sizeof(z32) : sizeof(z64);
    const double t1 = now_ms();
    {
        op->ipiv = (int*)op_alloc(op, sizeof(int) * n);
        op->perm = (int*)op_alloc(op, sizeof(int) * n);
        op->inv = op_alloc(op, esz * 2 * 4096 * (size_t)((n + 63) / 64));
        op->LU = op_alloc(op, esz * (size_t)n * n);

        // equilibration + finiteness + matrix norm
        double* rowsum = (double*)op->vden;
        if (prob->equilibrate) {
            op->rows = (double*)op_alloc(op, sizeof(double) * n);
            op->cols = (double*)op_alloc(op, sizeof(double) * n);
        }
        double* rows = op->rows ? op->rows : (double*)op->vt;
        if (op->single) launch_row_stats<float>(s, (const z32*)op->A, n, n, rows, rowsum);
        else launch_row_stats<double>(s, (const z64*)op->A, n, n, rows, rowsum);
        {
            std::vector<double> hr(n), hs(n);
            HIPCHECK(hipMemcpyAsync(hr.data(), rows, sizeof(double) * n, hipMemcpyDeviceToHost, s));
            HIPCHECK(hipMemcpyAsync(hs.data(), rowsum, sizeof(double) * n, hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            double mn = 0;
            for (int i = 0; i < n; ++i) {
                if (!std::isfinite(hr[i])) throw BemError(BEM_ERR_ARGUMENT, "matrix must be finite");
                if (!std::isfinite(hs[i])) throw BemError(BEM_ERR_ARGUMENT, "matrix norm overflow");
                mn = std::max(mn, hs[i]);
            }
            op->matrix_norm = mn;
        }
        if (op->rows) {
            launch_fix_scale(s, n, op->rows);
            if (op->single) launch_col_max<float>(s, (const z32*)op->A, n, n, op->rows, op->cols);
            else launch_col_max<double>(s, (const z64*)op->A, n, n, op->rows, op->cols);
            launch_fix_scale(s, n, op->cols);
        }
        if (op->single) launch_scale_copy<float, float>(s, (const z32*)op->A, n, n, op->rows, op->cols, (z32*)op->LU, n);
        else launch_scale_copy<double, double>(s, (const z64*)op->A, n, n, op->rows, op->cols, (z64*)op->LU, n);
        double anorm_low = 0;
        const bool guard = op->single && prob->condition_guard >= 0;
        if (guard) {
            // ||Ahat||_inf of the equilibrated single-precision matrix
            launch_row_stats<float>(s, (const z32*)op->LU, n, n, (double*)op->vt, rowsum);
            std::vector<double> hs(n);
            HIPCHECK(hipMemcpyAsync(hs.data(), rowsum, sizeof(double) * n, hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            for (double v : hs) anorm_low = std::max(anorm_low, v);
        }
        try {
            if (op->single) factor_into<float>(op, op->LU, op->ipiv, op->perm, op->inv);
            else factor_into<double>(op, op->LU, op->ipiv, op->perm, op->inv);
        } catch (const BemError& e) {
            if (e.code != BEM_ERR_NUMERIC) throw;
            if (!op->single) throw BemError(BEM_ERR_NUMERIC, "matrix is singular (zero pivot)");
            promote(op, BEM_REASON_SINGLE_SINGULAR);
        }
        if (guard && op->precision == 1) {
            op->rcond = estimate_rcond(op, anorm_low);
            if (!std::isfinite(op->rcond) || op->rcond <= prob->condition_guard) promote(op, BEM_REASON_CONDITION);
        }
        if (op->single && op->precision == 1 && prob->force_promotion) promote(op, BEM_REASON_CONDITION);
        HIPCHECK(hipStreamSynchronize(s));
        op->factor_ms = now_ms() - t1;
    }
}

// One right-hand side: plane wave with device direction ddir (3 doubles);
// scattered traces into P, Q (nt each) and the report into dg.
// Plane-wave traces (op->vt: p_inc then q_inc) and the right-hand side b
// (op->vb) of the operator's equations for device direction ddir
// (full.jl:144-153,442-454).
void build_rhs(bem_operator op, const double* ddir) {
    hipStream_t s = op->lane->s;
    const int nt = op->nt, n = op->n;
    const bem_problem& pr = op->prob;
    const bool fluid = pr.boundary == BEM_FLUID;
    z64* pinc = op->vt;        // nt
    z64* qinc = op->vt + nt;   // nt
    launch_incident(s, op->mesh->d, pr.k, 1, ddir, pinc, qinc);
    // right-hand side b (n) in op->vb
    if (fluid && pr.formulation == BEM_MULLER) {
        HIPCHECK(hipMemcpyAsync(op->vb, pinc, sizeof(z64) * nt, hipMemcpyDeviceToDevice, s));
        HIPCHECK(hipMemcpyAsync(op->vb + nt, qinc, sizeof(z64) * nt, hipMemcpyDeviceToDevice, s));
    } else if (fluid) {
        // [0; -A[N+1:2N, :] * [p_inc; q_inc]]   (full.jl:448-453)
        launch_fill_double(s, (double*)op->vb, 2 * (int64_t)nt, 0.0);
        if (op->single)
            launch_matvec<float>(s, (const z32*)op->A + (size_t)nt * n, n, nt, n, op->vt, op->vr, nullptr, nullptr);
        else
            launch_matvec<double>(s, (const z64*)op->A + (size_t)nt * n, n, nt, n, op->vt, op->vr, nullptr, nullptr);
        launch_copy_z(s, nt, op->vr, op->vb + nt, -1.0, nullptr);
    } else {
        const z64* src = pr.boundary == BEM_RIGID ? qinc : pinc;
        if (op->single) launch_matvec<float>(s, (const z32*)op->B, n, n, n, src, op->vb, nullptr, nullptr);
        else launch_matvec<double>(s, (const z64*)op->B, n, n, n, src, op->vb, nullptr, nullptr);
    }
}

void solve_rhs(bem_operator op, const double* ddir, z64* P, z64* Q, bem_diagnostics& dg) {
    hipStream_t s = op->lane->s;
    const int nt = op->nt, n = op->n;
    const bem_problem& pr = op->prob;
    const bool fluid = pr.boundary == BEM_FLUID;
    const double t0 = now_ms();
    std::memset(&dg, 0, sizeof dg);
    dg.unknown_count = n;
    dg.quadrature_nodes = nt;
    z64* pinc = op->vt;        // nt
    z64* qinc = op->vt + nt;   // nt
    build_rhs(op, ddir);
    if (pr.precision == BEM_MIXED) {
        solve_refined(op, dg);
    } else {
        if (!correction(op, op->vb, op->vx, 0, nullptr))
            throw BemError(BEM_ERR_ARGUMENT, "right-hand side must be finite");
        dg.converged = 1;
    }
    dg.factor_precision = op->precision;
    dg.method = pr.precision == BEM_MIXED ? 1 : 0;
    dg.derivative_evaluation = (fluid && pr.formulation == BEM_MULLER) ? (op->calderon ? 2 : 1) : 0;
    dg.refinement_reason = op->reason;
    dg.single_reciprocal_condition = op->rcond;
    // residual report on the original equations: r = b - A x
    residual(op, op->vx, op->vb, op->vr, nullptr);
    launch_norm_stats(s, n, op->vr, op->vb, op->rows, op->red);
    {
        std::vector<double> h(4 * kRedBlocks);
        int expired = 0;
        HIPCHECK(hipMemcpyAsync(h.data(), op->red, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(&expired, op->lane->ws.status + 2, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        tri_check(op->lane, expired);
        double ss[4] = {0, 0, 0, 0};
        for (int b = 0; b < kRedBlocks; ++b)
            for (int q = 0; q < 4; ++q) ss[q] += h[4 * b + q];
        dg.absolute_residual = std::sqrt(ss[0]);
        const double bn = std::sqrt(ss[1]);
        dg.relative_residual = bn == 0 ? (dg.absolute_residual == 0 ? 0.0 : INFINITY) : dg.absolute_residual / bn;
        dg.scaled_absolute_residual = std::sqrt(ss[2]);
        const double sbn = std::sqrt(ss[3]);
        dg.scaled_relative_residual =
            sbn == 0 ? (dg.scaled_absolute_residual == 0 ? 0.0 : INFINITY) : dg.scaled_absolute_residual / sbn;
    }
    // traces (full.jl:160-161,472-477)
    if (fluid) {
        const bool mul = pr.formulation == BEM_MULLER;
        launch_copy_z(s, nt, op->vx, P, 1.0, mul ? pinc : nullptr);
        launch_copy_z(s, nt, op->vx + nt, Q, 1.0, mul ? qinc : nullptr);
    } else if (pr.boundary == BEM_RIGID) {
        launch_copy_z(s, nt, op->vx, P, 1.0, nullptr);
        launch_copy_z(s, nt, qinc, Q, -1.0, nullptr);
    } else {
        launch_copy_z(s, nt, pinc, P, -1.0, nullptr);
        launch_copy_z(s, nt, op->vx, Q, 1.0, nullptr);
    }
    HIPCHECK(hipStreamSynchronize(s));
    dg.assembly_ms = op->assembly_ms;
    dg.factor_ms = op->factor_ms;
    dg.solve_ms = now_ms() - t0;
}

void operator_solve(bem_operator op, int nrhs, const double* angles, bem_solution* out) {
    if (nrhs <= 0 || !angles || !out) throw BemError(BEM_ERR_ARGUMENT, "need at least one incidence");
    for (int r = 0; r < 2 * nrhs; ++r)
        if (!std::isfinite(angles[r])) throw BemError(BEM_ERR_ARGUMENT, "incidence angles must be finite");
    bem_ctx ctx = op->ctx;
    hipStream_t s = op->lane->s;
    const int nt = op->nt;
    auto* sol = new bem_solution_s();
    sol->ctx = ctx;
    sol->mesh = op->mesh;
    op->mesh->refs++;
    sol->k = op->prob.k;
    sol->prob = op->prob;
    sol->nrhs = nrhs;
    sol->nt = nt;
    sol->angles.assign(angles, angles + 2 * nrhs);
    sol->diag.resize(nrhs);
    double* ddir = nullptr;
    try {
        sol->p = (z64*)ctx->alloc(sizeof(z64) * (size_t)nt * nrhs);
        sol->q = (z64*)ctx->alloc(sizeof(z64) * (size_t)nt * nrhs);
        std::vector<double> dirs(3 * nrhs);
        for (int r = 0; r < nrhs; ++r) direction(angles[2 * r], angles[2 * r + 1], &dirs[3 * r]);
        ddir = (double*)ctx->alloc(sizeof(double) * 3 * nrhs);
        HIPCHECK(hipMemcpyAsync(ddir, dirs.data(), sizeof(double) * 3 * nrhs, hipMemcpyHostToDevice, s));
        for (int r = 0; r < nrhs; ++r)
            solve_rhs(op, ddir + 3 * r, sol->p + (size_t)r * nt, sol->q + (size_t)r * nt, sol->diag[r]);
        sol->hp.resize((size_t)nt * nrhs);
        sol->hq.resize((size_t)nt * nrhs);
        HIPCHECK(hipMemcpyAsync(sol->hp.data(), sol->p, sizeof(z64) * nt * nrhs, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(sol->hq.data(), sol->q, sizeof(z64) * nt * nrhs, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        ctx->release(ddir);
    } catch (...) {
        hipStreamSynchronize(s);  // nothing pending may touch blocks that return to the pool
        ctx->release(ddir);
        ctx->release(sol->p);
        ctx->release(sol->q);
        mesh_release(sol->mesh);
        delete sol;
        throw;
    }
    *out = sol;
}

// Unknowns of one sweep item's dense system (Mueller fluid systems carry
// pressure and normal derivative per panel).
size_t sweep_unknowns(const bem_mesh m, const bem_problem& p) {
    const size_t nt = (size_t)m->d.nt;
    return p.boundary == BEM_FLUID ? 2 * nt : nt;
}

// From this many unknowns on, the first trailing updates of an LU are GEMMs of
// over 15000 128 x 128 tiles (about 60 per CU), and the factorization is bound
// by the matrix cores rather than by its panel chain (at 10240 unknowns the
// panel chain takes about three times the GEMM time).
constexpr size_t kHeavyUnknowns = 16384;

// Device bytes one sweep item holds in a lane arena.
size_t sweep_need(const bem_mesh m, const bem_problem& p) {
    const size_t nt = (size_t)m->d.nt;
    const size_t n = sweep_unknowns(m, p);
    const size_t esz = p.precision == BEM_MIXED ? sizeof(z32) : sizeof(z64);
    size_t b = esz * n * n * (p.boundary == BEM_FLUID ? 2 : 3);    // A, LU (+ B)
    b += esz * 2 * 4096 * ((n + 63) / 64);                           // block inverses
    if (p.precision == BEM_MIXED) b += lu_split_bytes((int64_t)n);  // bf16 planes of the outer updates
    b += sizeof(z64) * n * 8 + sizeof(double) * n * 3 + sizeof(int) * n * 2;
    b += sizeof(z64) * (2 * nt + kFarBlocks + 4) + 64 * 1024;
    if (calderon_applies(m->h, p)) b += calderon_bytes(m->d.nt);
    b += nys_bytes(m);
    return b;
}

// Device bytes of the buffers backscatter() needs for nt panels.
size_t backscatter_scratch(int nt) {
    return sizeof(z64) * ((size_t)2 * nt + kFarBlocks + 2) + sizeof(double) * 8 + 4 * 256;
}

// Backscatter amplitude of one plane wave (incidence ang[0], ang[1]) through
// op on its lane: traces by solve_rhs, then the FP64 far field at -d.
void backscatter(bem_operator op, char* scratch, const double* ang, double* amp, bem_diagnostics* dgo) {
    Lane* L = op->lane;
    const int nt = op->nt;
    char* cur = scratch;
    auto take = [&](size_t b) {
        char* q = cur;
        cur += (b + 255) & ~(size_t)255;
        return q;
    };
    z64* P = (z64*)take(sizeof(z64) * nt);
    z64* Q = (z64*)take(sizeof(z64) * nt);
    z64* part = (z64*)take(sizeof(z64) * kFarBlocks);
    z64* res = (z64*)take(sizeof(z64) * 2);
    double* dd = (double*)take(sizeof(double) * 8);
    double hd[6];
    direction(ang[0], ang[1], hd);
    for (int c = 0; c < 3; ++c) hd[3 + c] = -hd[c];
    HIPCHECK(hipMemcpyAsync(dd, hd, sizeof hd, hipMemcpyHostToDevice, L->s));
    bem_diagnostics dg;
    solve_rhs(op, dd, P, Q, dg);
    prof_begin(&L->pl, L->s, C_FAR);
    launch_far_field(L->s, op->mesh->d, op->prob.k, 1, 1, dd + 3, P, Q, part, res);
    prof_end(&L->pl, L->s, C_FAR, (double)nt * 7, (double)nt * 7 * 8);
    HIPCHECK(hipMemcpyAsync(amp, res, sizeof(z64), hipMemcpyDeviceToHost, L->s));
    HIPCHECK(hipStreamSynchronize(L->s));
    if (dgo) *dgo = dg;
}

// One sweep item on lane L: assemble, factor, solve, backscatter amplitude.
void sweep_item(bem_ctx ctx, Lane* L, bem_mesh mesh, const bem_problem& prob, const double* ang, double* amp,
                bem_diagnostics* dgo) {
    L->off = 0;
    bem_operator op = nullptr;
    operator_create(ctx, L, mesh, &prob, &op, true);
    try {
        char* scratch = (char*)op_alloc(op, backscatter_scratch(op->nt));
        backscatter(op, scratch, ang, amp, dgo);
    } catch (...) {
        op_free(op);
        throw;
    }
    op_free(op);
}

// ---------------------------------------------------------------------------
// serving (bem_server_*): worker lanes, HBM admission, resident operators
// ---------------------------------------------------------------------------
struct ServeReq {
    int64_t id = 0;
    bem_mesh mesh = nullptr;
    bem_problem prob{};
    double ang[2] = {0, 0};
    size_t need = 0;
    double t_submit = 0, t_start = 0, t_end = 0;
    int status = BEM_OK;
    std::string err;
    double amp[2] = {0, 0};
    bem_diagnostics diag{};
    bool done = false;
};

// An operator kept resident after its request: it lives in `block`.
struct Resident {
    bem_mesh mesh = nullptr;
    bem_problem prob{};
    bem_operator op = nullptr;
    char* block = nullptr;
    size_t bytes = 0;
    bool busy = false;
    uint64_t stamp = 0;
};

struct ServeLane {
    Lane* L = nullptr;
    char* scratch = nullptr;
    size_t scap = 0;
};

bool same_problem(const bem_problem& a, const bem_problem& b) {
    return a.boundary == b.boundary && a.formulation == b.formulation && a.precision == b.precision &&
           a.equilibrate == b.equilibrate && a.k == b.k && a.density_contrast == b.density_contrast &&
           a.soundspeed_contrast == b.soundspeed_contrast && a.refine_tolerance == b.refine_tolerance &&
           a.refine_maxiter == b.refine_maxiter && a.force_promotion == b.force_promotion &&
           a.condition_guard == b.condition_guard;
}

}  // namespace

struct bem_server_s {
    bem_ctx ctx = nullptr;
    bool keep = false;
    std::vector<ServeLane> lanes;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ServeReq*> queue;
    std::unordered_map<int64_t, ServeReq*> reqs;
    std::list<Resident> resident;
    int64_t next_id = 1, submitted = 0, completed = 0, failed = 0, running = 0, reuses = 0;
    uint64_t stamp = 0;
    bool stop = false;
    std::string msg;
};

namespace {

void evict(bem_server srv, std::list<Resident>::iterator it) {
    op_free(it->op);  // resident operators own no arena: this frees heap buffers only
    srv->ctx->release(it->block);
    mesh_release(it->mesh);
    srv->resident.erase(it);
}

void finish(bem_server srv, ServeReq* r, int status, const std::string& err) {
    r->status = status;
    r->err = err;
    r->done = true;
    if (r->t_end == 0) r->t_end = now_ms();
    if (status == BEM_OK) ++srv->completed;
    else ++srv->failed;
    mesh_release(r->mesh);
    r->mesh = nullptr;
}

// Admission, with srv->mu held.  Oldest first: a request whose operator is
// resident and idle runs on it; otherwise it needs a device block of its
// estimated size, made from free budget plus idle resident operators evicted
// least recently used.  The oldest request that does not fit keeps a claim on
// memory: later requests are admitted only from memory beyond its need.
bool admit(bem_server srv, ServeReq*& out, std::list<Resident>::iterator& hit, char*& block) {
    bem_ctx ctx = srv->ctx;
    size_t claim = 0;
    hit = srv->resident.end();
    for (auto qi = srv->queue.begin(); qi != srv->queue.end(); ++qi) {
        ServeReq* r = *qi;
        auto res = srv->resident.end();
        for (auto it = srv->resident.begin(); it != srv->resident.end(); ++it)
            if (it->mesh == r->mesh && same_problem(it->prob, r->prob)) res = it;
        if (res != srv->resident.end()) {
            if (res->busy) continue;  // in use on another lane: wait for it
            res->busy = true;
            hit = res;
            out = r;
            srv->queue.erase(qi);
            return true;
        }
        uint64_t used, budget;
        {
            std::lock_guard<std::mutex> g(ctx->amu);
            used = ctx->used;
            budget = ctx->budget;
        }
        uint64_t freeb = budget > used ? budget - used : 0, idle = 0;
        for (auto& e : srv->resident)
            if (!e.busy) idle += e.bytes;
        if (r->need + claim <= freeb + idle) {
            while (freeb < r->need + claim) {
                auto lru = srv->resident.end();
                for (auto it = srv->resident.begin(); it != srv->resident.end(); ++it)
                    if (!it->busy && (lru == srv->resident.end() || it->stamp < lru->stamp)) lru = it;
                if (lru == srv->resident.end()) break;
                freeb += lru->bytes;
                evict(srv, lru);
            }
            try {
                block = (char*)ctx->alloc(r->need);
            } catch (const BemError&) {
                block = nullptr;
            }
            if (block) {
                out = r;
                srv->queue.erase(qi);
                return true;
            }
        }
        if (claim == 0) claim = r->need;
    }
    // nothing runs, so nothing will free memory: the oldest request cannot fit
    if (srv->running == 0 && !srv->queue.empty()) {
        bool busy_wait = false;
        for (auto& e : srv->resident) busy_wait |= e.busy;
        if (!busy_wait) {
            ServeReq* r = srv->queue.front();
            srv->queue.pop_front();
            finish(srv, r, BEM_ERR_ADMISSION, "request of " + std::to_string(r->need) +
                                                  " bytes does not fit in the free HBM budget");
            srv->cv_done.notify_all();
            return admit(srv, out, hit, block);  // the rest of the queue may fit now
        }
    }
    return false;
}

void serve_worker(bem_server srv, int li) {
    (void)hipSetDevice(srv->ctx->device);
    bem_ctx ctx = srv->ctx;
    ServeLane& SL = srv->lanes[li];
    Lane* L = SL.L;
    std::unique_lock<std::mutex> lk(srv->mu);
    for (;;) {
        ServeReq* r = nullptr;
        std::list<Resident>::iterator hit;
        char* block = nullptr;
        while (!srv->stop && !admit(srv, r, hit, block)) srv->cv_work.wait(lk);
        if (srv->stop && !r) return;
        const bool reuse = hit != srv->resident.end();
        ++srv->running;
        lk.unlock();
        bem_operator built = nullptr;
        int status = BEM_OK;
        std::string err;
        r->t_start = now_ms();
        try {
            bem_operator op;
            if (reuse) {
                op = hit->op;
                op->lane = L;
            } else {
                L->arena = block;
                L->cap = r->need;
                L->off = 0;
                operator_create(ctx, L, r->mesh, &r->prob, &op, true);
                built = op;
            }
            const size_t sb = backscatter_scratch(op->nt);
            if (SL.scap < sb) {
                ctx->release(SL.scratch);
                SL.scratch = nullptr;
                SL.scap = 0;
                SL.scratch = (char*)ctx->alloc(sb);
                SL.scap = sb;
            }
            backscatter(op, SL.scratch, r->ang, r->amp, &r->diag);
        } catch (const BemError& e) {
            status = e.code;
            err = e.what();
        } catch (const std::exception& e) {
            status = BEM_ERR_DEVICE;
            err = e.what();
        }
        r->t_end = now_ms();
        const bool keep = !reuse && built && srv->keep && status == BEM_OK;
        if (!reuse && !keep) {  // back to the context's pool, outside the server lock
            if (built) op_free(built);
            ctx->release(block);
        }
        lk.lock();
        if (reuse) {
            hit->busy = false;
            hit->stamp = ++srv->stamp;
            ++srv->reuses;
        } else if (keep) {
            built->use_arena = false;  // owned by the resident block from now on
            Resident e;
            e.mesh = r->mesh;
            e.mesh->refs++;
            e.prob = r->prob;
            e.op = built;
            e.block = block;
            e.bytes = r->need;
            e.stamp = ++srv->stamp;
            srv->resident.push_back(e);
        }
        L->arena = nullptr;
        L->cap = L->off = 0;
        finish(srv, r, status, err);
        --srv->running;
        srv->cv_work.notify_all();
        srv->cv_done.notify_all();
    }
}

template <typename F>
int srv_guard(bem_server srv, F&& f) {
    try {
        f();
        return BEM_OK;
    } catch (const BemError& e) {
        std::lock_guard<std::mutex> g(srv->mu);
        srv->msg = e.what();
        return e.code;
    } catch (const std::exception& e) {
        std::lock_guard<std::mutex> g(srv->mu);
        srv->msg = e.what();
        return BEM_ERR_DEVICE;
    }
}

}  // namespace

// ===========================================================================
// Hierarchical-matrix path: build (plan, dense leaves, ACA), apply, GMRES
// ===========================================================================
// One compressed operator over the shared plan: dense leaves (same layout for
// every operator of the plan), compact ACA factors, projection tiles and the
// per-leaf block lists of the apply.
struct HOp {
    int kind = HK_BM;
    std::vector<HBlk> blk;  // dense [0, nd), low-rank [nd, nd + nl)
    int ntiles = 0, nsplit = 0;
    HBlk* dblk = nullptr;
    z32* dval = nullptr;
    z32* uv = nullptr;
    int64_t* toff = nullptr;
    HTile* tiles = nullptr;
    HSplit* splits = nullptr;
    z64* P = nullptr;         // segment slabs of the split projection tiles
    int64_t* vrow = nullptr;  // V row offsets of the stacked projection rows
    int64_t v_entries = 0, lowrank_entries = 0;  // sum rank * n and sum rank * (m + n)
    z64* T = nullptr;
    HLeafs L;
    int max_rank = 0, unconverged = 0;
    double mean_rank = 0, aca_ms = 0, assembly_ms = 0;
    uint64_t lowrank_bytes = 0;
};

struct bem_hmatrix_s {
    bem_ctx ctx = nullptr;
    bem_mesh mesh = nullptr;
    bem_problem prob{};
    bem_hmatrix_options opt{};
    Lane* lane = nullptr;
    int64_t n = 0;                // unknowns: panels (rigid) or 2 x panels (fluid Mueller)
    int64_t nt = 0;               // panels
    bool fluid = false;
    HPar par{};
    int nd = 0, nl = 0, nleaf = 0;
    std::vector<void*> bufs;
    int* perm = nullptr;
    std::vector<int> hperm, leaf_lo, leaf_m, leaf_of_pos;
    std::vector<HBlk> plan_blk;   // first-pass layout (rank capacity 64)
    int64_t dense_count = 0, first_cap = 0;
    // near-pair and same-panel slots of the dense leaves (unknown indices)
    int *npi = nullptr, *npj = nullptr, *nsi = nullptr, *nsj = nullptr;
    int64_t *ndst = nullptr, *sdst = nullptr;
    int64_t nnear = 0, nself = 0;
    HOp A, B;        // system and right-hand-side operators
    z64* minv = nullptr;
    // two-level preconditioner: orthonormal coarse basis Q (n x nq, tree
    // order, column-major), Cc = Q - M^{-1} A Q and E^{-1} = (Q^H A Q)^{-1}
    int nq = 0;
    z64 *cq = nullptr, *cc = nullptr, *einv = nullptr;
    // recycled solutions (GMRES warm starts across solves): U and images C = A U
    // with C^H C = I, device (n x rcap) and host copies
    int rcap = 0, rcount = 0;
    z64 *ru = nullptr, *rc = nullptr;
    std::vector<std::complex<double>> hru, hrc;
    // Low-frequency Mueller (max(k, k_int) * radius <= pi/2): the system acts
    // on [p; q] (rows 0..N-1 and N..2N-1, each in tree order over a panel
    // tree) through compressed S and (1/2 I - D) of both media, with K' and H
    // from the Calderon identities applied by inner solves with S.
    bool calderon = false;
    int64_t inner_iters = 0;          // inner S solves' GMRES iterations (diagnostics)
    HOp cal[4];                       // S_e, (1/2 I - D_e), S_i, (1/2 I - D_i)
    z64 *minv_e = nullptr, *minv_i = nullptr;  // block Jacobi of S_e, S_i
    z64* lblk = nullptr;              // inverses of the local 2m x 2m Mueller blocks (nleaf x 128 x 128)
    // Nystrom meshes: density-interpolation corrections of each medium (0
    // exterior, 1 interior of fluid bodies)
    z64 *nys_sd[2] = {nullptr, nullptr}, *nys_kh[2] = {nullptr, nullptr};
    int nys_med = 0;  // medium of the single-medium operator being built (Calderon form)
    int64_t nsys = 0;                 // unknowns of the system (n, or 2N on the panel tree)
    std::vector<int> sysperm;         // unknown of each system row (nsys)
    int* dsysperm = nullptr;
    bem_hmatrix_info info{};
};

namespace {

// A linear system for the batched GMRES: y = A x and the right preconditioner
// y = P^{-1} x on n rows x nr right-hand sides (layout [row * nr + c]), with
// `work` scratch of work_size(nr) complex doubles for either.
struct HSys {
    int64_t n = 0;
    double tol = 1e-8;
    int restart = 50, maxiter = 1000;
    std::function<void(const z64*, int, z64*, z64*)> apply, prec;
    std::function<size_t(int)> work_size;
};

void hm_gmres(bem_hmatrix h, int nr, const z64* B, z64* X, std::vector<bem_diagnostics>& dg,
              const double* target = nullptr, const HSys* sys = nullptr);
HSys hm_system(bem_hmatrix h);
void hm_apply(bem_hmatrix h, HOp& op, const z64* X, int nr, z64* Y);
void hm_coarse_with(bem_hmatrix h, const std::function<void(const z64*, int, z64*)>& apply,
                    const std::function<void(const z64*, int, z64*)>& local);
void hm_coarse_correct(bem_hmatrix h, const z64* X, int nr, z64* Y, z64* work);

void* hm_alloc(bem_hmatrix h, size_t bytes) {
    void* p = h->ctx->alloc(bytes);
    h->bufs.push_back(p);
    return p;
}
template <typename T>
T* hm_upload(bem_hmatrix h, const std::vector<T>& v) {
    T* p = (T*)hm_alloc(h, sizeof(T) * std::max<size_t>(1, v.size()));
    if (!v.empty()) HIPCHECK(hipMemcpyAsync(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, h->lane->s));
    return p;
}

void hm_free(bem_hmatrix h) {
    if (h->lane) hipStreamSynchronize(h->lane->s);
    for (void* p : h->bufs) h->ctx->release(p);
    h->bufs.clear();
    if (h->lane) free_lane(h->ctx, h->lane);
    h->lane = nullptr;
    if (h->mesh) mesh_release(h->mesh);
    h->mesh = nullptr;
}

// Plan: cluster tree, block tree, dense-leaf layout and near/self slots.
void hm_plan(bem_hmatrix h) {
    const HostMesh& hm = h->mesh->h;
    const int64_t nt = hm.nt;
    // fluid Mueller: unknown i is the pressure and nt + i the normal
    // derivative on panel i, both clustered at the panel centroid
    const int nq = (h->fluid && !h->calderon) ? 2 : 1;
    const int64_t n = nq * nt;
    h->n = n;
    h->nt = nt;
    const auto& o = h->opt;
    const double t0 = now_ms();
    std::vector<double> pts(hm.cen.begin(), hm.cen.end());
    if (nq == 2) pts.insert(pts.end(), hm.cen.begin(), hm.cen.end());
    // near pairs never fall in low-rank blocks (box distance > 2 hmax); on
    // Nystrom meshes the pairs of nodes of one element (the corrected ones)
    // never do (box distance > the largest element's node spread)
    double near = 0;
    if (hm.nq) {
        for (int64_t e = 0; e < hm.nel; ++e)
            for (int a = 0; a < hm.nq; ++a)
                for (int b = a + 1; b < hm.nq; ++b) {
                    const double* x = &hm.cen[3 * (e * hm.nq + a)];
                    const double* y = &hm.cen[3 * (e * hm.nq + b)];
                    near = std::max(near, std::sqrt((x[0] - y[0]) * (x[0] - y[0]) + (x[1] - y[1]) * (x[1] - y[1]) +
                                                    (x[2] - y[2]) * (x[2] - y[2])));
                }
        near *= 1.0 + 1e-9;
    } else {
        double hmax = 0;
        for (double v : hm.hmax) hmax = std::max(hmax, v);
        near = kNearEta * hmax * (1.0 + 1e-9);
    }
    HPlan P;
    hmatrix_plan(pts.data(), n, o.leaf_size, o.eta, near, P);
    std::vector<int> leaf_of_node(P.nodes.size(), -1);
    std::vector<std::pair<int64_t, int32_t>> lv;
    for (size_t id = 0; id < P.nodes.size(); ++id)
        if (P.nodes[id].child[0] < 0) lv.push_back({P.nodes[id].lo, (int32_t)id});
    std::sort(lv.begin(), lv.end());
    h->nleaf = (int)lv.size();
    h->leaf_of_pos.assign(n, 0);
    for (int l = 0; l < h->nleaf; ++l) {
        const HNode& nd = P.nodes[lv[l].second];
        leaf_of_node[lv[l].second] = l;
        h->leaf_lo.push_back((int)nd.lo);
        h->leaf_m.push_back((int)(nd.hi - nd.lo));
        for (int64_t q = nd.lo; q < nd.hi; ++q) h->leaf_of_pos[q] = l;
    }
    h->nd = (int)P.dense.size();
    h->nl = (int)P.lowrank.size();
    std::unordered_map<int64_t, int> dmap;
    int64_t doff = 0;
    for (int b = 0; b < h->nd; ++b) {
        const HNode &r = P.nodes[P.dense[b].row], &c = P.nodes[P.dense[b].col];
        HBlk x{(int)r.lo, (int)(r.hi - r.lo), (int)c.lo, (int)(c.hi - c.lo), 0, 0, doff};
        doff += (int64_t)x.m * x.n;
        h->plan_blk.push_back(x);
        dmap[(int64_t)leaf_of_node[P.dense[b].row] * h->nleaf + leaf_of_node[P.dense[b].col]] = b;
    }
    h->dense_count = doff;
    std::vector<int> npi, npj, nsi, nsj;
    std::vector<int64_t> ndst, sdst;
    auto slot = [&](int64_t I, int64_t J) {
        const int64_t pI = P.iperm[I], pJ = P.iperm[J];
        auto it = dmap.find((int64_t)h->leaf_of_pos[pI] * h->nleaf + h->leaf_of_pos[pJ]);
        if (it == dmap.end()) throw BemError(BEM_ERR_DEVICE, "near pair outside the dense leaves");
        const HBlk& b = h->plan_blk[it->second];
        return b.off + (pI - b.row_lo) * b.n + (pJ - b.col_lo);
    };
    for (int64_t I = 0; hm.nq && I < n; ++I) {  // Nystrom: the node pairs of each element (both quadrants)
        const int64_t e0 = ((I % nt) / hm.nq) * hm.nq;
        for (int q = 0; q < nq; ++q)
            for (int b = 0; b < hm.nq; ++b) {
                const int64_t J = e0 + b + q * nt;
                nsi.push_back((int)I);
                nsj.push_back((int)J);
                sdst.push_back(slot(I, J));
            }
    }
    for (int64_t I = 0; !hm.nq && I < n; ++I) {
        const int64_t i = I % nt;
        for (int32_t e = hm.near_ptr[i]; e < hm.near_ptr[i + 1]; ++e)
            for (int q = 0; q < nq; ++q) {
                const int64_t J = hm.near_col[e] + q * nt;
                npi.push_back((int)I);
                npj.push_back((int)J);
                ndst.push_back(slot(I, J));
            }
        for (int q = 0; q < nq; ++q) {  // unknowns on the same panel
            const int64_t J = i + q * nt;
            nsi.push_back((int)I);
            nsj.push_back((int)J);
            sdst.push_back(slot(I, J));
        }
    }
    const int cap1 = std::min(64, o.max_rank);
    const double kmax = std::max(h->par.k, h->par.ki);
    int64_t uoff = 0;
    for (int b = 0; b < h->nl; ++b) {
        const HNode &r = P.nodes[P.lowrank[b].row], &c = P.nodes[P.lowrank[b].col];
        HBlk x{(int)r.lo, (int)(r.hi - r.lo), (int)c.lo, (int)(c.hi - c.lo), 0, 0, uoff};
        // the largest blocks (long rank chains) start with a larger capacity,
        // so that few of them need the second pass, which runs them alone
        const int cap = std::min(x.m, x.n) >= kAcaWide ? std::min(kAcaWideCap, o.max_rank) : cap1;
        x.rmax = std::min({x.m, x.n, cap});
        if (h->fluid && !h->calderon) {
            double d2 = 0;
            for (int q = 0; q < 3; ++q) {
                const double gap = std::max({0.0, c.bmin[q] - r.bmax[q], r.bmin[q] - c.bmax[q]});
                d2 += gap * gap;
            }
            x.sig = (float)(kmax + 1.0 / std::sqrt(d2));  // admissible blocks have dist > 0
        }
        uoff += (int64_t)x.rmax * (x.m + x.n);
        h->plan_blk.push_back(x);
    }
    h->first_cap = uoff;
    h->hperm.assign(P.perm.begin(), P.perm.end());
    h->perm = hm_upload(h, h->hperm);
    h->npi = hm_upload(h, npi);
    h->npj = hm_upload(h, npj);
    h->ndst = hm_upload(h, ndst);
    h->nsi = hm_upload(h, nsi);
    h->nsj = hm_upload(h, nsj);
    h->sdst = hm_upload(h, sdst);
    h->nnear = (int64_t)npi.size();
    h->nself = (int64_t)nsi.size();
    h->info.n = n;
    h->info.nodes = (int64_t)P.nodes.size();
    h->info.leaves = h->nleaf;
    h->info.dense_blocks = h->nd;
    h->info.lowrank_blocks = h->nl;
    h->info.near_pairs = h->nnear;
    h->info.dense_bytes = sizeof(z32) * (uint64_t)doff;
    h->info.plan_ms = now_ms() - t0;
}

// Assemble one operator over the plan: dense leaves, ACA (two passes),
// compaction, projection tiles, per-leaf lists.
void hm_operator(bem_hmatrix h, HOp& op, int kind) {
    bem_ctx ctx = h->ctx;
    const DevMesh& dm = h->mesh->d;
    hipStream_t s = h->lane->s;
    const auto& o = h->opt;
    const int nd = h->nd, nl = h->nl;
    op.kind = kind;
    op.blk = h->plan_blk;
    const double t1 = now_ms();
    op.dblk = hm_upload(h, op.blk);
    op.dval = (z32*)hm_alloc(h, sizeof(z32) * std::max<int64_t>(1, h->dense_count));
    launch_hdense(s, dm, kind, h->par, h->perm, nd, op.dblk, op.dval);
    if (dm.nq)
        launch_hnys_self(s, dm, kind, h->par, h->nself, h->nsi, h->nsj, h->sdst, h->nys_sd[h->nys_med],
                         h->nys_kh[h->nys_med], h->nys_sd[1], h->nys_kh[1], op.dval);
    else
        launch_hnear_self(s, dm, kind, h->par, h->nnear, h->npi, h->npj, h->ndst, h->nself, h->nsi, h->nsj,
                          h->sdst, op.dval);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(s));
    const double t2 = now_ms();
    op.assembly_ms = t2 - t1;
    std::vector<int> list(nl);
    for (int b = 0; b < nl; ++b) list[b] = nd + b;
    std::stable_sort(list.begin(), list.end(), [&](int a, int b) {
        return (int64_t)op.blk[a].m * op.blk[a].n > (int64_t)op.blk[b].m * op.blk[b].n;
    });
    int* dlist = hm_upload(h, list);
    // the list's prefix of wide blocks (sorted by m n, so min(m, n) >= kAcaWide
    // only among the first ones), run on the side stream
    auto wide_prefix = [&](const std::vector<int>& l) {
        int nb = 0;
        while (nb < (int)l.size() && std::min(op.blk[l[nb]].m, op.blk[l[nb]].n) >= kAcaWide) ++nb;
        return nb;
    };
    int* conv = (int*)hm_alloc(h, sizeof(int) * (nd + nl + 1));
    z32* uv1 = (z32*)ctx->alloc(sizeof(z32) * std::max<int64_t>(1, h->first_cap));
    z32* uv2 = nullptr;
    std::vector<int> hconv(nd + nl);
    try {
        prof_begin(&h->lane->pl, s, C_HACA);
        launch_aca(s, h->lane->s2, h->lane->ev_ready, h->lane->ev_panel, dm, kind, h->par, h->perm, nl,
                   wide_prefix(list), dlist, op.dblk, uv1, conv, o.aca_tolerance);
        prof_end(&h->lane->pl, s, C_HACA, 0.0, 0.0);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(op.blk.data(), op.dblk, sizeof(HBlk) * op.blk.size(), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(hconv.data(), conv, sizeof(int) * hconv.size(), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        // pass 2: blocks that hit the first capacity, with the full cap
        std::vector<int> redo;
        int64_t uoff2 = 0;
        for (int b = nd; b < nd + nl; ++b) {
            HBlk& x = op.blk[b];
            const int cap2 = std::min({x.m, x.n, o.max_rank});
            if (!hconv[b] && cap2 > x.rmax) {
                x.off = uoff2;
                x.rmax = cap2;
                x.rank = 0;
                uoff2 += (int64_t)x.rmax * (x.m + x.n);
                redo.push_back(b);
            }
        }
        std::stable_sort(redo.begin(), redo.end(), [&](int a, int b) {
            return (int64_t)op.blk[a].m * op.blk[a].n > (int64_t)op.blk[b].m * op.blk[b].n;
        });
        if (!redo.empty()) {
            uv2 = (z32*)ctx->alloc(sizeof(z32) * uoff2);
            HIPCHECK(hipMemcpyAsync(op.dblk, op.blk.data(), sizeof(HBlk) * op.blk.size(), hipMemcpyHostToDevice, s));
            int* dredo = hm_upload(h, redo);
            launch_aca(s, h->lane->s2, h->lane->ev_ready, h->lane->ev_panel, dm, kind, h->par, h->perm,
                       (int)redo.size(), wide_prefix(redo), dredo, op.dblk, uv2, conv, o.aca_tolerance);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipMemcpyAsync(op.blk.data(), op.dblk, sizeof(HBlk) * op.blk.size(), hipMemcpyDeviceToHost, s));
            HIPCHECK(hipMemcpyAsync(hconv.data(), conv, sizeof(int) * hconv.size(), hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
        }
        // compact factors: U (m x rank) then V (n x rank) per block
        std::vector<HBlk> compact = op.blk, o1, n1, o2, n2;
        std::vector<char> in2(nd + nl, 0);
        for (int b : redo) in2[b] = 1;
        int64_t coff = 0;
        int rmx = 0, unconv = 0;
        double rsum = 0;
        for (int b = nd; b < nd + nl; ++b) {
            HBlk& x = compact[b];
            x.off = coff;
            x.rmax = x.rank;
            coff += (int64_t)x.rank * (x.m + x.n);
            (in2[b] ? o2 : o1).push_back(op.blk[b]);
            (in2[b] ? n2 : n1).push_back(x);
            rmx = std::max(rmx, x.rank);
            rsum += x.rank;
            unconv += hconv[b] ? 0 : 1;
        }
        op.uv = (z32*)hm_alloc(h, sizeof(z32) * std::max<int64_t>(1, coff));
        HBlk* do1 = hm_upload(h, o1);
        HBlk* dn1 = hm_upload(h, n1);
        HBlk* do2 = hm_upload(h, o2);
        HBlk* dn2 = hm_upload(h, n2);
        launch_hcompact(s, (int)o1.size(), do1, dn1, uv1, op.uv);
        if (uv2) launch_hcompact(s, (int)o2.size(), do2, dn2, uv2, op.uv);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(s));
        ctx->release(uv1);
        ctx->release(uv2);
        uv1 = uv2 = nullptr;
        op.blk = compact;
        HIPCHECK(hipMemcpyAsync(op.dblk, op.blk.data(), sizeof(HBlk) * op.blk.size(), hipMemcpyHostToDevice, s));
        if (kind == HK_MULLER) launch_hunscale(s, nl, op.dblk, dlist, h->perm, (int)h->nt, op.uv);
        op.max_rank = rmx;
        op.unconverged = unconv;
        op.mean_rank = nl ? rsum / nl : 0.0;
        op.lowrank_bytes = sizeof(z32) * (uint64_t)coff;
        // Projection tiles: the low-rank blocks are grouped by column cluster
        // (same col_lo and n), their V rows stacked in block order and their
        // products T laid out in that order, so one tile of up to kHProjRows
        // stacked rows reads its X columns once for several blocks.
        std::map<std::pair<int, int>, std::vector<int>> groups;
        for (int b = nd; b < nd + nl; ++b)
            if (op.blk[b].rank > 0) groups[{op.blk[b].col_lo, op.blk[b].n}].push_back(b);
        std::vector<HTile> tiles;
        std::vector<HSplit> splits;
        int nparts = 0;
        std::vector<int64_t> vrow, toff(nd + nl, 0);
        int64_t tsum = 0;
        for (const auto& g : groups) {
            const int64_t g0 = tsum;
            for (int b : g.second) {
                const HBlk& x = op.blk[b];
                toff[b] = tsum * kHRhs;
                for (int l = 0; l < x.rank; ++l) vrow.push_back(x.off + (int64_t)x.rank * x.m + (int64_t)l * x.n);
                tsum += x.rank;
                op.v_entries += (int64_t)x.rank * x.n;
                op.lowrank_entries += (int64_t)x.rank * (x.m + x.n);
            }
            // clusters longer than kHProjSeg: segments of equal length (a
            // multiple of 32 columns), each tile's partials summed in order
            const int ncol = g.first.second;
            const int nseg = (ncol + kHProjSeg - 1) / kHProjSeg;
            const int seg = nseg > 1 ? ((ncol + nseg - 1) / nseg + 31) / 32 * 32 : ncol;
            for (int64_t r0 = g0; r0 < tsum; r0 += kHProjRows) {
                const int rows = (int)std::min<int64_t>(kHProjRows, tsum - r0);
                if (nseg == 1) {
                    tiles.push_back({r0, rows, g.first.first, ncol, 0, -1});
                    continue;
                }
                splits.push_back({r0, rows, nparts, 0});
                for (int v0 = 0; v0 < ncol; v0 += seg, ++nparts, ++splits.back().nseg)
                    tiles.push_back({r0, rows, g.first.first + v0, std::min(seg, ncol - v0), v0, nparts});
            }
        }
        // most work first (columns x rows of the tile), so the long tiles do
        // not trail the launch
        std::stable_sort(tiles.begin(), tiles.end(), [](const HTile& x, const HTile& y) {
            return (int64_t)x.n * x.rows > (int64_t)y.n * y.rows;
        });
        op.ntiles = (int)tiles.size();
        op.tiles = hm_upload(h, tiles);
        op.nsplit = (int)splits.size();
        if (op.nsplit) {
            op.splits = hm_upload(h, splits);
            op.P = (z64*)hm_alloc(h, sizeof(z64) * (size_t)nparts * kHProjRows * kHRhs);
        }
        op.vrow = hm_upload(h, vrow);
        op.toff = hm_upload(h, toff);
        op.T = (z64*)hm_alloc(h, sizeof(z64) * std::max<int64_t>(1, tsum * kHRhs));
    } catch (...) {
        hipStreamSynchronize(s);
        ctx->release(uv1);
        ctx->release(uv2);
        throw;
    }
    op.aca_ms = now_ms() - t2;
    // leaf row lists: dense blocks of the row, low-rank blocks covering it
    const int nleaf = h->nleaf;
    std::vector<std::vector<int>> dl(nleaf);
    std::vector<std::vector<HRef>> ll(nleaf);
    std::vector<int> diag(nleaf, -1);
    for (int b = 0; b < nd; ++b) {
        const int lr = h->leaf_of_pos[op.blk[b].row_lo], lc = h->leaf_of_pos[op.blk[b].col_lo];
        dl[lr].push_back(b);
        if (lr == lc) diag[lr] = b;
    }
    for (int b = nd; b < nd + nl; ++b) {
        const HBlk& x = op.blk[b];
        if (x.rank == 0) continue;
        for (int l = h->leaf_of_pos[x.row_lo]; l < nleaf && h->leaf_lo[l] < x.row_lo + x.m; ++l)
            ll[l].push_back({b, h->leaf_lo[l] - x.row_lo});
    }
    std::vector<int> dptr{0}, dlst, lptr{0};
    std::vector<HRef> llst;
    for (int l = 0; l < nleaf; ++l) {
        dlst.insert(dlst.end(), dl[l].begin(), dl[l].end());
        dptr.push_back((int)dlst.size());
        llst.insert(llst.end(), ll[l].begin(), ll[l].end());
        lptr.push_back((int)llst.size());
    }
    op.L.count = nleaf;
    op.L.lo = hm_upload(h, h->leaf_lo);
    op.L.m = hm_upload(h, h->leaf_m);
    op.L.diag = hm_upload(h, diag);
    op.L.dptr = hm_upload(h, dptr);
    op.L.dlist = hm_upload(h, dlst);
    op.L.lptr = hm_upload(h, lptr);
    op.L.llist = hm_upload(h, llst);
    // launch order of the row kernel: leaves by decreasing K chunks
    std::vector<int64_t> work(nleaf, 0);
    for (int l = 0; l < nleaf; ++l) {
        for (int b : dl[l]) work[l] += (op.blk[b].n + 31) / 32;
        for (const HRef& r : ll[l]) work[l] += (op.blk[r.blk].rank + 31) / 32;
    }
    std::vector<int> order(nleaf);
    for (int l = 0; l < nleaf; ++l) order[l] = l;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work[x] > work[y]; });
    op.L.order = hm_upload(h, order);
    HIPCHECK(hipStreamSynchronize(s));
}

void hm_coarse(bem_hmatrix h);

void hm_build_calderon(bem_hmatrix h);

// Nystrom meshes: the corrections of S and D (and of K', H when `deriv`) of
// each medium, from the naive defects over all node pairs; regular-wave
// basis under the dense path's rule (Mueller at low frequency), else point
// sources.
void hm_nys_corrections(bem_hmatrix h, bool deriv) {
    const bool regular = nys_regular(h->mesh->h, h->prob);
    const size_t nc = sizeof(z64) * 2 * (size_t)h->mesh->d.nt * h->mesh->d.nq;
    const NysWork w = nys_work([&](size_t b) { return hm_alloc(h, b); }, h->mesh, regular, deriv);
    for (int med = 0; med < (h->fluid ? 2 : 1); ++med) {
        h->nys_sd[med] = (z64*)hm_alloc(h, nc);
        if (deriv) h->nys_kh[med] = (z64*)hm_alloc(h, nc);
        nys_medium(h->lane->s, &h->lane->pl, h->mesh, w, med == 0 ? h->par.k : h->par.ki, regular, deriv,
                   h->nys_sd[med], h->nys_kh[med]);
    }
    HIPCHECK(hipGetLastError());
}

void hm_build(bem_hmatrix h) {
    const bool bm = h->prob.formulation == BEM_BURTON_MILLER;
    if (h->calderon) {
        hm_build_calderon(h);
        return;
    }
    hm_plan(h);
    h->nsys = h->n;
    h->sysperm = h->hperm;
    h->dsysperm = h->perm;
    if (h->mesh->d.nq) hm_nys_corrections(h, bm || h->fluid);
    if (h->fluid) {
        // Mueller: the right-hand side is the incident traces themselves
        hm_operator(h, h->A, HK_MULLER);
    } else {
        hm_operator(h, h->A, bm ? HK_BM : HK_CBIE);
        hm_operator(h, h->B, bm ? HK_BM_RHS : HK_CBIE_RHS);
    }
    if (h->opt.preconditioner >= 1) {
        h->minv = (z64*)hm_alloc(h, sizeof(z64) * 4096 * (size_t)h->nleaf);
        int* st = (int*)hm_alloc(h, sizeof(int));
        launch_fill_int(h->lane->s, st, 1, 0);
        launch_hjacobi_inv(h->lane->s, h->A.L, h->A.dblk, h->A.dval, h->minv, st);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(h->lane->s));
    }
    if (h->opt.preconditioner == 2) hm_coarse(h);
    const int64_t rbytes = 64ll * 1024 * 1024 / (32 * h->n);
    h->rcap = (int)std::min<int64_t>(std::min<int64_t>(h->opt.recycle_dimension, h->n), rbytes);
    if (h->rcap > 0) {
        h->ru = (z64*)hm_alloc(h, sizeof(z64) * h->n * h->rcap);
        h->rc = (z64*)hm_alloc(h, sizeof(z64) * h->n * h->rcap);
    }
    h->info.assembly_ms = h->A.assembly_ms + h->B.assembly_ms;
    h->info.aca_ms = h->A.aca_ms + h->B.aca_ms;
    h->info.max_rank = std::max(h->A.max_rank, h->B.max_rank);
    h->info.unconverged_blocks = h->A.unconverged + h->B.unconverged;
    h->info.mean_rank = h->A.mean_rank;
    h->info.lowrank_bytes = h->A.lowrank_bytes + h->B.lowrank_bytes;
    if (!h->fluid) h->info.dense_bytes *= 2;
}

// ---------------------------------------------------------------------------
// Low-frequency compressed Mueller (reference compression/fluid.jl:20-106 in
// the Calderon regime, max(k, k_int) * radius <= pi/2).  Compressed S and
// C = 1/2 I - D of each medium over one panel tree; with t_e = S_e q - D_e p
// and t_i = g S_i q - D_i p the system is
//   y_p = p + t_e - t_i
//   y_q = q + S_e^{-1} (D_e t_e + p/4) - S_i^{-1} (D_i t_i + p/4) / g,
// which is [I - D_e + D_i, S_e - g S_i; -H_e + H_i/g, I + K'_e - K'_i] with
// K' = S^{-1} D S and H = S^{-1} (D^2 - I/4) (full.jl:201-207), never formed.
// The reference applies S^{-1} through a hierarchical LU at a hundredth of
// the compression tolerance; here each S^{-1} is an inner GMRES on the
// compressed S (block Jacobi on its diagonal leaves) to a hundredth of the
// outer tolerance.  The outer preconditioner inverts the local 2m x 2m
// Mueller blocks of the diagonal leaves, built from the leaves' S and D
// blocks by the same identities (fluid.jl:57-72), plus the coarse space.
// ---------------------------------------------------------------------------
void hm_build_calderon(bem_hmatrix h) {
    hipStream_t s = h->lane->s;
    hm_plan(h);  // panel tree (h->calderon set): n = N
    const int64_t N = h->n;
    h->nsys = 2 * N;
    h->sysperm.resize(2 * N);
    for (int64_t pos = 0; pos < N; ++pos) {
        h->sysperm[pos] = h->hperm[pos];
        h->sysperm[N + pos] = (int)(N + h->hperm[pos]);
    }
    h->dsysperm = hm_upload(h, h->sysperm);
    if (h->mesh->d.nq) hm_nys_corrections(h, false);  // corrected S and D of both media
    const HPar keep = h->par;
    const double km[2] = {keep.k, keep.ki};
    for (int med = 0; med < 2; ++med) {
        h->par = HPar{km[med], km[med], 1.0};
        h->nys_med = med;
        hm_operator(h, h->cal[2 * med], HK_CBIE_RHS);  // S
        hm_operator(h, h->cal[2 * med + 1], HK_CBIE);  // 1/2 I - D
    }
    h->nys_med = 0;
    h->par = keep;
    // block Jacobi of S_e and S_i for the inner solves
    int* st = (int*)hm_alloc(h, sizeof(int));
    launch_fill_int(s, st, 1, 0);
    h->minv_e = (z64*)hm_alloc(h, sizeof(z64) * 4096 * (size_t)h->nleaf);
    h->minv_i = (z64*)hm_alloc(h, sizeof(z64) * 4096 * (size_t)h->nleaf);
    launch_hjacobi_inv(s, h->cal[0].L, h->cal[0].dblk, h->cal[0].dval, h->minv_e, st);
    launch_hjacobi_inv(s, h->cal[2].L, h->cal[2].dblk, h->cal[2].dval, h->minv_i, st);
    HIPCHECK(hipGetLastError());
    if (h->opt.preconditioner >= 1) {
        // local Mueller blocks of the diagonal leaves, inverted on the host
        using zc = std::complex<double>;
        const double g = keep.g;
        std::vector<int> diag(h->nleaf, -1);
        for (int b = 0; b < h->nd; ++b) {
            const HBlk& x = h->plan_blk[b];
            if (x.row_lo == x.col_lo && x.m == x.n) diag[h->leaf_of_pos[x.row_lo]] = b;
        }
        std::vector<std::vector<z32>> dv(4, std::vector<z32>(std::max<int64_t>(1, h->dense_count)));
        for (int q = 0; q < 4; ++q)
            HIPCHECK(hipMemcpyAsync(dv[q].data(), h->cal[q].dval, sizeof(z32) * h->dense_count,
                                    hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        std::vector<z64> lb((size_t)h->nleaf * 128 * 128, z64{0, 0});
        auto inv = [](std::vector<zc>& A, int n) {  // in-place Gauss-Jordan with partial pivoting
            std::vector<int> piv(n);
            for (int k = 0; k < n; ++k) {
                int p = k;
                for (int i = k + 1; i < n; ++i)
                    if (std::abs(A[(size_t)i * n + k]) > std::abs(A[(size_t)p * n + k])) p = i;
                if (!(std::abs(A[(size_t)p * n + k]) > 0)) return false;
                piv[k] = p;
                if (p != k)
                    for (int j = 0; j < n; ++j) std::swap(A[(size_t)p * n + j], A[(size_t)k * n + j]);
                const zc ip = 1.0 / A[(size_t)k * n + k];
                A[(size_t)k * n + k] = 1.0;
                for (int j = 0; j < n; ++j) A[(size_t)k * n + j] *= ip;
                for (int i = 0; i < n; ++i)
                    if (i != k) {
                        const zc f = A[(size_t)i * n + k];
                        if (f == 0.0) continue;
                        A[(size_t)i * n + k] = 0.0;
                        for (int j = 0; j < n; ++j) A[(size_t)i * n + j] -= f * A[(size_t)k * n + j];
                    }
            }
            for (int k = n - 1; k >= 0; --k)
                if (piv[k] != k)
                    for (int i = 0; i < n; ++i) std::swap(A[(size_t)i * n + k], A[(size_t)i * n + piv[k]]);
            return true;
        };
        auto mul = [](const std::vector<zc>& A, const std::vector<zc>& B, int n) {
            std::vector<zc> C((size_t)n * n, 0.0);
            for (int i = 0; i < n; ++i)
                for (int k = 0; k < n; ++k) {
                    const zc a = A[(size_t)i * n + k];
                    for (int j = 0; j < n; ++j) C[(size_t)i * n + j] += a * B[(size_t)k * n + j];
                }
            return C;
        };
        for (int l = 0; l < h->nleaf; ++l) {
            const int b = diag[l];
            if (b < 0) throw BemError(BEM_ERR_DEVICE, "leaf without a diagonal block");
            const int m = h->plan_blk[b].m;
            const int64_t off = h->plan_blk[b].off;
            std::vector<zc> se(m * m), de(m * m), si(m * m), di(m * m);
            for (int e = 0; e < m * m; ++e) {
                const int i = e / m, j = e % m;
                const double half = i == j ? 0.5 : 0.0;
                se[e] = zc(dv[0][off + e].x, dv[0][off + e].y);
                de[e] = zc(half - dv[1][off + e].x, -dv[1][off + e].y);
                si[e] = zc(dv[2][off + e].x, dv[2][off + e].y);
                di[e] = zc(half - dv[3][off + e].x, -dv[3][off + e].y);
            }
            auto sinv_e = se, sinv_i = si;
            if (!inv(sinv_e, m) || !inv(sinv_i, m)) throw BemError(BEM_ERR_NUMERIC, "singular local single layer");
            auto dde = mul(de, de, m), ddi = mul(di, di, m);
            for (int i = 0; i < m; ++i) {
                dde[(size_t)i * m + i] -= 0.25;
                ddi[(size_t)i * m + i] -= 0.25;
            }
            const auto ke = mul(sinv_e, mul(de, se, m), m), he = mul(sinv_e, dde, m);
            const auto ki = mul(sinv_i, mul(di, si, m), m), hi = mul(sinv_i, ddi, m);
            std::vector<zc> B((size_t)4 * m * m);
            const int n2 = 2 * m;
            for (int i = 0; i < m; ++i)
                for (int j = 0; j < m; ++j) {
                    const size_t e = (size_t)i * m + j;
                    const double d1 = i == j ? 1.0 : 0.0;
                    B[(size_t)i * n2 + j] = d1 - de[e] + di[e];
                    B[(size_t)i * n2 + m + j] = se[e] - g * si[e];
                    B[(size_t)(m + i) * n2 + j] = -he[e] + hi[e] / g;
                    B[(size_t)(m + i) * n2 + m + j] = d1 + ke[e] - ki[e];
                }
            if (!inv(B, n2)) throw BemError(BEM_ERR_NUMERIC, "singular local Mueller block");
            for (int i = 0; i < n2; ++i)
                for (int j = 0; j < n2; ++j) {
                    const zc v = B[(size_t)i * n2 + j];
                    lb[(size_t)l * 128 * 128 + (size_t)i * 128 + j] = z64{v.real(), v.imag()};
                }
        }
        h->lblk = hm_upload(h, lb);
    }
    if (h->opt.preconditioner == 2) {
        const HSys S = hm_system(h);
        z64* work = (z64*)hm_alloc(h, sizeof(z64) * std::max<size_t>(1, S.work_size(18)));
        hm_coarse_with(h, [&](const z64* X, int nr, z64* Y) { S.apply(X, nr, Y, work); },
                       [&](const z64* X, int nr, z64* Y) {
                           launch_hblock2_apply(s, h->cal[0].L, N, h->lblk, X, nr, Y);
                       });
    }
    const int64_t rbytes = 64ll * 1024 * 1024 / (32 * h->nsys);
    h->rcap = (int)std::min<int64_t>(std::min<int64_t>(h->opt.recycle_dimension, h->nsys), rbytes);
    if (h->rcap > 0) {
        h->ru = (z64*)hm_alloc(h, sizeof(z64) * h->nsys * h->rcap);
        h->rc = (z64*)hm_alloc(h, sizeof(z64) * h->nsys * h->rcap);
    }
    HIPCHECK(hipStreamSynchronize(s));
    h->info.assembly_ms = h->info.aca_ms = 0;
    h->info.max_rank = 0;
    h->info.unconverged_blocks = 0;
    h->info.lowrank_bytes = 0;
    for (const HOp& op : h->cal) {
        h->info.assembly_ms += op.assembly_ms;
        h->info.aca_ms += op.aca_ms;
        h->info.max_rank = std::max(h->info.max_rank, op.max_rank);
        h->info.unconverged_blocks += op.unconverged;
        h->info.lowrank_bytes += op.lowrank_bytes;
    }
    h->info.mean_rank = h->cal[0].mean_rank;
    h->info.dense_bytes *= 4;
    h->info.n = h->nsys;
}

// Inner solve S x = r of one medium (block Jacobi on S's diagonal leaves).
HSys hm_single_layer_system(bem_hmatrix h, int med) {
    HSys S;
    S.n = h->n;
    S.tol = std::max(1e-14, 0.01 * h->opt.gmres_tolerance);
    S.restart = 100;
    S.maxiter = 5000;
    HOp* op = &h->cal[2 * med];
    z64* minv = med == 0 ? h->minv_e : h->minv_i;
    S.apply = [h, op](const z64* X, int nr, z64* Y, z64*) { hm_apply(h, *op, X, nr, Y); };
    S.prec = [h, op, minv](const z64* X, int nr, z64* Y, z64*) {
        launch_hjacobi_apply(h->lane->s, op->L, minv, X, nr, Y);
    };
    S.work_size = [](int) { return (size_t)0; };
    return S;
}

HSys hm_calderon_system(bem_hmatrix h) {
    HSys S;
    S.n = h->nsys;
    S.tol = h->opt.gmres_tolerance;
    S.restart = std::max(1, h->opt.gmres_restart);
    S.maxiter = h->opt.gmres_maxiter;
    S.apply = [h](const z64* X, int nr, z64* Y, z64*) {
        hipStream_t s = h->lane->s;
        const int64_t N = h->n, vs = N * nr;
        const double g = h->par.g;
        std::vector<void*> tmp;
        auto al = [&]() { void* p = h->ctx->alloc(sizeof(z64) * vs); tmp.push_back(p); return (z64*)p; };
        try {
            z64 *A1 = al(), *A2 = al(), *te = al(), *ti = al(), *re = al(), *ri = al(), *de = al(), *di = al();
            const z64 *Xp = X, *Xq = X + vs;
            z64 *Yp = Y, *Yq = Y + vs;
            hm_apply(h, h->cal[0], Xq, nr, A1);
            hm_apply(h, h->cal[1], Xp, nr, A2);
            launch_lincomb3(s, vs, 1.0, A1, 1.0, A2, -0.5, Xp, te);  // S_e q - D_e p
            hm_apply(h, h->cal[2], Xq, nr, A1);
            hm_apply(h, h->cal[3], Xp, nr, A2);
            launch_lincomb3(s, vs, g, A1, 1.0, A2, -0.5, Xp, ti);    // g S_i q - D_i p
            launch_lincomb3(s, vs, 1.0, Xp, 1.0, te, -1.0, ti, Yp);
            hm_apply(h, h->cal[1], te, nr, A1);
            launch_lincomb3(s, vs, 0.5, te, -1.0, A1, 0.25, Xp, re);  // D_e t_e + p/4
            hm_apply(h, h->cal[3], ti, nr, A2);
            launch_lincomb3(s, vs, 0.5, ti, -1.0, A2, 0.25, Xp, ri);  // D_i t_i + p/4
            HIPCHECK(hipGetLastError());
            std::vector<bem_diagnostics> dgi(nr);
            const HSys Se = hm_single_layer_system(h, 0), Si = hm_single_layer_system(h, 1);
            hm_gmres(h, nr, re, de, dgi, nullptr, &Se);
            for (const auto& d : dgi) h->inner_iters += d.iterations;
            hm_gmres(h, nr, ri, di, dgi, nullptr, &Si);
            for (const auto& d : dgi) h->inner_iters += d.iterations;
            launch_lincomb3(s, vs, 1.0, Xq, 1.0, de, -1.0 / g, di, Yq);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipStreamSynchronize(s));
        } catch (...) {
            hipStreamSynchronize(s);
            for (void* p : tmp) h->ctx->release(p);
            throw;
        }
        for (void* p : tmp) h->ctx->release(p);
    };
    S.prec = [h](const z64* X, int nr, z64* Y, z64* work) {
        hipStream_t s = h->lane->s;
        if (h->lblk) launch_hblock2_apply(s, h->cal[0].L, h->n, h->lblk, X, nr, Y);
        else HIPCHECK(hipMemcpyAsync(Y, X, sizeof(z64) * h->nsys * nr, hipMemcpyDeviceToDevice, s));
        hm_coarse_correct(h, X, nr, Y, work);
    };
    S.work_size = [h](int nr) { return (size_t)(2 + 512) * std::max(1, h->nq) * nr; };
    return S;
}

// Y = op X on the device (tree-position order, nr <= kHRhs)
void hm_apply(bem_hmatrix h, HOp& op, const z64* X, int nr, z64* Y) {
    // algorithmic traffic: every block once, plus the input and output vectors
    const double bytes = (double)h->dense_count * sizeof(z32) + (double)op.lowrank_bytes +
                         2.0 * (double)h->n * nr * sizeof(z64);
    const double flops = 8.0 * nr * ((double)h->dense_count + (double)op.lowrank_bytes / sizeof(z32));
    // split of the work: projections read the V factors and X, the rows
    // kernel the dense leaves, the U factors and T
    const double pf = 8.0 * nr * (double)op.v_entries;
    const double pb = sizeof(z32) * (double)op.v_entries + (double)h->n * nr * sizeof(z64);
    prof_begin(&h->lane->pl, h->lane->s, C_HPROJ);
    launch_hproj(h->lane->s, op.ntiles, op.tiles, op.vrow, op.uv, X, nr, op.T, op.nsplit, op.splits, op.P);
    prof_end(&h->lane->pl, h->lane->s, C_HPROJ, pf, pb);
    prof_begin(&h->lane->pl, h->lane->s, C_HROWS);
    launch_hrows(h->lane->s, op.L, op.dblk, op.dval, op.uv, op.toff, op.T, X, nr, Y);
    prof_end(&h->lane->pl, h->lane->s, C_HROWS, flops - pf, bytes - pb);
}

// Y = P^{-1} X for the right preconditioner: identity, block Jacobi M^{-1},
// or the two-level form of the reference's coarse preconditioner
// (compression/fluid.jl:166-181,274-289):
//   P^{-1} x = M^{-1} x + (Q - M^{-1} A Q) E^{-1} Q^H x,  E = Q^H A Q.
// `work` holds 2 * nq * nr + kGmPartial * nq * nr complex doubles.
constexpr int kGmPartial = 512;
void hm_prec(bem_hmatrix h, const z64* X, int nr, z64* Y, z64* work) {
    hipStream_t s = h->lane->s;
    if (!h->minv) {
        HIPCHECK(hipMemcpyAsync(Y, X, sizeof(z64) * h->n * nr, hipMemcpyDeviceToDevice, s));
        return;
    }
    launch_hjacobi_apply(s, h->A.L, h->minv, X, nr, Y);
    hm_coarse_correct(h, X, nr, Y, work);
}

// Y += (Q - M^{-1} A Q) E^{-1} Q^H X when the coarse space is built
void hm_coarse_correct(bem_hmatrix h, const z64* X, int nr, z64* Y, z64* work) {
    if (h->nq <= 0) return;
    hipStream_t s = h->lane->s;
    z64* w = work;
    z64* w2 = work + (size_t)h->nq * nr;
    z64* part = work + 2 * (size_t)h->nq * nr;
    launch_basis_dots(s, h->nsys, nr, h->nq, h->cq, X, part, w);
    launch_small_apply(s, h->nq, nr, h->einv, w, w2);
    launch_basis_combine(s, h->nsys, nr, h->nq, h->cc, w2, Y);
}
size_t hm_prec_work(bem_hmatrix h, int nr) { return (size_t)(2 + kGmPartial) * std::max(1, h->nq) * nr; }


// Coarse space of the two-level preconditioner: the regular harmonic
// polynomials of degree <= 2 (the span of the reference's nine regular
// Helmholtz traces at k = 0, quadrature/fluid.jl:19-61, sampled at the panel
// centroids relative to their centre and radius) on each trace's unknowns,
// 9 columns for a rigid surface and 18 for a fluid body
// (compression/fluid.jl:193-205), orthonormalised.
void hm_coarse(bem_hmatrix h) {
    hm_coarse_with(h, [h](const z64* X, int nr, z64* Y) { hm_apply(h, h->A, X, nr, Y); },
                   [h](const z64* X, int nr, z64* Y) { launch_hjacobi_apply(h->lane->s, h->A.L, h->minv, X, nr, Y); });
}

// The coarse space over the system's unknowns (h->nsys rows; unknown u of
// row `pos` is h->sysperm[pos]: panel u % N, trace u / N), with the
// system's apply and local preconditioner.
void hm_coarse_with(bem_hmatrix h, const std::function<void(const z64*, int, z64*)>& apply,
                    const std::function<void(const z64*, int, z64*)>& local) {
    hipStream_t s = h->lane->s;
    const int64_t n = h->nsys;
    const int nt = (int)h->nt;
    const int nq = h->fluid ? 18 : 9;
    std::vector<double> c(3 * (size_t)nt);
    HIPCHECK(hipMemcpyAsync(c.data(), h->mesh->d.cen, sizeof(double) * 3 * nt, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    double ctr[3] = {0, 0, 0};
    for (int i = 0; i < nt; ++i)
        for (int d = 0; d < 3; ++d) ctr[d] += c[(size_t)d * nt + i] / nt;
    double rad = 0;
    for (int i = 0; i < nt; ++i) {
        double r2 = 0;
        for (int d = 0; d < 3; ++d) r2 += (c[(size_t)d * nt + i] - ctr[d]) * (c[(size_t)d * nt + i] - ctr[d]);
        rad = std::max(rad, std::sqrt(r2));
    }
    if (!(rad > 0)) rad = 1.0;
    std::vector<std::complex<double>> Q((size_t)n * nq, 0.0);
    for (int64_t pos = 0; pos < n; ++pos) {
        const int idx = h->sysperm[pos];
        const int panel = idx % nt;
        const int block = idx / nt;  // 0 pressure, 1 normal derivative (fluid)
        const double x = (c[panel] - ctr[0]) / rad, y = (c[(size_t)nt + panel] - ctr[1]) / rad,
                     z = (c[2 * (size_t)nt + panel] - ctr[2]) / rad;
        const double hv[9] = {1.0, x, y, z, x * y, y * z, x * z, x * x - y * y, 2 * z * z - x * x - y * y};
        for (int j = 0; j < 9; ++j) Q[(size_t)(9 * block + j) * n + pos] = hv[j];
    }
    // orthonormal columns: modified Gram-Schmidt, twice
    for (int j = 0; j < nq; ++j) {
        std::complex<double>* qj = &Q[(size_t)j * n];
        for (int pass = 0; pass < 2; ++pass)
            for (int i = 0; i < j; ++i) {
                const std::complex<double>* qi = &Q[(size_t)i * n];
                std::complex<double> d = 0;
                for (int64_t p = 0; p < n; ++p) d += std::conj(qi[p]) * qj[p];
                for (int64_t p = 0; p < n; ++p) qj[p] -= d * qi[p];
            }
        double nn = 0;
        for (int64_t p = 0; p < n; ++p) nn += std::norm(qj[p]);
        nn = std::sqrt(nn);
        if (!(nn > 0)) throw BemError(BEM_ERR_NUMERIC, "degenerate coarse space");
        for (int64_t p = 0; p < n; ++p) qj[p] /= nn;
    }
    // A Q and M^{-1} A Q in the batched vector layout [pos * nq + j]
    std::vector<z64> X((size_t)n * nq);
    for (int64_t p = 0; p < n; ++p)
        for (int j = 0; j < nq; ++j) X[p * nq + j] = z64{Q[(size_t)j * n + p].real(), Q[(size_t)j * n + p].imag()};
    z64* dX = (z64*)h->ctx->alloc(sizeof(z64) * n * nq);
    z64* dY = nullptr;
    z64* dZ = nullptr;
    try {
        dY = (z64*)h->ctx->alloc(sizeof(z64) * n * nq);
        dZ = (z64*)h->ctx->alloc(sizeof(z64) * n * nq);
        HIPCHECK(hipMemcpyAsync(dX, X.data(), sizeof(z64) * n * nq, hipMemcpyHostToDevice, s));
        apply(dX, nq, dY);
        local(dY, nq, dZ);
        std::vector<z64> AQ((size_t)n * nq), MAQ((size_t)n * nq);
        HIPCHECK(hipMemcpyAsync(AQ.data(), dY, sizeof(z64) * n * nq, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(MAQ.data(), dZ, sizeof(z64) * n * nq, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        // E = Q^H A Q and its inverse (Gauss-Jordan with partial pivoting)
        std::vector<std::complex<double>> E((size_t)nq * nq, 0.0), Ei((size_t)nq * nq, 0.0);
        for (int i = 0; i < nq; ++i)
            for (int j = 0; j < nq; ++j) {
                std::complex<double> d = 0;
                for (int64_t p = 0; p < n; ++p)
                    d += std::conj(Q[(size_t)i * n + p]) * std::complex<double>(AQ[p * nq + j].x, AQ[p * nq + j].y);
                E[(size_t)i * nq + j] = d;
            }
        for (int i = 0; i < nq; ++i) Ei[(size_t)i * nq + i] = 1.0;
        for (int k = 0; k < nq; ++k) {
            int pr = k;
            for (int i = k + 1; i < nq; ++i)
                if (std::abs(E[(size_t)i * nq + k]) > std::abs(E[(size_t)pr * nq + k])) pr = i;
            if (!(std::abs(E[(size_t)pr * nq + k]) > 0)) throw BemError(BEM_ERR_NUMERIC, "singular coarse operator");
            if (pr != k)
                for (int j = 0; j < nq; ++j) {
                    std::swap(E[(size_t)pr * nq + j], E[(size_t)k * nq + j]);
                    std::swap(Ei[(size_t)pr * nq + j], Ei[(size_t)k * nq + j]);
                }
            const std::complex<double> ip = 1.0 / E[(size_t)k * nq + k];
            for (int j = 0; j < nq; ++j) {
                E[(size_t)k * nq + j] *= ip;
                Ei[(size_t)k * nq + j] *= ip;
            }
            for (int i = 0; i < nq; ++i)
                if (i != k) {
                    const std::complex<double> f = E[(size_t)i * nq + k];
                    if (f == 0.0) continue;
                    for (int j = 0; j < nq; ++j) {
                        E[(size_t)i * nq + j] -= f * E[(size_t)k * nq + j];
                        Ei[(size_t)i * nq + j] -= f * Ei[(size_t)k * nq + j];
                    }
                }
        }
        std::vector<z64> hq((size_t)n * nq), hc((size_t)n * nq), he((size_t)nq * nq);
        for (int j = 0; j < nq; ++j)
            for (int64_t p = 0; p < n; ++p) {
                const std::complex<double> q = Q[(size_t)j * n + p];
                hq[(size_t)j * n + p] = z64{q.real(), q.imag()};
                hc[(size_t)j * n + p] = z64{q.real() - MAQ[p * nq + j].x, q.imag() - MAQ[p * nq + j].y};
            }
        for (size_t e = 0; e < he.size(); ++e) he[e] = z64{Ei[e].real(), Ei[e].imag()};
        h->cq = hm_upload(h, hq);
        h->cc = hm_upload(h, hc);
        h->einv = hm_upload(h, he);
        h->nq = nq;
        HIPCHECK(hipStreamSynchronize(s));
    } catch (...) {
        hipStreamSynchronize(s);
        h->ctx->release(dX); h->ctx->release(dY); h->ctx->release(dZ);
        throw;
    }
    h->ctx->release(dX); h->ctx->release(dY); h->ctx->release(dZ);
}

__global__ void k_inv_norm(int nr, const z64* sq, double* scale, double* norm) {
    const int c = threadIdx.x;
    if (c >= nr) return;
    const double v = sqrt(fmax(sq[c].x, 0.0));
    norm[c] = v;
    scale[c] = v > 0 ? 1.0 / v : 0.0;
}

// Restarted GMRES (right-preconditioned, CGS2 Arnoldi, complex Givens) for nr
// right-hand sides B at once, all in FP64.  X receives the solution.
HSys hm_calderon_system(bem_hmatrix h);

// The hierarchical matrix's own system (A with its preconditioner).
HSys hm_system(bem_hmatrix h) {
    if (h->calderon) return hm_calderon_system(h);
    HSys S;
    S.n = h->n;
    S.tol = h->opt.gmres_tolerance;
    S.restart = std::max(1, h->opt.gmres_restart);
    S.maxiter = h->opt.gmres_maxiter;
    S.apply = [h](const z64* X, int nr, z64* Y, z64*) { hm_apply(h, h->A, X, nr, Y); };
    S.prec = [h](const z64* X, int nr, z64* Y, z64* w) { hm_prec(h, X, nr, Y, w); };
    S.work_size = [h](int nr) { return hm_prec_work(h, nr); };
    return S;
}

// target: absolute residual target per column (null: tolerance * ||B||);
// sys: the system (null: the hierarchical matrix's own).
void hm_gmres(bem_hmatrix h, int nr, const z64* B, z64* X, std::vector<bem_diagnostics>& dg, const double* target,
              const HSys* sys) {
    const HSys own = sys ? HSys{} : hm_system(h);
    const HSys& S = sys ? *sys : own;
    hipStream_t s = h->lane->s;
    const int64_t n = S.n;
    const int m = std::max(1, S.restart);
    const double tol = S.tol;
    const int64_t vs = n * nr;
    z64* Vb = (z64*)h->ctx->alloc(sizeof(z64) * vs * (m + 1));
    std::vector<void*> tmp{Vb};
    auto al = [&](size_t b) { void* p = h->ctx->alloc(b); tmp.push_back(p); return p; };
    try {
        z64* U = (z64*)al(sizeof(z64) * vs);
        z64* W = (z64*)al(sizeof(z64) * vs);
        z64* Z = (z64*)al(sizeof(z64) * vs);
        z64* R = (z64*)al(sizeof(z64) * vs);
        z64* part = (z64*)al(sizeof(z64) * 512 * (m + 1) * nr);
        z64* h1 = (z64*)al(sizeof(z64) * (m + 1) * nr);
        z64* h2 = (z64*)al(sizeof(z64) * (m + 1) * nr);
        z64* sq = (z64*)al(sizeof(z64) * nr);
        double* scale = (double*)al(sizeof(double) * nr);
        double* norm = (double*)al(sizeof(double) * nr);
        z64* ycoef = (z64*)al(sizeof(z64) * m * nr);
        z64* pw = (z64*)al(sizeof(z64) * std::max<size_t>(1, S.work_size(nr)));
        launch_fill_double(s, (double*)U, 2 * vs, 0.0);
        launch_fill_double(s, (double*)X, 2 * vs, 0.0);
        // ||b||
        std::vector<double> bn(nr), rn(nr);
        launch_gm_dots(s, n, nr, 1, B, vs, B, part, sq);
        std::vector<z64> hs(nr);
        HIPCHECK(hipMemcpyAsync(hs.data(), sq, sizeof(z64) * nr, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        for (int c = 0; c < nr; ++c) bn[c] = std::sqrt(std::max(hs[c].x, 0.0));
        std::vector<double> thr(nr);
        for (int c = 0; c < nr; ++c) thr[c] = target ? target[c] : tol * bn[c];
        HIPCHECK(hipMemcpyAsync(R, B, sizeof(z64) * vs, hipMemcpyDeviceToDevice, s));
        int iters = 0;
        bool done = false;
        std::vector<std::complex<double>> H((size_t)(m + 1) * m * nr), g((size_t)(m + 1) * nr), sn((size_t)m * nr);
        std::vector<double> cs((size_t)m * nr);
        std::vector<z64> hh1((m + 1) * nr), hh2((m + 1) * nr);
        for (;;) {
            // residual norms of this cycle's start
            launch_gm_dots(s, n, nr, 1, R, vs, R, part, sq);
            hipLaunchKernelGGL(k_inv_norm, dim3(1), dim3(64), 0, s, nr, sq, scale, norm);
            HIPCHECK(hipMemcpyAsync(rn.data(), norm, sizeof(double) * nr, hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            bool all = true;
            for (int c = 0; c < nr; ++c) all &= !(rn[c] > thr[c]);
            if (all || iters >= S.maxiter) {
                done = all;
                break;
            }
            launch_gm_scale(s, n, nr, R, scale, Vb);
            std::fill(g.begin(), g.end(), std::complex<double>(0));
            for (int c = 0; c < nr; ++c) g[(size_t)c * (m + 1)] = rn[c];
            int jend = 0;
            for (int j = 0; j < m; ++j) {
                z64* Vj = Vb + (int64_t)j * vs;
                S.prec(Vj, nr, Z, pw);
                S.apply(Z, nr, W, pw);
                // classical Gram-Schmidt, twice
                prof_begin(&h->lane->pl, s, C_HORTHO);
                launch_gm_dots(s, n, nr, j + 1, Vb, vs, W, part, h1);
                launch_gm_combine(s, n, nr, j + 1, Vb, vs, h1, -1.0, W);
                launch_gm_dots(s, n, nr, j + 1, Vb, vs, W, part, h2);
                launch_gm_combine(s, n, nr, j + 1, Vb, vs, h2, -1.0, W);
                launch_gm_dots(s, n, nr, 1, W, vs, W, part, sq);
                hipLaunchKernelGGL(k_inv_norm, dim3(1), dim3(64), 0, s, nr, sq, scale, norm);
                launch_gm_scale(s, n, nr, W, scale, Vb + (int64_t)(j + 1) * vs);
                prof_end(&h->lane->pl, s, C_HORTHO, 16.0 * 4 * (j + 1) * (double)vs,
                         (4.0 * (j + 1) + 6.0) * (double)vs * sizeof(z64));
                HIPCHECK(hipMemcpyAsync(hh1.data(), h1, sizeof(z64) * (j + 1) * nr, hipMemcpyDeviceToHost, s));
                HIPCHECK(hipMemcpyAsync(hh2.data(), h2, sizeof(z64) * (j + 1) * nr, hipMemcpyDeviceToHost, s));
                HIPCHECK(hipMemcpyAsync(rn.data(), norm, sizeof(double) * nr, hipMemcpyDeviceToHost, s));
                HIPCHECK(hipGetLastError());
                HIPCHECK(hipStreamSynchronize(s));
                ++iters;
                jend = j + 1;
                bool conv = true;
                for (int c = 0; c < nr; ++c) {
                    auto col = [&](int i) -> std::complex<double>& { return H[((size_t)c * (m + 1) + i) * m + j]; };
                    for (int i = 0; i <= j; ++i)
                        col(i) = std::complex<double>(hh1[i * nr + c].x + hh2[i * nr + c].x,
                                                      hh1[i * nr + c].y + hh2[i * nr + c].y);
                    col(j + 1) = rn[c];
                    for (int i = 0; i < j; ++i) {  // earlier rotations
                        const std::complex<double> a = col(i), b = col(i + 1);
                        col(i) = cs[(size_t)c * m + i] * a + sn[(size_t)c * m + i] * b;
                        col(i + 1) = -std::conj(sn[(size_t)c * m + i]) * a + cs[(size_t)c * m + i] * b;
                    }
                    const std::complex<double> a = col(j), b = col(j + 1);
                    const double na = std::abs(a), rho = std::hypot(na, std::abs(b));
                    double cc;
                    std::complex<double> ss;
                    if (rho == 0) {
                        cc = 1;
                        ss = 0;
                    } else if (na == 0) {
                        cc = 0;
                        ss = std::conj(b) / std::abs(b);
                        col(j) = std::abs(b);
                    } else {
                        cc = na / rho;
                        ss = (a / na) * std::conj(b) / rho;
                        col(j) = (a / na) * rho;
                    }
                    col(j + 1) = 0;
                    cs[(size_t)c * m + j] = cc;
                    sn[(size_t)c * m + j] = ss;
                    std::complex<double>& gj = g[(size_t)c * (m + 1) + j];
                    g[(size_t)c * (m + 1) + j + 1] = -std::conj(ss) * gj;
                    gj = cc * gj;
                    conv &= !(std::abs(g[(size_t)c * (m + 1) + j + 1]) > thr[c]);
                }
                if (conv || iters >= S.maxiter) break;
            }
            // y = H^{-1} g per column; U += V y; X = M^{-1} U; R = B - A X
            std::vector<z64> y((size_t)jend * nr);
            for (int c = 0; c < nr; ++c) {
                std::vector<std::complex<double>> yc(jend);
                for (int i = jend - 1; i >= 0; --i) {
                    std::complex<double> v = g[(size_t)c * (m + 1) + i];
                    for (int q = i + 1; q < jend; ++q) v -= H[((size_t)c * (m + 1) + i) * m + q] * yc[q];
                    const std::complex<double> d = H[((size_t)c * (m + 1) + i) * m + i];
                    yc[i] = d == 0.0 ? 0.0 : v / d;
                }
                for (int i = 0; i < jend; ++i) y[(size_t)i * nr + c] = z64{yc[i].real(), yc[i].imag()};
            }
            HIPCHECK(hipMemcpyAsync(ycoef, y.data(), sizeof(z64) * y.size(), hipMemcpyHostToDevice, s));
            launch_gm_combine(s, n, nr, jend, Vb, vs, ycoef, 1.0, U);
            S.prec(U, nr, X, pw);
            S.apply(X, nr, W, pw);
            launch_gm_sub(s, vs, B, W, R);
            HIPCHECK(hipGetLastError());
        }
        for (int c = 0; c < nr; ++c) {
            bem_diagnostics& d = dg[c];
            std::memset(&d, 0, sizeof d);
            d.unknown_count = n;
            d.quadrature_nodes = n;
            d.absolute_residual = rn[c];
            d.relative_residual = bn[c] > 0 ? rn[c] / bn[c] : rn[c];
            d.iterations = iters;
            d.method = 2;
            d.factor_precision = 1;
            d.converged = !(rn[c] > thr[c]) ? 1 : 0;
        }
        (void)done;
    } catch (...) {
        hipStreamSynchronize(s);
        for (void* p : tmp) h->ctx->release(p);
        throw;
    }
    hipStreamSynchronize(s);
    for (void* p : tmp) h->ctx->release(p);
}

// ---------------------------------------------------------------------------
// GMRES warm-started from recycled solutions (reference compression/fluid.jl:
// 303-419).  The hierarchical matrix keeps up to rcap accepted solutions U
// with their images C = A U, orthonormalised in image space (C^H C = I, the
// same transformations applied to U), so x0 = U C^H b minimises the residual
// over the kept space.  A batch of right-hand sides starts from x0 (or from
// zero when the space is empty or does not reduce every residual), GMRES
// solves A d = b - A x0 to the original targets tolerance * ||b||, the true
// residual of x0 + d is checked, and a recycled attempt that misses it is
// repeated from zero with the space cleared.  Accepted solutions enter the
// space while it has room.
// ---------------------------------------------------------------------------
void hm_norms(bem_hmatrix h, int nr, const z64* V, z64* part, z64* sq, std::vector<double>& out) {
    hipStream_t s = h->lane->s;
    const int64_t vs = h->nsys * nr;
    launch_gm_dots(s, h->nsys, nr, 1, V, vs, V, part, sq);
    std::vector<z64> hs(nr);
    HIPCHECK(hipMemcpyAsync(hs.data(), sq, sizeof(z64) * nr, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    out.resize(nr);
    for (int c = 0; c < nr; ++c) out[c] = std::sqrt(std::max(hs[c].x, 0.0));
}

void hm_solve_recycled(bem_hmatrix h, int nr, const z64* B, z64* X, std::vector<bem_diagnostics>& dg) {
    if (h->rcap == 0) {
        hm_gmres(h, nr, B, X, dg);
        return;
    }
    hipStream_t s = h->lane->s;
    const HSys S = hm_system(h);
    const int64_t n = S.n, vs = n * nr;
    const double tol = h->opt.gmres_tolerance;
    std::vector<void*> tmp;
    auto al = [&](size_t b) { void* p = h->ctx->alloc(b); tmp.push_back(p); return p; };
    try {
        z64* R0 = (z64*)al(sizeof(z64) * vs);
        z64* W = (z64*)al(sizeof(z64) * vs);
        z64* pw = (z64*)al(sizeof(z64) * std::max<size_t>(1, S.work_size(nr)));
        z64* D = (z64*)al(sizeof(z64) * vs);
        z64* part = (z64*)al(sizeof(z64) * 512 * (size_t)std::max(1, h->rcap) * nr);
        z64* coef = (z64*)al(sizeof(z64) * (size_t)std::max(1, h->rcap) * nr);
        z64* ones = (z64*)al(sizeof(z64) * nr);
        z64* sq = (z64*)al(sizeof(z64) * nr);
        std::vector<z64> hones(nr, z64{1.0, 0.0});
        HIPCHECK(hipMemcpyAsync(ones, hones.data(), sizeof(z64) * nr, hipMemcpyHostToDevice, s));
        std::vector<double> bn, rn, tgt(nr);
        hm_norms(h, nr, B, part, sq, bn);
        for (int c = 0; c < nr; ++c) tgt[c] = tol * bn[c];
        launch_fill_double(s, (double*)X, 2 * vs, 0.0);
        auto report = [&](const std::vector<double>& r, int extra_iters) {
            for (int c = 0; c < nr; ++c) {
                bem_diagnostics& d = dg[c];
                const int it = d.iterations + extra_iters;
                std::memset(&d, 0, sizeof d);
                d.unknown_count = n;
                d.quadrature_nodes = n;
                d.absolute_residual = r[c];
                d.relative_residual = bn[c] > 0 ? r[c] / bn[c] : r[c];
                d.iterations = it;
                d.method = 2;
                d.factor_precision = 1;
                d.converged = !(r[c] > tgt[c]) ? 1 : 0;
            }
        };
        bool reused = false;
        if (h->rcount > 0) {
            // x0 = U C^H b and its residual
            launch_basis_dots(s, n, nr, h->rcount, h->rc, B, part, coef);
            launch_basis_combine(s, n, nr, h->rcount, h->ru, coef, X);
            S.apply(X, nr, W, pw);
            launch_gm_sub(s, vs, B, W, R0);
            hm_norms(h, nr, R0, part, sq, rn);
            bool all = true, shrinks = true;
            for (int c = 0; c < nr; ++c) {
                all &= !(rn[c] > tgt[c]);
                shrinks &= rn[c] < bn[c] || bn[c] == 0.0;
            }
            if (all) {
                for (auto& d : dg) std::memset(&d, 0, sizeof d);
                report(rn, 0);
                for (void* p : tmp) h->ctx->release(p);
                return;
            }
            if (shrinks) {
                reused = true;
            } else {
                launch_fill_double(s, (double*)X, 2 * vs, 0.0);  // the space does not fit: start over
                h->rcount = 0;
            }
        }
        if (reused) {
            hm_gmres(h, nr, R0, D, dg, tgt.data(), &S);
            launch_gm_combine(s, n, nr, 1, D, vs, ones, 1.0, X);  // x = x0 + d
        } else {
            hm_gmres(h, nr, B, X, dg, tgt.data(), &S);
        }
        // true residual of the returned solution
        S.apply(X, nr, W, pw);
        launch_gm_sub(s, vs, B, W, R0);
        hm_norms(h, nr, R0, part, sq, rn);
        bool accepted = true;
        for (int c = 0; c < nr; ++c) accepted &= !(rn[c] > tgt[c]);
        if (reused && !accepted) {
            // a recycled attempt that misses the target is repeated from zero;
            // its iterations count, and it must not seed later solves
            std::vector<int> first(nr);
            for (int c = 0; c < nr; ++c) first[c] = dg[c].iterations;
            h->rcount = 0;
            hm_gmres(h, nr, B, X, dg, tgt.data(), &S);
            for (int c = 0; c < nr; ++c) dg[c].iterations += first[c];
            S.apply(X, nr, W, pw);
            launch_gm_sub(s, vs, B, W, R0);
            hm_norms(h, nr, R0, part, sq, rn);
            accepted = true;
            for (int c = 0; c < nr; ++c) accepted &= !(rn[c] > tgt[c]);
        }
        report(rn, 0);
        // accepted solutions enter the space while it has room: images
        // orthonormalised, the same combination applied to the solutions
        for (int c = 0; c < nr && h->rcount < h->rcap; ++c) {
            if (rn[c] > tgt[c] || !(bn[c] > 0)) continue;
            std::vector<z64> xu(n), xc(n);
            // D and R0 hold n * nr >= n entries each; neither is needed any more
            launch_gather_col(s, n, nr, c, X, D);
            HIPCHECK(hipMemcpyAsync(xu.data(), D, sizeof(z64) * n, hipMemcpyDeviceToHost, s));
            launch_gather_col(s, n, nr, c, W, R0);
            HIPCHECK(hipMemcpyAsync(xc.data(), R0, sizeof(z64) * n, hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            std::vector<std::complex<double>> u(n), cimg(n);
            for (int64_t p = 0; p < n; ++p) {
                u[p] = {xu[p].x, xu[p].y};
                cimg[p] = {xc[p].x, xc[p].y};
            }
            double orig = 0;
            for (const auto& v : cimg) orig += std::norm(v);
            orig = std::sqrt(orig);
            for (int pass = 0; pass < 2; ++pass)
                for (int j = 0; j < h->rcount; ++j) {
                    const std::complex<double>* cj = &h->hrc[(size_t)j * n];
                    const std::complex<double>* uj = &h->hru[(size_t)j * n];
                    std::complex<double> d = 0;
                    for (int64_t p = 0; p < n; ++p) d += std::conj(cj[p]) * cimg[p];
                    for (int64_t p = 0; p < n; ++p) {
                        cimg[p] -= d * cj[p];
                        u[p] -= d * uj[p];
                    }
                }
            double nc = 0;
            for (const auto& v : cimg) nc += std::norm(v);
            nc = std::sqrt(nc);
            if (!(std::isfinite(nc) && nc > std::sqrt(2.220446049250313e-16) * orig)) continue;
            const int j = h->rcount;
            h->hru.resize((size_t)(j + 1) * n);
            h->hrc.resize((size_t)(j + 1) * n);
            for (int64_t p = 0; p < n; ++p) {
                h->hru[(size_t)j * n + p] = u[p] / nc;
                h->hrc[(size_t)j * n + p] = cimg[p] / nc;
                xu[p] = z64{(u[p] / nc).real(), (u[p] / nc).imag()};
                xc[p] = z64{(cimg[p] / nc).real(), (cimg[p] / nc).imag()};
            }
            HIPCHECK(hipMemcpyAsync(h->ru + (size_t)j * n, xu.data(), sizeof(z64) * n, hipMemcpyHostToDevice, s));
            HIPCHECK(hipMemcpyAsync(h->rc + (size_t)j * n, xc.data(), sizeof(z64) * n, hipMemcpyHostToDevice, s));
            HIPCHECK(hipStreamSynchronize(s));
            h->rcount = j + 1;
        }
        if (h->rcount == 0) {
            h->hru.clear();
            h->hrc.clear();
        }
        for (void* p : tmp) h->ctx->release(p);
    } catch (...) {
        hipStreamSynchronize(s);
        for (void* p : tmp) h->ctx->release(p);
        throw;
    }
}

// ---------------------------------------------------------------------------
// Reduced-basis frequency sweep (reference src/postprocessing/sweeps.jl:
// 113-306, FrequencyReduction): dense FP64 CBIE fluid equations per
// frequency; training solves fill a bounded orthonormal snapshot space, which
// is frozen; each query assembles its own matrix, projects onto the basis by
// least squares on the row-scaled images A V / r (pivoted QR), accepts when
// both the original and the row-scaled relative residuals meet the
// tolerance, and otherwise falls back to the full equilibrated LU solve, which
// must meet the same residual.  Assembly, images, residuals, LU and far field
// run on the device; the m <= capacity column QR and the snapshot
// orthogonalisation run on the host.
// ---------------------------------------------------------------------------
using zc = std::complex<double>;

// Householder QR with column pivoting (Businger and Golub 1965; largest
// remaining column norm first) of W (n x m, column-major), then the
// least-squares solution of W y = c when the numerical rank is m.  Returns the
// rank: count of |R_jj| > rank_tol * max |R_jj|.
int pivoted_qr_solve(std::vector<zc>& W, int n, int m, std::vector<zc> c, double rank_tol, std::vector<zc>& y) {
    std::vector<int> piv(m);
    for (int j = 0; j < m; ++j) piv[j] = j;
    std::vector<double> rdiag(m, 0.0);
    for (int j = 0; j < m; ++j) {
        // pivot: the remaining column with the largest norm below row j
        int best = j;
        double bn = -1.0;
        for (int q = j; q < m; ++q) {
            double v = 0.0;
            for (int i = j; i < n; ++i) v += std::norm(W[(size_t)q * n + i]);
            if (v > bn) { bn = v; best = q; }
        }
        if (best != j) {
            for (int i = 0; i < n; ++i) std::swap(W[(size_t)j * n + i], W[(size_t)best * n + i]);
            std::swap(piv[j], piv[best]);
        }
        zc* a = &W[(size_t)j * n];
        double xn = 0.0;
        for (int i = j; i < n; ++i) xn += std::norm(a[i]);
        xn = std::sqrt(xn);
        if (xn == 0.0) { rdiag[j] = 0.0; continue; }
        const zc alpha = a[j] == zc(0) ? zc(-xn) : -std::exp(zc(0, std::arg(a[j]))) * xn;  // R_jj
        // v = a - alpha e_j, H = I - 2 v v^H / (v^H v)
        a[j] -= alpha;
        double vv = 0.0;
        for (int i = j; i < n; ++i) vv += std::norm(a[i]);
        for (int q = j + 1; q < m; ++q) {
            zc* b = &W[(size_t)q * n];
            zc d = 0.0;
            for (int i = j; i < n; ++i) d += std::conj(a[i]) * b[i];
            const zc f = 2.0 * d / vv;
            for (int i = j; i < n; ++i) b[i] -= f * a[i];
        }
        zc d = 0.0;
        for (int i = j; i < n; ++i) d += std::conj(a[i]) * c[i];
        const zc f = 2.0 * d / vv;
        for (int i = j; i < n; ++i) c[i] -= f * a[i];
        rdiag[j] = std::abs(alpha);
        a[j] = alpha;  // R_jj; rows above j of later columns hold R
    }
    double mx = 0.0;
    for (double v : rdiag) mx = std::max(mx, v);
    int rank = 0;
    for (double v : rdiag)
        if (v > rank_tol * mx) ++rank;
    if (rank < m) return rank;
    // back substitution R z = (Q^H c)[0:m], then y[piv[j]] = z[j]
    std::vector<zc> z(m);
    for (int j = m - 1; j >= 0; --j) {
        zc acc = c[j];
        for (int q = j + 1; q < m; ++q) acc -= W[(size_t)q * n + j] * z[q];
        z[j] = acc / W[(size_t)j * n + j];
    }
    y.assign(m, 0.0);
    for (int j = 0; j < m; ++j) y[piv[j]] = z[j];
    return rank;
}

// Snapshot space of the training solves (sweeps.jl:115-147): two passes of
// Gram-Schmidt against the kept columns; kept when the remainder of the
// normalised snapshot exceeds rank_tol.
bool add_snapshot(std::vector<std::vector<zc>>& V, int capacity, const std::vector<zc>& x, double rank_tol) {
    if ((int)V.size() >= capacity) return false;
    double mag = 0.0;
    for (const zc& v : x) mag += std::norm(v);
    mag = std::sqrt(mag);
    if (!std::isfinite(mag)) throw BemError(BEM_ERR_ARGUMENT, "snapshot norm overflow");
    if (mag == 0.0) return false;
    std::vector<zc> v(x.size());
    for (size_t i = 0; i < x.size(); ++i) v[i] = x[i] / mag;
    for (int pass = 0; pass < 2; ++pass)
        for (const auto& q : V) {
            zc d = 0.0;
            for (size_t i = 0; i < v.size(); ++i) d += std::conj(q[i]) * v[i];
            for (size_t i = 0; i < v.size(); ++i) v[i] -= d * q[i];
        }
    double rem = 0.0;
    for (const zc& e : v) rem += std::norm(e);
    rem = std::sqrt(rem);
    if (!(rem > rank_tol)) return false;
    for (zc& e : v) e /= rem;
    V.push_back(std::move(v));
    return true;
}

double rel_or_inf(double r, double b) { return b == 0.0 ? (r == 0.0 ? 0.0 : INFINITY) : r / b; }

// One equation of the reduced sweep at wavenumber prob.k: returns the solution
// x (host) and fills the report and the backscatter amplitude.
std::vector<zc> reduced_equation(bem_ctx ctx, bem_mesh mesh, const bem_problem& prob, const double* ddir,
                                 const double* dback, const std::vector<std::vector<zc>>& V, double tol,
                                 double rank_tol, bem_reduced_report& rep, double* amp) {
    Lane* L = ctx->lane0();
    hipStream_t s = L->s;
    bem_operator op = nullptr;
    operator_create(ctx, L, mesh, &prob, &op, false, true);
    std::vector<void*> tmp;
    try {
        const int n = op->n, nt = op->nt;
        const int m = (int)V.size();
        build_rhs(op, ddir);
        double* rw = (double*)op_alloc(op, sizeof(double) * n);
        launch_row_stats<double>(s, (const z64*)op->A, n, n, rw, op->vden);
        launch_fix_scale(s, n, rw);
        std::vector<double> hr(n);
        std::vector<zc> hb(n), x(n, 0.0);
        HIPCHECK(hipMemcpyAsync(hr.data(), rw, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(hb.data(), op->vb, sizeof(z64) * n, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        for (double v : hr)
            if (!std::isfinite(v)) throw BemError(BEM_ERR_ARGUMENT, "row norm overflow");
        double bnorm = 0.0;
        for (const zc& v : hb) bnorm += std::norm(v);
        rep.reason = m == 0 ? BEM_REDUCED_EMPTY : BEM_REDUCED_RESIDUAL;
        rep.rank = 0;
        if (bnorm != 0.0 && m > 0) {
            // images A V / rows, recomputed at this frequency
            z64* dV = (z64*)ctx->alloc(sizeof(z64) * (size_t)n * m);
            tmp.push_back(dV);
            z64* dW = (z64*)ctx->alloc(sizeof(z64) * (size_t)n * m);
            tmp.push_back(dW);
            for (int c = 0; c < m; ++c)
                HIPCHECK(hipMemcpyAsync(dV + (size_t)c * n, V[c].data(), sizeof(z64) * n, hipMemcpyHostToDevice, s));
            for (int c = 0; c < m; ++c)
                launch_matvec<double>(s, (const z64*)op->A, n, n, n, dV + (size_t)c * n, dW + (size_t)c * n, nullptr,
                                      nullptr);
            HIPCHECK(hipGetLastError());
            std::vector<zc> W((size_t)n * m);
            HIPCHECK(hipMemcpyAsync(W.data(), dW, sizeof(z64) * n * m, hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            bool finite = true;
            for (int c = 0; c < m; ++c)
                for (int i = 0; i < n; ++i) {
                    zc& w = W[(size_t)c * n + i];
                    w /= hr[i];
                    finite = finite && std::isfinite(w.real()) && std::isfinite(w.imag());
                }
            if (finite) {
                std::vector<zc> cvec(n), y;
                for (int i = 0; i < n; ++i) cvec[i] = hb[i] / hr[i];
                rep.rank = pivoted_qr_solve(W, n, m, cvec, rank_tol, y);
                if (rep.rank == m) {
                    for (int c = 0; c < m; ++c)
                        for (int i = 0; i < n; ++i) x[i] += V[c][i] * y[c];
                } else {
                    rep.reason = BEM_REDUCED_RANK;
                }
            } else {
                rep.reason = BEM_REDUCED_NONFINITE;
            }
        }
        auto residuals = [&](double& orig, double& scaled) {
            launch_matvec<double>(s, (const z64*)op->A, n, n, n, op->vx, op->vr, op->vb, nullptr);
            launch_norm_stats(s, n, op->vr, op->vb, rw, op->red);
            std::vector<double> h(4 * kRedBlocks);
            HIPCHECK(hipMemcpyAsync(h.data(), op->red, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            double ss[4] = {0, 0, 0, 0};
            for (int b = 0; b < kRedBlocks; ++b)
                for (int q = 0; q < 4; ++q) ss[q] += h[4 * b + q];
            orig = rel_or_inf(std::sqrt(ss[0]), std::sqrt(ss[1]));
            scaled = rel_or_inf(std::sqrt(ss[2]), std::sqrt(ss[3]));
        };
        HIPCHECK(hipMemcpyAsync(op->vx, x.data(), sizeof(z64) * n, hipMemcpyHostToDevice, s));
        residuals(rep.projected_original, rep.projected_scaled);
        bool xfinite = true;
        for (const zc& v : x) xfinite = xfinite && std::isfinite(v.real()) && std::isfinite(v.imag());
        const bool accepted =
            xfinite && std::max(rep.projected_original, rep.projected_scaled) <= tol;
        if (accepted) {
            rep.reduced = 1;
            rep.reason = BEM_REDUCED_ACCEPTED;
            rep.residual_original = rep.projected_original;
            rep.residual_scaled = rep.projected_scaled;
        } else {
            if (!xfinite) rep.reason = BEM_REDUCED_NONFINITE;
            rep.reduced = 0;
            operator_factor(op);
            if (!correction(op, op->vb, op->vx, 0, nullptr))
                throw BemError(BEM_ERR_ARGUMENT, "right-hand side must be finite");
            residuals(rep.residual_original, rep.residual_scaled);
            if (!(std::max(rep.residual_original, rep.residual_scaled) <= tol))
                throw BemError(BEM_ERR_NUMERIC, "full solve failed current-equation residual acceptance");
            HIPCHECK(hipMemcpyAsync(x.data(), op->vx, sizeof(z64) * n, hipMemcpyDeviceToHost, s));
        }
        // backscatter amplitude of the CBIE traces x = [p_scat; dp_scat/dn] (far_field, full.jl:509-519)
        z64* part = (z64*)ctx->alloc(sizeof(z64) * kFarBlocks);
        tmp.push_back(part);
        z64* fo = (z64*)ctx->alloc(sizeof(z64));
        tmp.push_back(fo);
        launch_far_field(s, mesh->d, prob.k, 1, 1, dback, op->vx, op->vx + nt, part, fo);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(amp, fo, sizeof(z64), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        for (void* p : tmp) ctx->release(p);
        tmp.clear();
        op_free(op);
        return x;
    } catch (...) {
        hipStreamSynchronize(s);
        for (void* p : tmp) ctx->release(p);
        op_free(op);
        throw;
    }
}

// ---------------------------------------------------------------------------
// Fixed device scratch of bem_far_field_noalloc: the staged directions, the
// per-block partial sums and the amplitudes of at most kNoAllocMaxDirs
// directions, laid out at fixed offsets.  One copy per device lives for the
// life of the process.  In host code the array's name is only a stand-in:
// each context resolves its device's address once, at creation, with
// hipGetSymbolAddress and passes that pointer to the kernels.
// ---------------------------------------------------------------------------
constexpr int64_t kNoAllocMaxDirs = 1024;
constexpr size_t kScratchDirsOff = 0;                                              // 3 doubles per direction
constexpr size_t kScratchPartOff = 32 * 1024;                                      // kFarBlocks z64 per direction
constexpr size_t kScratchResOff = kScratchPartOff + sizeof(z64) * kFarBlocks * kNoAllocMaxDirs;
constexpr size_t kScratchBytes = kScratchResOff + sizeof(z64) * kNoAllocMaxDirs;  // 4.05 MiB
static_assert(sizeof(double) * 3 * kNoAllocMaxDirs <= kScratchPartOff, "direction staging overlaps partials");
static_assert(kScratchResOff % 256 == 0, "amplitudes must stay aligned");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// This device's address of `scratch` (the calling thread has the device
// current); resolved once per device and kept.
void* device_scratch(int device) {
    static std::mutex mu;
    static std::map<int, void*> addr;
    std::lock_guard<std::mutex> g(mu);
    auto it = addr.find(device);
    if (it != addr.end()) return it->second;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        throw BemError(BEM_ERR_DEVICE, std::string("scratch symbol address: ") + hipGetErrorString(e));
    addr[device] = ptr;
    return ptr;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
// ---------------------------------------------------------------------------
// Axisymmetric BEM (reference src/solvers/bem/axisymmetric.jl): meridian
// panels, Fourier modes, one dense FP64 system per mode.
// ---------------------------------------------------------------------------
struct bem_axi_s {
    bem_ctx ctx = nullptr;
    int n = 0;                       // panels
    bem_problem prob{};
    double beta = 0;
    std::vector<int> modes;          // 0 (axial) or 0 .. m_max
    std::vector<double> geo;         // host SoA panels: r1 z1 r2 z2 rm zm L nr nz (9 n)
    double* dgeo = nullptr;          // device copy
    z64 *dp = nullptr, *dd = nullptr;  // device traces [mode][panel]
    std::vector<z64> hp, hd;          // host traces
    double max_residual = 0;
    double assembly_ms = 0, solve_ms = 0;
};

namespace {

AxiRules axi_rules() {
    AxiRules R;
    double x[16], w[16];
    gauss_legendre(8, x, w);
    for (int i = 0; i < 8; ++i) {
        R.x8[i] = 0.5 * (x[i] + 1.0);
        R.w8[i] = 0.5 * w[i];
    }
    gauss_legendre(16, x, w);
    for (int i = 0; i < 16; ++i) {
        R.x16[i] = 0.5 * (x[i] + 1.0);
        R.w16[i] = 0.5 * w[i];
    }
    return R;
}

// Meridian checks (reference src/geometry/meridian.jl:16-80 for an open
// meridian): finite coordinates, at least three panels, both ends on the
// axis and every other node strictly off it, no zero-length panel, and the
// north-to-south traversal that makes (-dz, drho) / L the outward normal
// (the profile closed along the axis runs clockwise in the (rho, z) plane).
void axi_validate(int64_t nn, const double* rho, const double* z) {
    if (nn < 4) throw BemError(BEM_ERR_ARGUMENT, "a meridian needs at least three panels");
    double scale = 0, area2 = 0;
    for (int64_t i = 0; i < nn; ++i) {
        if (!std::isfinite(rho[i]) || !std::isfinite(z[i]))
            throw BemError(BEM_ERR_ARGUMENT, "meridian coordinates must be finite");
        scale = std::max(scale, std::max(std::fabs(rho[i]), std::fabs(z[i])));
    }
    if (!(scale > 0)) throw BemError(BEM_ERR_ARGUMENT, "meridian has zero extent");
    const double tol = 64 * 2.220446049250313e-16 * scale;
    if (std::fabs(rho[0]) > tol || std::fabs(rho[nn - 1]) > tol)
        throw BemError(BEM_ERR_ARGUMENT, "an open meridian must meet the axis only at its two endpoints");
    for (int64_t i = 1; i + 1 < nn; ++i)
        if (!(rho[i] > tol))
            throw BemError(BEM_ERR_ARGUMENT, "an open meridian must meet the axis only at its two endpoints");
    for (int64_t i = 0; i + 1 < nn; ++i) {
        if (!(std::hypot(rho[i + 1] - rho[i], z[i + 1] - z[i]) > tol))
            throw BemError(BEM_ERR_ARGUMENT, "meridian has a repeated or unresolved vertex");
        area2 += rho[i] * z[i + 1] - rho[i + 1] * z[i];
    }
    area2 += rho[nn - 1] * z[0] - rho[0] * z[nn - 1];
    if (!(area2 < 0))
        throw BemError(BEM_ERR_ARGUMENT,
                       "the meridian must run from the north pole to the south pole (outward normals)");
}

void axi_upload(bem_axi h, int64_t nn, const double* rho, const double* z) {
    const int n = (int)(nn - 1);
    h->n = n;
    h->geo.assign(9 * (size_t)n, 0.0);
    for (int j = 0; j < n; ++j) {
        const double r1 = rho[j], z1 = z[j], r2 = rho[j + 1], z2 = z[j + 1];
        const double dr = r2 - r1, dz = z2 - z1, L = std::hypot(dr, dz);
        const double v[9] = {r1, z1, r2, z2, 0.5 * (r1 + r2), 0.5 * (z1 + z2), L, -dz / L, dr / L};
        for (int c = 0; c < 9; ++c) h->geo[(size_t)c * n + j] = v[c];
    }
    h->dgeo = (double*)h->ctx->alloc(sizeof(double) * h->geo.size());
    HIPCHECK(hipMemcpy(h->dgeo, h->geo.data(), sizeof(double) * h->geo.size(), hipMemcpyHostToDevice));
}

AxiPanels axi_panels(bem_axi h) {
    const int n = h->n;
    const double* g = h->dgeo;
    return AxiPanels{g, g + n, g + 2 * n, g + 3 * n, g + 4 * n, g + 5 * n, g + 6 * n, g + 7 * n, g + 8 * n, n};
}

void axi_free(bem_axi h) {
    if (!h) return;
    bem_ctx c = h->ctx;
    hipStreamSynchronize(c->lane0()->s);
    c->release(h->dgeo);
    c->release(h->dp);
    c->release(h->dd);
    delete h;
}

// Incident plane wave e^{ik(z cos b + rho sin b cos phi)}, mode m, at the
// panel midpoints (axisymmetric.jl:967-979,1021-1025).
void axi_incident(const bem_axi h, double k, int m, std::vector<std::complex<double>>& p,
                  std::vector<std::complex<double>>& dp) {
    const int n = h->n;
    const double* g = h->geo.data();
    const double sb = std::sin(h->beta), cb = std::cos(h->beta);
    const double eps = m == 0 ? 1.0 : 2.0;
    std::complex<double> im_m(1.0, 0.0);
    for (int q = 0; q < (m & 3); ++q) im_m *= std::complex<double>(0.0, 1.0);
    p.resize(n);
    dp.resize(n);
    for (int j = 0; j < n; ++j) {
        const double rm = g[4 * n + j], zm = g[5 * n + j], nr = g[7 * n + j], nz = g[8 * n + j];
        const double u = k * rm * sb;
        const double jm = std::cyl_bessel_j((double)m, u);
        const double jd = m == 0 ? -std::cyl_bessel_j(1.0, u)
                                 : 0.5 * (std::cyl_bessel_j((double)(m - 1), u) - std::cyl_bessel_j((double)(m + 1), u));
        const std::complex<double> ph = eps * im_m * std::exp(std::complex<double>(0.0, k * zm * cb));
        p[j] = ph * jm;
        dp[j] = ph * (std::complex<double>(0.0, k * cb * nz) * jm + k * sb * nr * jd);
    }
}

}  // namespace

extern "C" {

const char* bem_version(void) { return "bemhip 0.1.0 (gfx950)"; }

void bem_problem_defaults(bem_problem* p, int32_t bc) {
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->boundary = bc;
    p->formulation = bc == BEM_FLUID ? BEM_MULLER : BEM_BURTON_MILLER;
    p->precision = BEM_DOUBLE;
    p->equilibrate = 1;
    p->k = 1.0;
    p->density_contrast = 1.0;
    p->soundspeed_contrast = 1.0;
    p->refine_tolerance = 1e-13;
    p->refine_maxiter = 8;
    p->force_promotion = 0;
    p->condition_guard = 10.0 * 1.1920928955078125e-07;
}

int bem_ctx_create(int32_t device, uint64_t budget, bem_ctx* out) {
    if (!out) return BEM_ERR_ARGUMENT;
    *out = nullptr;
    auto* c = new bem_ctx_s();
    int rc = guarded(c, [&] {
        int count = 0;
        HIPCHECK(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw BemError(BEM_ERR_ARGUMENT, "invalid device");
        HIPCHECK(hipSetDevice(device));
        c->device = device;
        hipDeviceProp_t prop;
        HIPCHECK(hipGetDeviceProperties(&prop, device));
        c->cu_count = prop.multiProcessorCount;
        size_t fr = 0, tot = 0;
        HIPCHECK(hipMemGetInfo(&fr, &tot));
        c->budget = budget ? budget : (uint64_t)(0.9 * (double)fr);
        if (upload_rules() != 0) throw BemError(BEM_ERR_DEVICE, "constant table upload failed");
        c->scratch = device_scratch(device);
        c->lanes.push_back(make_lane(c));
        HIPCHECK(hipDeviceSynchronize());
    });
    if (rc != BEM_OK) {
        std::fprintf(stderr, "bem_ctx_create: %s\n", c->err.c_str());
        for (Lane* L : c->lanes) free_lane(c, L);
        delete c;
        return rc;
    }
    *out = c;
    return BEM_OK;
}

int bem_ctx_destroy(bem_ctx ctx) {
    if (!ctx) return BEM_ERR_ARGUMENT;
    for (Lane* L : ctx->lanes) hipStreamSynchronize(L->s);
    for (Lane* L : ctx->lanes) free_lane(ctx, L);
    ctx->trim_all();
    delete ctx;
    return BEM_OK;
}

const char* bem_last_error(bem_ctx ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int bem_ctx_memory(bem_ctx ctx, uint64_t* used, uint64_t* budget) {
    if (!ctx) return BEM_ERR_ARGUMENT;
    if (used) *used = ctx->used;
    if (budget) *budget = ctx->budget;
    return BEM_OK;
}

int bem_ctx_set_profiling(bem_ctx ctx, int32_t enable) {
    if (!ctx) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    ctx->prof.on = enable != 0;
    return BEM_OK;
}

int bem_ctx_kernel_stats(bem_ctx ctx, const char* name, int64_t* launches, double* total_ms, double* flops,
                         double* bytes) {
    if (!ctx || !name) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    ctx->prof.drain();
    if (std::strcmp(name, "h_apply") == 0) {
        const Prof& P = ctx->prof;
        if (launches) *launches = P.launches[C_HROWS];
        if (total_ms) *total_ms = P.ms[C_HPROJ] + P.ms[C_HROWS];
        if (flops) *flops = P.flops[C_HPROJ] + P.flops[C_HROWS];
        if (bytes) *bytes = P.bytes[C_HPROJ] + P.bytes[C_HROWS];
        return BEM_OK;
    }
    for (int c = 0; c < kNumClasses; ++c)
        if (std::strcmp(name, kClassNames[c]) == 0) {
            if (launches) *launches = ctx->prof.launches[c];
            if (total_ms) *total_ms = ctx->prof.ms[c];
            if (flops) *flops = ctx->prof.flops[c];
            if (bytes) *bytes = ctx->prof.bytes[c];
            return BEM_OK;
        }
    ctx->err = std::string("unknown kernel class ") + name;
    return BEM_ERR_ARGUMENT;
}

int bem_ctx_set_lanes(bem_ctx ctx, int32_t max_lanes) {
    if (!ctx || max_lanes < 1 || max_lanes > 8) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    ctx->max_lanes = max_lanes;
    return BEM_OK;
}

int bem_ctx_set_heavy_lanes(bem_ctx ctx, int32_t max_heavy) {
    if (!ctx || max_heavy < 1 || max_heavy > 8) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    ctx->max_heavy = max_heavy;
    return BEM_OK;
}

int bem_ctx_set_lu_tuning(bem_ctx ctx, int32_t mode) {
    if (!ctx || mode < 0 || mode > 2) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    std::lock_guard<std::mutex> t(ctx->tmu);
    ctx->lu_mode = mode;
    return BEM_OK;
}

int bem_ctx_autotune(bem_ctx ctx, int64_t n, int32_t precision, bem_lu_tune* out) {
    if (!ctx || n < 64 || n > 524288 || (precision != BEM_DOUBLE && precision != BEM_MIXED)) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&] {
        const bem_lu_tune r = precision == BEM_MIXED ? tune_lu<float>(ctx, (int)n) : tune_lu<double>(ctx, (int)n);
        if (out) *out = r;
    });
}

int bem_ctx_lu_tune(bem_ctx ctx, int64_t n, int32_t precision, bem_lu_tune* out) {
    if (!ctx || !out || n < 1 || n > 524288 || (precision != BEM_DOUBLE && precision != BEM_MIXED))
        return BEM_ERR_ARGUMENT;
    const bool single = precision == BEM_MIXED;
    {
        std::lock_guard<std::mutex> g(ctx->tmu);
        auto it = ctx->tunes.find({(int)n, single ? 1 : 0});
        if (ctx->lu_mode == 2 && it != ctx->tunes.end()) {
            *out = it->second;
            return BEM_OK;
        }
    }
    const LuTune t = lu_tune_for(ctx, (int)n, single);
    *out = bem_lu_tune{};
    out->unknowns = n;
    out->precision = precision;
    out->outer = t.outer;
    out->group = t.group;
    out->split = t.split;
    out->ahead = t.ahead;
    out->graph = t.graph;
    return BEM_OK;
}

int bem_ctx_reset_stats(bem_ctx ctx) {
    if (!ctx) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    ctx->prof.drain();
    for (int c = 0; c < kNumClasses; ++c) {
        ctx->prof.launches[c] = 0;
        ctx->prof.ms[c] = ctx->prof.flops[c] = ctx->prof.bytes[c] = 0;
    }
    return BEM_OK;
}

int bem_icosphere_size(int32_t level, int64_t* nv, int64_t* nt) {
    if (level < 0 || level > 9) return BEM_ERR_ARGUMENT;
    int64_t f = 20;
    for (int i = 0; i < level; ++i) f *= 4;
    if (nt) *nt = f;
    if (nv) *nv = f / 2 + 2;
    return BEM_OK;
}

int bem_icosphere(int32_t level, double ax, double ay, double az, double* nodes, int64_t* tri) {
    if (level < 0 || level > 9 || !nodes || !tri) return BEM_ERR_ARGUMENT;
    if (!(ax > 0 && ay > 0 && az > 0)) return BEM_ERR_ARGUMENT;
    std::vector<double> v;
    std::vector<int64_t> f;
    icosphere(level, ax, ay, az, v, f);
    std::memcpy(nodes, v.data(), sizeof(double) * v.size());
    for (size_t i = 0; i < f.size(); ++i) tri[i] = f[i] + 1;
    return BEM_OK;
}

int bem_revolution_size(int32_t ns, int32_t nphi, int64_t* nv, int64_t* nt) {
    if (ns < 3 || nphi < 3) return BEM_ERR_ARGUMENT;
    if (nv) *nv = (int64_t)(ns - 2) * nphi + 2;
    if (nt) *nt = 2 * (int64_t)(ns - 2) * nphi;
    return BEM_OK;
}

// Rings s = 1..ns-2 with n_phi nodes each (phi_m = 2 pi m / n_phi, node at
// (x_s, r_s cos phi, r_s sin phi)), pole nodes at s = 0 (first) and s = ns-1
// (last).  Faces are ordered outward for a meridian running from -x to +x.
int bem_revolution(int32_t ns, const double* xs, const double* rs, int32_t nphi, double* nodes, int64_t* tri) {
    if (ns < 3 || nphi < 3 || !xs || !rs || !nodes || !tri) return BEM_ERR_ARGUMENT;
    const int rings = ns - 2;
    int64_t v = 0;
    nodes[3 * v] = xs[0]; nodes[3 * v + 1] = 0; nodes[3 * v + 2] = 0; ++v;
    for (int s = 1; s <= rings; ++s)
        for (int m = 0; m < nphi; ++m) {
            const double ph = 2.0 * M_PI * m / nphi;
            nodes[3 * v] = xs[s];
            nodes[3 * v + 1] = rs[s] * std::cos(ph);
            nodes[3 * v + 2] = rs[s] * std::sin(ph);
            ++v;
        }
    nodes[3 * v] = xs[ns - 1]; nodes[3 * v + 1] = 0; nodes[3 * v + 2] = 0;
    const int64_t last = v;
    auto ring = [&](int s, int m) -> int64_t { return 1 + (int64_t)(s - 1) * nphi + ((m % nphi) + nphi) % nphi; };
    int64_t t = 0;
    auto put = [&](int64_t a, int64_t b, int64_t c) {
        tri[3 * t] = a + 1; tri[3 * t + 1] = b + 1; tri[3 * t + 2] = c + 1; ++t;
    };
    for (int m = 0; m < nphi; ++m) put(0, ring(1, m + 1), ring(1, m));
    for (int s = 1; s < rings; ++s)
        for (int m = 0; m < nphi; ++m) {
            put(ring(s, m), ring(s, m + 1), ring(s + 1, m + 1));
            put(ring(s, m), ring(s + 1, m + 1), ring(s + 1, m));
        }
    for (int m = 0; m < nphi; ++m) put(last, ring(rings, m), ring(rings, m + 1));
    return BEM_OK;
}

int bem_mesh_validate(int64_t nv, const double* nodes, int64_t nt, const int64_t* tri, double* resolution,
                      char* message, int32_t mlen) {
    if (!nodes || !tri || nv <= 0) {
        if (message && mlen > 0) std::snprintf(message, mlen, "invalid mesh arrays");
        return BEM_ERR_ARGUMENT;
    }
    std::vector<int64_t> t0(3 * (size_t)std::max<int64_t>(nt, 0));
    for (size_t i = 0; i < t0.size(); ++i) t0[i] = tri[i] - 1;
    std::string msg;
    double vol = 0;
    int code = validate_topology(nv, nodes, nt, t0.data(), resolution, &vol, msg);
    if (message && mlen > 0) std::snprintf(message, mlen, "%s", msg.c_str());
    return code ? BEM_ERR_ARGUMENT : BEM_OK;
}

int bem_mesh_create(bem_ctx ctx, int64_t nv, const double* nodes, int64_t nt, const int64_t* tri, bem_mesh* out) {
    if (!ctx || !out) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&] {
        if (!nodes || !tri || nv <= 0 || nt <= 0) throw BemError(BEM_ERR_ARGUMENT, "surface must contain triangles");
        if (nt > (int64_t)INT_MAX / 2) throw BemError(BEM_ERR_ARGUMENT, "too many triangles");
        auto* m = new bem_mesh_s();
        m->ctx = ctx;
        HostMesh& h = m->h;
        h.nv = nv;
        h.nt = nt;
        h.nodes.assign(nodes, nodes + 3 * nv);
        h.tri.resize(3 * nt);
        for (int64_t i = 0; i < 3 * nt; ++i) h.tri[i] = tri[i] - 1;
        std::string msg;
        int code = validate_topology(nv, h.nodes.data(), nt, h.tri.data(), &h.resolution, &h.volume, msg);
        if (code) {
            delete m;
            throw BemError(BEM_ERR_ARGUMENT, msg);
        }
        build_geometry(h);
        build_near_list(h);
        try {
            hipStream_t s = ctx->lane0()->s;
            DevMesh& d = m->d;
            d.nt = (int)nt;
            auto up = [&](const void* src, size_t bytes) {
                void* p = ctx->alloc(bytes);
                m->bufs.push_back(p);
                HIPCHECK(hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, s));
                return p;
            };
            std::vector<double> cen(3 * nt), nrm(3 * nt), gp(21 * nt);
            for (int64_t j = 0; j < nt; ++j)
                for (int c = 0; c < 3; ++c) {
                    cen[c * nt + j] = h.cen[3 * j + c];
                    nrm[c * nt + j] = h.nrm[3 * j + c];
                }
            for (int64_t j = 0; j < nt; ++j)
                for (int q = 0; q < 7; ++q)
                    for (int c = 0; c < 3; ++c) gp[(3 * q + c) * nt + j] = h.gauss[3 * (7 * j + q) + c];
            std::vector<float> cen32(cen.begin(), cen.end()), nrm32(nrm.begin(), nrm.end()),
                gp32(gp.begin(), gp.end()), area32(h.area.begin(), h.area.end());
            d.cen = (double*)up(cen.data(), sizeof(double) * cen.size());
            d.nrm = (double*)up(nrm.data(), sizeof(double) * nrm.size());
            d.area = (double*)up(h.area.data(), sizeof(double) * nt);
            d.gp = (double*)up(gp.data(), sizeof(double) * gp.size());
            d.vert = (double*)up(h.vert.data(), sizeof(double) * h.vert.size());
            d.cen32 = (float*)up(cen32.data(), sizeof(float) * cen32.size());
            d.nrm32 = (float*)up(nrm32.data(), sizeof(float) * nrm32.size());
            d.area32 = (float*)up(area32.data(), sizeof(float) * nt);
            d.gp32 = (float*)up(gp32.data(), sizeof(float) * gp32.size());
            std::vector<int32_t> rowof(h.near_col.size());
            for (int64_t i = 0; i < nt; ++i)
                for (int32_t e = h.near_ptr[i]; e < h.near_ptr[i + 1]; ++e) rowof[e] = (int32_t)i;
            d.nnear = (int64_t)h.near_col.size();
            d.near_ptr = (int*)up(h.near_ptr.data(), sizeof(int32_t) * h.near_ptr.size());
            d.near_col = (int*)up(h.near_col.data(), sizeof(int32_t) * std::max<size_t>(1, h.near_col.size()));
            d.near_row = (int*)up(rowof.data(), sizeof(int32_t) * std::max<size_t>(1, rowof.size()));
            HIPCHECK(hipStreamSynchronize(s));
        } catch (...) {
            hipStreamSynchronize(ctx->lane0()->s);
            for (void* p : m->bufs) ctx->release(p);
            delete m;
            throw;
        }
        *out = m;
    });
}

int bem_nystrom_quadrature(int64_t nv, const double* nodes, int64_t nt, int32_t npt, const int64_t* tri,
                           int32_t qorder, int64_t* n_nodes, double* coords, double* normals, double* weights,
                           char* message, int32_t mlen) {
    std::string msg;
    auto fail = [&](const std::string& m) {
        if (message && mlen > 0) std::snprintf(message, mlen, "%s", m.c_str());
        return BEM_ERR_ARGUMENT;
    };
    if (!nodes || !tri || nv <= 0 || nt <= 0) return fail("surface must contain triangles");
    try {
        HostMesh h;
        h.nv = nv;
        h.nodes.assign(nodes, nodes + 3 * nv);
        h.npt = npt;
        h.qorder = qorder;
        h.nel = nt;
        if (npt != 3 && npt != 6 && npt != 10) return fail("triangles must have 3, 6 or 10 rows in Gmsh ordering");
        h.elem.resize((size_t)npt * nt);
        for (size_t i = 0; i < h.elem.size(); ++i) h.elem[i] = tri[i] - 1;
        if (validate_curved(nv, h.nodes.data(), nt, npt, h.elem.data(), &h.resolution, &h.volume, msg))
            return fail(msg);
        if (build_nystrom(h, msg)) return fail(msg);
        if (n_nodes) *n_nodes = h.nt;
        for (int64_t p = 0; p < h.nt; ++p) {
            for (int c = 0; c < 3; ++c) {
                if (coords) coords[3 * p + c] = h.cen[3 * p + c];
                if (normals) normals[3 * p + c] = h.nrm[3 * p + c];
            }
            if (weights) weights[p] = h.area[p];
        }
    } catch (const std::exception& e) {
        return fail(e.what());
    }
    if (message && mlen > 0) message[0] = 0;
    return BEM_OK;
}

int bem_mesh_create_nystrom(bem_ctx ctx, int64_t nv, const double* nodes, int64_t nt, int32_t npt,
                            const int64_t* tri, int32_t qorder, bem_mesh* out) {
    if (!ctx || !out) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&] {
        if (!nodes || !tri || nv <= 0 || nt <= 0) throw BemError(BEM_ERR_ARGUMENT, "surface must contain triangles");
        if (npt != 3 && npt != 6 && npt != 10)
            throw BemError(BEM_ERR_ARGUMENT, "triangles must have 3, 6 or 10 rows in Gmsh ordering");
        if (qorder < 1 || qorder > 8) throw BemError(BEM_ERR_ARGUMENT, "qorder must be between 1 and 8");
        if (nt > (int64_t)INT_MAX / 64) throw BemError(BEM_ERR_ARGUMENT, "too many triangles");
        auto* m = new bem_mesh_s();
        m->ctx = ctx;
        HostMesh& h = m->h;
        h.nv = nv;
        h.nodes.assign(nodes, nodes + 3 * nv);
        h.npt = npt;
        h.qorder = qorder;
        h.nel = nt;
        h.elem.resize((size_t)npt * nt);
        for (size_t i = 0; i < h.elem.size(); ++i) h.elem[i] = tri[i] - 1;
        std::string msg;
        if (validate_curved(nv, h.nodes.data(), nt, npt, h.elem.data(), &h.resolution, &h.volume, msg) ||
            build_nystrom(h, msg)) {
            delete m;
            throw BemError(BEM_ERR_ARGUMENT, msg);
        }
        if (h.nt > (int64_t)INT_MAX / 2) {
            delete m;
            throw BemError(BEM_ERR_ARGUMENT, "too many quadrature nodes");
        }
        try {
            hipStream_t s = ctx->lane0()->s;
            DevMesh& d = m->d;
            d.nt = (int)h.nt;
            d.nq = h.nq;
            auto up = [&](const void* src, size_t bytes) {
                void* p = ctx->alloc(bytes);
                m->bufs.push_back(p);
                HIPCHECK(hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, s));
                return p;
            };
            const int64_t n = h.nt;
            std::vector<double> cen(3 * n), nrm(3 * n);
            for (int64_t j = 0; j < n; ++j)
                for (int c = 0; c < 3; ++c) {
                    cen[c * n + j] = h.cen[3 * j + c];
                    nrm[c * n + j] = h.nrm[3 * j + c];
                }
            d.cen = (double*)up(cen.data(), sizeof(double) * cen.size());
            d.nrm = (double*)up(nrm.data(), sizeof(double) * nrm.size());
            d.area = (double*)up(h.area.data(), sizeof(double) * n);
            // FP32 copies for the compressed path's regular entries
            std::vector<float> cen32(cen.begin(), cen.end()), nrm32(nrm.begin(), nrm.end()),
                area32(h.area.begin(), h.area.end());
            d.cen32 = (float*)up(cen32.data(), sizeof(float) * cen32.size());
            d.nrm32 = (float*)up(nrm32.data(), sizeof(float) * nrm32.size());
            d.area32 = (float*)up(area32.data(), sizeof(float) * n);
            m->sources = (double*)up(h.sources.data(), sizeof(double) * h.sources.size());
            std::vector<double> subs;
            sub_triangles(h, subs);
            d.vert = (double*)up(subs.data(), sizeof(double) * subs.size());
            d.nsub = (int)(subs.size() / 9);
            HIPCHECK(hipStreamSynchronize(s));
        } catch (...) {
            hipStreamSynchronize(ctx->lane0()->s);
            for (void* p : m->bufs) ctx->release(p);
            delete m;
            throw;
        }
        *out = m;
    });
}

int bem_mesh_nystrom_info(bem_mesh m, int64_t* n_elements, int32_t* npt, int32_t* nq, int32_t* qorder) {
    if (!m) return BEM_ERR_ARGUMENT;
    if (n_elements) *n_elements = m->h.nq ? m->h.nel : m->h.nt;
    if (npt) *npt = m->h.npt;
    if (nq) *nq = m->h.nq;
    if (qorder) *qorder = m->h.qorder;
    return BEM_OK;
}

int bem_curved_icosphere_size(int32_t level, int32_t order, int64_t* nv, int64_t* nt) {
    if (level < 0 || level > 9 || order < 1 || order > 3 || !nv || !nt) return BEM_ERR_ARGUMENT;
    const int64_t f = 20 * ((int64_t)1 << (2 * level));
    const int64_t v = 10 * ((int64_t)1 << (2 * level)) + 2;
    const int64_t e = 3 * f / 2;
    *nt = f;
    *nv = order == 1 ? v : order == 2 ? v + e : v + 2 * e + f;
    return BEM_OK;
}

int bem_curved_icosphere(int32_t level, double ax, double ay, double az, int32_t order, double* nodes,
                         int64_t* tri) {
    if (level < 0 || level > 9 || order < 1 || order > 3 || !nodes || !tri) return BEM_ERR_ARGUMENT;
    std::vector<double> v;
    std::vector<int64_t> t;
    curved_icosphere(level, ax, ay, az, order, v, t);
    std::memcpy(nodes, v.data(), sizeof(double) * v.size());
    for (size_t i = 0; i < t.size(); ++i) tri[i] = t[i] + 1;
    return BEM_OK;
}

int bem_mesh_destroy(bem_mesh mesh) {
    if (!mesh) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(mesh->ctx->mu);
    mesh_release(mesh);
    return BEM_OK;
}

int bem_mesh_info(bem_mesh m, int64_t* nt, int64_t* nv, double* res, double* vol) {
    if (!m) return BEM_ERR_ARGUMENT;
    if (nt) *nt = m->h.nt;
    if (nv) *nv = m->h.nv;
    if (res) *res = m->h.resolution;
    if (vol) *vol = m->h.volume;
    return BEM_OK;
}

int bem_mesh_panels(bem_mesh m, double* c, double* n, double* a) {
    if (!m) return BEM_ERR_ARGUMENT;
    if (c) std::memcpy(c, m->h.cen.data(), sizeof(double) * m->h.cen.size());
    if (n) std::memcpy(n, m->h.nrm.data(), sizeof(double) * m->h.nrm.size());
    if (a) std::memcpy(a, m->h.area.data(), sizeof(double) * m->h.area.size());
    return BEM_OK;
}

int bem_mesh_quadrature(bem_mesh m, double* coords, double* normals, double* weights) {
    if (!m) return BEM_ERR_ARGUMENT;
    if (m->h.nq) {
        for (int64_t p = 0; p < m->h.nt; ++p) {
            for (int c = 0; c < 3; ++c) {
                if (coords) coords[3 * p + c] = m->h.cen[3 * p + c];
                if (normals) normals[3 * p + c] = m->h.nrm[3 * p + c];
            }
            if (weights) weights[p] = m->h.area[p];
        }
        return BEM_OK;
    }
    double l[7][3], w[7];
    q7_rule(l, w);
    const int64_t nt = m->h.nt;
    for (int64_t j = 0; j < nt; ++j)
        for (int q = 0; q < 7; ++q) {
            const int64_t p = 7 * j + q;
            for (int c = 0; c < 3; ++c) {
                if (coords) coords[3 * p + c] = m->h.gauss[3 * p + c];
                if (normals) normals[3 * p + c] = m->h.nrm[3 * j + c];
            }
            if (weights) weights[p] = w[q] * m->h.area[j];
        }
    return BEM_OK;
}

int bem_mesh_near_count(bem_mesh m, int64_t* count) {
    if (!m || !count) return BEM_ERR_ARGUMENT;
    *count = (int64_t)m->h.near_col.size();
    return BEM_OK;
}

int bem_operator_create(bem_ctx ctx, bem_mesh mesh, const bem_problem* problem, bem_operator* out) {
    if (!ctx) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&] { operator_create(ctx, ctx->lane0(), mesh, problem, out); });
}

int bem_operator_destroy(bem_operator op) {
    if (!op) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(op->ctx->mu);
    op_free(op);
    return BEM_OK;
}

int bem_operator_size(bem_operator op, int64_t* n) {
    if (!op || !n) return BEM_ERR_ARGUMENT;
    *n = op->n;
    return BEM_OK;
}

int bem_operator_matrix(bem_operator op, double* Aout, int64_t ld) {
    if (!op || !Aout || ld < op->n) return BEM_ERR_ARGUMENT;
    bem_ctx ctx = op->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&] {
        const int n = op->n;
        std::vector<z64> h((size_t)n * n);
        if (op->single) {
            std::vector<z32> f((size_t)n * n);
            HIPCHECK(hipMemcpy(f.data(), op->A, sizeof(z32) * f.size(), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < f.size(); ++i) h[i] = z64{f[i].x, f[i].y};
        } else {
            HIPCHECK(hipMemcpy(h.data(), op->A, sizeof(z64) * h.size(), hipMemcpyDeviceToHost));
        }
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                Aout[2 * (i + j * ld)] = h[(size_t)i * n + j].x;
                Aout[2 * (i + j * ld) + 1] = h[(size_t)i * n + j].y;
            }
    });
}

int bem_operator_solve(bem_operator op, int32_t nrhs, const double* angles, bem_solution* out) {
    if (!op) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(op->ctx->mu);
    return guarded(op->ctx, [&] { operator_solve(op, nrhs, angles, out); });
}

int bem_solve(bem_ctx ctx, bem_mesh mesh, const bem_problem* problem, double beta, double alpha, bem_solution* out) {
    if (!ctx) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&] {
        bem_operator op = nullptr;
        operator_create(ctx, ctx->lane0(), mesh, problem, &op);
        const double ang[2] = {beta, alpha};
        try {
            operator_solve(op, 1, ang, out);
        } catch (...) {
            op_free(op);
            throw;
        }
        op_free(op);
    });
}

int bem_solution_destroy(bem_solution sol) {
    if (!sol) return BEM_ERR_ARGUMENT;
    bem_ctx ctx = sol->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    ctx->release(sol->p);
    ctx->release(sol->q);
    mesh_release(sol->mesh);
    delete sol;
    return BEM_OK;
}

int bem_solution_count(bem_solution sol, int32_t* nrhs, int64_t* np) {
    if (!sol) return BEM_ERR_ARGUMENT;
    if (nrhs) *nrhs = sol->nrhs;
    if (np) *np = sol->nt;
    return BEM_OK;
}

int bem_solution_traces(bem_solution sol, int32_t r, double* p, double* q) {
    if (!sol || r < 0 || r >= sol->nrhs) return BEM_ERR_ARGUMENT;
    const size_t nt = sol->nt;
    if (p) std::memcpy(p, sol->hp.data() + r * nt, sizeof(double) * 2 * nt);
    if (q) std::memcpy(q, sol->hq.data() + r * nt, sizeof(double) * 2 * nt);
    return BEM_OK;
}

int bem_solution_diagnostics(bem_solution sol, int32_t r, bem_diagnostics* out) {
    if (!sol || !out || r < 0 || r >= sol->nrhs) return BEM_ERR_ARGUMENT;
    *out = sol->diag[r];
    return BEM_OK;
}

int bem_far_field(bem_solution sol, int32_t r, int64_t ndir, const double* dirs, double* out) {
    if (!sol || !out || r < 0 || r >= sol->nrhs || ndir <= 0 || ndir > (1 << 20)) return BEM_ERR_ARGUMENT;
    bem_ctx ctx = sol->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&] {
        std::vector<double> hd(3 * ndir);
        if (dirs) {
            for (int64_t d = 0; d < ndir; ++d) {
                const double* x = dirs + 3 * d;
                if (!(std::isfinite(x[0]) && std::isfinite(x[1]) && std::isfinite(x[2])))
                    throw BemError(BEM_ERR_ARGUMENT, "direction must be finite");
            }
            std::memcpy(hd.data(), dirs, sizeof(double) * 3 * ndir);
        } else {
            double d[3];
            direction(sol->angles[2 * r], sol->angles[2 * r + 1], d);
            for (int64_t q = 0; q < ndir; ++q)
                for (int c = 0; c < 3; ++c) hd[3 * q + c] = -d[c];
        }
        hipStream_t s = ctx->lane0()->s;
        double* dd = (double*)ctx->alloc(sizeof(double) * hd.size());
        z64* part = nullptr;
        z64* res = nullptr;
        try {
            part = (z64*)ctx->alloc(sizeof(z64) * ndir * kFarBlocks);
            res = (z64*)ctx->alloc(sizeof(z64) * ndir);
            HIPCHECK(hipMemcpyAsync(dd, hd.data(), sizeof(double) * hd.size(), hipMemcpyHostToDevice, s));
            prof_begin(&ctx->lane0()->pl, s, C_FAR);
            launch_far_field(s, sol->mesh->d, sol->k, 1, (int)ndir, dd, sol->p + (size_t)r * sol->nt,
                             sol->q + (size_t)r * sol->nt, part, res);
            prof_end(&ctx->lane0()->pl, s, C_FAR, (double)sol->nt * 7 * ndir, (double)sol->nt * ndir * 7 * 8);
            HIPCHECK(hipMemcpyAsync(out, res, sizeof(z64) * ndir, hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
        } catch (...) {
            hipStreamSynchronize(s);
            ctx->release(dd); ctx->release(part); ctx->release(res);
            throw;
        }
        ctx->release(dd); ctx->release(part); ctx->release(res);
    });
}

int bem_far_field_noalloc(bem_solution sol, int32_t r, int64_t ndir, const double* dirs, double* out) {
    if (!sol || !out || r < 0 || r >= sol->nrhs || ndir <= 0) return BEM_ERR_ARGUMENT;
    bem_ctx ctx = sol->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&] {
        if (ndir > kNoAllocMaxDirs)
            throw BemError(BEM_ERR_ARGUMENT, "bem_far_field_noalloc: at most 1024 directions per call");
        if (!ctx->scratch) throw BemError(BEM_ERR_STATE, "context has no scratch address");
        hipStream_t s = ctx->lane0()->s;
        unsigned char* base = (unsigned char*)ctx->scratch;
        double* dd = (double*)(base + kScratchDirsOff);
        z64* part = (z64*)(base + kScratchPartOff);
        z64* res = (z64*)(base + kScratchResOff);
        // Every element the kernels read is written earlier in this call
        // (directions by the copy, partials by k_far_partial, amplitudes by
        // k_far_final), so nothing left by a previous call is read.
        int64_t nd = ndir;
        if (dirs) {
            for (int64_t d = 0; d < ndir; ++d) {
                const double* x = dirs + 3 * d;
                if (!(std::isfinite(x[0]) && std::isfinite(x[1]) && std::isfinite(x[2])))
                    throw BemError(BEM_ERR_ARGUMENT, "direction must be finite");
            }
            HIPCHECK(hipMemcpyAsync(dd, dirs, sizeof(double) * 3 * ndir, hipMemcpyHostToDevice, s));
        } else {
            // backscatter: one direction, its amplitude repeated below
            double d[3];
            direction(sol->angles[2 * r], sol->angles[2 * r + 1], d);
            for (double& c : d) c = -c;
            nd = 1;
            HIPCHECK(hipMemcpyAsync(dd, d, sizeof d, hipMemcpyHostToDevice, s));
            HIPCHECK(hipStreamSynchronize(s));
        }
        launch_far_field(s, sol->mesh->d, sol->k, 1, (int)nd, dd, sol->p + (size_t)r * sol->nt,
                         sol->q + (size_t)r * sol->nt, part, res);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(out, res, sizeof(z64) * nd, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        for (int64_t d = nd; d < ndir; ++d) {
            out[2 * d] = out[0];
            out[2 * d + 1] = out[1];
        }
    });
}

int bem_pressure(bem_solution sol, int32_t r, int64_t npts, const double* points, int32_t field, double* out) {
    if (!sol || !out || !points || r < 0 || r >= sol->nrhs || npts <= 0 || npts > ((int64_t)1 << 28))
        return BEM_ERR_ARGUMENT;
    bem_ctx ctx = sol->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&] {
        if (field < BEM_FIELD_TOTAL || field > BEM_FIELD_INTERIOR)
            throw BemError(BEM_ERR_ARGUMENT, "field must be :total, :scattered, :incident or :interior");
        for (int64_t i = 0; i < 3 * npts; ++i)
            if (!std::isfinite(points[i])) throw BemError(BEM_ERR_ARGUMENT, "point coordinates must be finite");
        if (sol->mesh->d.nq && field != BEM_FIELD_INCIDENT) {
            nys_pressure(sol, r, npts, points, field, out);
            return;
        }
        FieldParams fp{};
        fp.k = sol->k;
        fp.fluid = sol->prob.boundary == BEM_FLUID ? 1 : 0;
        fp.g = fp.fluid ? sol->prob.density_contrast : 1.0;
        fp.kint = fp.fluid ? sol->k / sol->prob.soundspeed_contrast : sol->k;
        fp.field = field;
        direction(sol->angles[2 * r], sol->angles[2 * r + 1], fp.d);
        hipStream_t s = ctx->lane0()->s;
        const int64_t chunk = std::min<int64_t>(npts, 1 << 20);
        double* dp = (double*)ctx->alloc(sizeof(double) * 3 * chunk);
        z64* dout = nullptr;
        int* dreg = nullptr;
        double* dwind = nullptr;
        try {
            dout = (z64*)ctx->alloc(sizeof(z64) * chunk);
            dreg = (int*)ctx->alloc(sizeof(int) * chunk);
            dwind = (double*)ctx->alloc(sizeof(double) * chunk);
            std::vector<int> reg(chunk);
            for (int64_t i0 = 0; i0 < npts; i0 += chunk) {
                const int nc = (int)std::min<int64_t>(chunk, npts - i0);
                HIPCHECK(hipMemcpyAsync(dp, points + 3 * i0, sizeof(double) * 3 * nc, hipMemcpyHostToDevice, s));
                launch_pressure(s, sol->mesh->d, fp, nc, dp, sol->p + (size_t)r * sol->nt,
                                sol->q + (size_t)r * sol->nt, dout, dreg, dwind);
                HIPCHECK(hipGetLastError());
                HIPCHECK(hipMemcpyAsync(out + 2 * i0, dout, sizeof(z64) * nc, hipMemcpyDeviceToHost, s));
                HIPCHECK(hipMemcpyAsync(reg.data(), dreg, sizeof(int) * nc, hipMemcpyDeviceToHost, s));
                HIPCHECK(hipStreamSynchronize(s));
                for (int i = 0; i < nc; ++i)
                    if (reg[i] == REGION_NONE)
                        throw BemError(BEM_ERR_ARGUMENT, "the selected field has no acoustic pressure at this point");
            }
        } catch (...) {
            hipStreamSynchronize(s);
            ctx->release(dp); ctx->release(dout); ctx->release(dreg); ctx->release(dwind);
            throw;
        }
        ctx->release(dp); ctx->release(dout); ctx->release(dreg); ctx->release(dwind);
    });
}

void bem_reduced_defaults(bem_reduced_options* o) {
    if (!o) return;
    o->tolerance = 1e-10;
    o->rank_tolerance = 1e-12;
    o->capacity = 16;
    o->maxbytes = 64ll * 1024 * 1024;
}

int bem_reduced_sweep(bem_ctx ctx, bem_mesh mesh, double g, double h, int32_t n_train, const double* train_k,
                      int32_t n_query, const double* query_k, double beta, double alpha,
                      const bem_reduced_options* opts, double* amplitudes, bem_reduced_report* train,
                      bem_reduced_report* query, int32_t* basis_rank) {
    if (!ctx || !mesh || !train_k || !query_k || !opts || !amplitudes) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> gd(ctx->mu);
    return guarded(ctx, [&] {
        if (n_train <= 0 || n_query <= 0)
            throw BemError(BEM_ERR_ARGUMENT, "query and training frequencies must be nonempty, finite and positive");
        for (int i = 0; i < n_train; ++i)
            if (!(std::isfinite(train_k[i]) && train_k[i] > 0))
                throw BemError(BEM_ERR_ARGUMENT, "converted wavenumbers must be finite and positive");
        for (int i = 0; i < n_query; ++i)
            if (!(std::isfinite(query_k[i]) && query_k[i] > 0))
                throw BemError(BEM_ERR_ARGUMENT, "converted wavenumbers must be finite and positive");
        if (!(std::isfinite(beta) && std::isfinite(alpha)))
            throw BemError(BEM_ERR_ARGUMENT, "incidence angles must be finite");
        const double tol = opts->tolerance, rtol = opts->rank_tolerance;
        if (!(std::isfinite(tol) && tol > 0 && tol < 1 && std::isfinite(rtol) && rtol > 0 && rtol < 1))
            throw BemError(BEM_ERR_ARGUMENT, "residual and rank tolerances must lie in (0, 1)");
        if (opts->capacity < 0 || opts->maxbytes < 0) throw BemError(BEM_ERR_ARGUMENT, "negative storage bound");
        bem_problem prob;
        bem_problem_defaults(&prob, BEM_FLUID);
        prob.formulation = BEM_CBIE;
        prob.precision = BEM_DOUBLE;
        prob.density_contrast = g;
        prob.soundspeed_contrast = h;
        const int64_t n = 2 * mesh->d.nt;
        const int capacity = (int)std::min<int64_t>(
            std::min<int64_t>(opts->capacity, n), opts->maxbytes / (int64_t)sizeof(z64) / n);
        hipStream_t s = ctx->lane0()->s;
        double d[6];
        direction(beta, alpha, d);
        for (int c = 0; c < 3; ++c) d[3 + c] = -d[c];
        double* dd = (double*)ctx->alloc(sizeof(double) * 6);
        try {
            HIPCHECK(hipMemcpyAsync(dd, d, sizeof d, hipMemcpyHostToDevice, s));
            std::vector<std::vector<zc>> V;
            const std::vector<std::vector<zc>> empty;
            for (int i = 0; i < n_train; ++i) {
                bem_reduced_report rep{};
                rep.k = train_k[i];
                prob.k = train_k[i];
                validate_problem(prob);
                double amp[2];
                std::vector<zc> x = reduced_equation(ctx, mesh, prob, dd, dd + 3, empty, tol, rtol, rep, amp);
                rep.kept = add_snapshot(V, capacity, x, rtol) ? 1 : 0;
                rep.rank = (int)V.size();
                if (train) train[i] = rep;
            }
            for (int i = 0; i < n_query; ++i) {
                bem_reduced_report rep{};
                rep.k = query_k[i];
                prob.k = query_k[i];
                validate_problem(prob);
                reduced_equation(ctx, mesh, prob, dd, dd + 3, V, tol, rtol, rep, amplitudes + 2 * i);
                if (query) query[i] = rep;
            }
            if (basis_rank) *basis_rank = (int32_t)V.size();
        } catch (...) {
            hipStreamSynchronize(s);
            ctx->release(dd);
            throw;
        }
        ctx->release(dd);
    });
}

int bem_sweep(bem_ctx ctx, int32_t n, const bem_mesh* meshes, const bem_problem* problems, const double* angles,
              double* amplitudes, double* latency_ms, bem_diagnostics* diagnostics) {
    if (!ctx || n <= 0 || !meshes || !problems || !angles || !amplitudes) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&] {
        for (int i = 0; i < n; ++i) {
            if (!meshes[i]) throw BemError(BEM_ERR_ARGUMENT, "null mesh");
            validate_problem(problems[i]);
            if (!std::isfinite(angles[2 * i]) || !std::isfinite(angles[2 * i + 1]))
                throw BemError(BEM_ERR_ARGUMENT, "incidence angles must be finite");
        }
        // largest first, so lanes finish together
        std::vector<int> order(n);
        std::vector<size_t> need(n);
        for (int i = 0; i < n; ++i) {
            order[i] = i;
            need[i] = sweep_need(meshes[i], problems[i]);
        }
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return need[a] > need[b]; });
        // the tuner's first pass over each size class (lane 0, before the lanes start)
        if (ctx->lu_mode == 2) {
            std::set<std::pair<int, int>> classes;
            for (int i = 0; i < n; ++i)
                classes.insert({(int)sweep_unknowns(meshes[i], problems[i]), problems[i].precision == BEM_MIXED});
            for (const auto& cl : classes) {
                bool known;
                {
                    std::lock_guard<std::mutex> t(ctx->tmu);
                    known = ctx->tunes.count(cl) != 0;
                }
                if (!known) {
                    if (cl.second) tune_lu<float>(ctx, cl.first);
                    else tune_lu<double>(ctx, cl.first);
                }
            }
        }
        const size_t max_need = need[order[0]];
        // lanes: as many as the HBM budget holds arenas for (at most 8)
        uint64_t arena_now = 0;
        for (Lane* L : ctx->lanes) arena_now += L->cap;
        const uint64_t avail = ctx->budget - (ctx->used - arena_now);
        int S = (int)std::min<uint64_t>((uint64_t)ctx->max_lanes, (uint64_t)(0.9 * (double)avail) / max_need);
        S = std::max(1, std::min(S, n));
        while ((int)ctx->lanes.size() < S) ctx->lanes.push_back(make_lane(ctx));
        // release arenas that are too small or unused, then (re)allocate
        bool realloc = false;
        for (int l = 0; l < (int)ctx->lanes.size(); ++l) {
            Lane* L = ctx->lanes[l];
            if (L->arena && (l >= S || L->cap < max_need || L->cap > max_need + max_need / 4)) {
                ctx->release(L->arena);
                L->arena = nullptr;
                L->cap = 0;
            }
            realloc |= l < S && !L->arena;
        }
        // New arenas start from an empty pool: blocks pooled by an earlier
        // sweep of another size (an FP64 sweep's arenas are twice a mixed
        // one's) would otherwise stay mapped beside the new ones and fill HBM
        // (a mixed sweep after an FP64 one ran 16 % slower so).  The lanes are
        // idle here, so hipFree's device-wide wait costs nothing.
        if (realloc) ctx->trim_all();
        for (int l = 0; l < S; ++l) {
            Lane* L = ctx->lanes[l];
            if (!L->arena) {
                L->arena = (char*)ctx->alloc(max_need);
                L->cap = max_need;
            }
            L->off = 0;
        }
        // At most ctx->max_heavy heavy items (kHeavyUnknowns) run at once;
        // lanes that may not start one take light items (latency-bound panel
        // chains) meanwhile.  Both lists run largest first.
        std::vector<int> heavy, light;
        for (int i : order) (sweep_unknowns(meshes[i], problems[i]) >= kHeavyUnknowns ? heavy : light).push_back(i);
        size_t hnext = 0, lnext = 0;
        int heavy_running = 0;
        const int max_heavy = ctx->max_heavy;
        std::mutex qmu;
        std::condition_variable qcv;
        std::atomic<bool> failed{false};
        int err_code = BEM_OK;
        std::string err_msg;
        // next item for a lane, or -1 when none is left for it (or after a failure)
        auto take = [&](bool& is_heavy) -> int {
            std::unique_lock<std::mutex> lk(qmu);
            for (;;) {
                if (failed.load()) return -1;
                if (heavy_running < max_heavy && hnext < heavy.size()) {
                    ++heavy_running;
                    is_heavy = true;
                    return heavy[hnext++];
                }
                if (lnext < light.size()) {
                    is_heavy = false;
                    return light[lnext++];
                }
                if (hnext >= heavy.size()) return -1;
                qcv.wait(lk);  // only heavy items remain and max_heavy are running
            }
        };
        auto done = [&](bool is_heavy, int code, const char* msg) {
            std::lock_guard<std::mutex> lk(qmu);
            if (is_heavy) --heavy_running;
            if (msg && !failed.exchange(true)) {
                err_code = code;
                err_msg = msg;
            }
            qcv.notify_all();
        };
        auto worker = [&](int li) {
            (void)hipSetDevice(ctx->device);
            Lane* L = ctx->lanes[li];
            for (;;) {
                bool is_heavy = false;
                const int i = take(is_heavy);
                if (i < 0) break;
                const double t0 = now_ms();
                try {
                    sweep_item(ctx, L, meshes[i], problems[i], angles + 2 * i, amplitudes + 2 * i,
                               diagnostics ? diagnostics + i : nullptr);
                } catch (const BemError& e) {
                    done(is_heavy, e.code, e.what());
                    break;
                } catch (const std::exception& e) {
                    done(is_heavy, BEM_ERR_DEVICE, e.what());
                    break;
                }
                if (latency_ms) latency_ms[i] = now_ms() - t0;
                done(is_heavy, BEM_OK, nullptr);
            }
        };
        std::vector<std::thread> th;
        for (int l = 1; l < S; ++l) th.emplace_back(worker, l);
        worker(0);
        for (auto& t : th) t.join();
        if (failed.load()) throw BemError(err_code, err_msg);
    });
}

int bem_server_create(bem_ctx ctx, const bem_server_options* options, bem_server* out) {
    if (!ctx || !out) return BEM_ERR_ARGUMENT;
    const int nl = options ? options->lanes : 4;
    if (nl < 1 || nl > 16) {
        ctx->err = "server lanes must be in 1..16";
        return BEM_ERR_ARGUMENT;
    }
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&] {
        auto* srv = new bem_server_s();
        srv->ctx = ctx;
        srv->keep = options ? options->keep_factors != 0 : true;
        try {
            for (int l = 0; l < nl; ++l) {
                ServeLane sl;
                sl.L = make_lane(ctx);
                srv->lanes.push_back(sl);
            }
            for (int l = 0; l < nl; ++l) srv->workers.emplace_back(serve_worker, srv, l);
        } catch (...) {
            {
                std::lock_guard<std::mutex> lk(srv->mu);
                srv->stop = true;
            }
            srv->cv_work.notify_all();
            for (auto& t : srv->workers) t.join();
            for (auto& sl : srv->lanes) free_lane(ctx, sl.L);
            delete srv;
            throw;
        }
        *out = srv;
    });
}

int bem_server_destroy(bem_server srv) {
    if (!srv) return BEM_ERR_ARGUMENT;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
        for (ServeReq* r : srv->queue) finish(srv, r, BEM_ERR_STATE, "server destroyed before the request ran");
        srv->queue.clear();
    }
    srv->cv_work.notify_all();
    srv->cv_done.notify_all();
    for (auto& t : srv->workers) t.join();
    while (!srv->resident.empty()) evict(srv, srv->resident.begin());
    for (auto& sl : srv->lanes) {
        srv->ctx->release(sl.scratch);
        free_lane(srv->ctx, sl.L);
    }
    for (auto& kv : srv->reqs) {
        if (kv.second->mesh) mesh_release(kv.second->mesh);
        delete kv.second;
    }
    delete srv;
    return BEM_OK;
}

int bem_server_submit(bem_server srv, bem_mesh mesh, const bem_problem* problem, double beta, double alpha,
                      int64_t* ticket) {
    if (!srv || !mesh || !problem || !ticket) return BEM_ERR_ARGUMENT;
    return srv_guard(srv, [&] {
        if (mesh->ctx != srv->ctx) throw BemError(BEM_ERR_ARGUMENT, "mesh belongs to another context");
        validate_problem(*problem);
        if (!std::isfinite(beta) || !std::isfinite(alpha))
            throw BemError(BEM_ERR_ARGUMENT, "incidence angles must be finite");
        auto* r = new ServeReq();
        r->mesh = mesh;
        r->prob = *problem;
        r->ang[0] = beta;
        r->ang[1] = alpha;
        r->need = sweep_need(mesh, *problem);
        if (r->need > srv->ctx->budget) {
            delete r;
            throw BemError(BEM_ERR_ADMISSION, "request exceeds the context's HBM budget");
        }
        mesh->refs++;
        r->t_submit = now_ms();
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            r->id = srv->next_id++;
            srv->queue.push_back(r);
            srv->reqs[r->id] = r;
            ++srv->submitted;
        }
        srv->cv_work.notify_all();
        *ticket = r->id;
    });
}

int bem_server_poll(bem_server srv, int64_t ticket, int32_t* done) {
    if (!srv || !done) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(ticket);
    if (it == srv->reqs.end()) {
        srv->msg = "unknown ticket";
        return BEM_ERR_STATE;
    }
    *done = it->second->done ? 1 : 0;
    return BEM_OK;
}

int bem_server_wait(bem_server srv, int64_t ticket, double timeout_ms, double* amplitude, bem_diagnostics* diag,
                    double* timing) {
    if (!srv) return BEM_ERR_ARGUMENT;
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(ticket);
    if (it == srv->reqs.end()) {
        srv->msg = "unknown ticket";
        return BEM_ERR_STATE;
    }
    ServeReq* r = it->second;
    auto ready = [&] { return r->done; };
    if (timeout_ms < 0) {
        srv->cv_done.wait(lk, ready);
    } else if (!srv->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready)) {
        srv->msg = "request not finished";
        return BEM_ERR_STATE;
    }
    if (r->status != BEM_OK) return r->status;  // kept for bem_server_error
    if (amplitude) {
        amplitude[0] = r->amp[0];
        amplitude[1] = r->amp[1];
    }
    if (diag) *diag = r->diag;
    if (timing) {
        timing[0] = r->t_start - r->t_submit;
        timing[1] = r->t_end - r->t_start;
        timing[2] = r->t_end - r->t_submit;
    }
    srv->reqs.erase(it);
    delete r;
    return BEM_OK;
}

const char* bem_server_error(bem_server srv, int64_t ticket) {
    if (!srv) return "null server";
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(ticket);
    if (it != srv->reqs.end() && it->second->done && it->second->status != BEM_OK) {
        srv->msg = it->second->err;
        delete it->second;
        srv->reqs.erase(it);
    }
    return srv->msg.c_str();
}

int bem_server_stats_get(bem_server srv, bem_server_stats* out) {
    if (!srv || !out) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(srv->mu);
    out->submitted = srv->submitted;
    out->completed = srv->completed;
    out->failed = srv->failed;
    out->queued = (int64_t)srv->queue.size();
    out->running = srv->running;
    out->resident_operators = (int64_t)srv->resident.size();
    out->factor_reuses = srv->reuses;
    uint64_t b = 0;
    for (auto& e : srv->resident) b += e.bytes;
    out->resident_bytes = b;
    return BEM_OK;
}

void bem_hmatrix_defaults(bem_hmatrix_options* o) {
    if (!o) return;
    o->leaf_size = 64;
    o->eta = 2.0;
    o->aca_tolerance = 1e-6;
    o->max_rank = 256;
    o->gmres_restart = 50;
    o->gmres_maxiter = 1000;
    o->gmres_tolerance = 1e-8;
    o->preconditioner = 2;
    o->recycle_dimension = 8;
}

int bem_hmatrix_plan(int64_t n, const double* points, int32_t leaf, double eta, double near_radius, int64_t* perm,
                     int64_t* n_nodes, int64_t* n_dense, int64_t* n_lowrank, int64_t* blocks, int64_t cap) {
    if (n <= 0 || !points || leaf < 1 || !(eta > 0) || !(near_radius >= 0)) return BEM_ERR_ARGUMENT;
    try {
        std::vector<double> aos(3 * (size_t)n);
        for (int64_t i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) aos[3 * i + c] = points[3 * i + c];
        HPlan P;
        hmatrix_plan(aos.data(), n, leaf, eta, near_radius, P);
        if (perm) std::copy(P.perm.begin(), P.perm.end(), perm);
        if (n_nodes) *n_nodes = (int64_t)P.nodes.size();
        if (n_dense) *n_dense = (int64_t)P.dense.size();
        if (n_lowrank) *n_lowrank = (int64_t)P.lowrank.size();
        const int64_t tot = (int64_t)(P.dense.size() + P.lowrank.size());
        if (blocks && cap >= 5 * tot) {
            int64_t e = 0;
            auto put = [&](int lr, const HBlock& b) {
                const HNode &r = P.nodes[b.row], &c = P.nodes[b.col];
                blocks[e++] = lr;
                blocks[e++] = r.lo;
                blocks[e++] = r.hi - r.lo;
                blocks[e++] = c.lo;
                blocks[e++] = c.hi - c.lo;
            };
            for (auto& b : P.dense) put(0, b);
            for (auto& b : P.lowrank) put(1, b);
        }
    } catch (...) {
        return BEM_ERR_DEVICE;
    }
    return BEM_OK;
}

int bem_hmatrix_create(bem_ctx ctx, bem_mesh mesh, const bem_problem* problem, const bem_hmatrix_options* options,
                       bem_hmatrix* out) {
    if (!ctx || !mesh || !problem || !out) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&] {
        validate_problem(*problem);
        const bool fluid = problem->boundary == BEM_FLUID;

        if (problem->boundary != BEM_RIGID && !fluid)
            throw BemError(BEM_ERR_ARGUMENT, "the hierarchical path solves rigid surfaces and fluid bodies");
        if (fluid) {
            // compression/fluid.jl:20-24
            if (problem->formulation != BEM_MULLER)
                throw BemError(BEM_ERR_ARGUMENT, "compressed fluid BEM requires formulation=:muller");
        } else if (problem->formulation != BEM_BURTON_MILLER && problem->formulation != BEM_CBIE) {
            throw BemError(BEM_ERR_ARGUMENT, "formulation must be :burton_miller or :cbie");
        }
        bem_hmatrix_options o;
        bem_hmatrix_defaults(&o);
        if (options) o = *options;
        if (o.leaf_size < 1 || o.leaf_size > 64) throw BemError(BEM_ERR_ARGUMENT, "leaf_size must be in 1..64");
        if (!(o.eta > 0)) throw BemError(BEM_ERR_ARGUMENT, "eta must be positive");
        if (!(o.aca_tolerance > 0 && o.aca_tolerance < 1)) throw BemError(BEM_ERR_ARGUMENT, "aca_tolerance must be in (0, 1)");
        if (o.max_rank < 1 || o.max_rank > kAcaCap) throw BemError(BEM_ERR_ARGUMENT, "max_rank must be in 1..256");
        if (o.gmres_restart < 1 || o.gmres_maxiter < 1 || !(o.gmres_tolerance > 0))
            throw BemError(BEM_ERR_ARGUMENT, "invalid GMRES options");
        if (o.preconditioner < 0 || o.preconditioner > 2)
            throw BemError(BEM_ERR_ARGUMENT, "preconditioner must be 0 (none), 1 (block Jacobi) or 2 (two-level)");
        if (o.recycle_dimension < 0) throw BemError(BEM_ERR_ARGUMENT, "recycle_dimension must be nonnegative");
        auto* h = new bem_hmatrix_s();
        h->ctx = ctx;
        h->mesh = mesh;
        mesh->refs++;
        h->prob = *problem;
        h->opt = o;
        h->fluid = fluid;
        // compression/fluid.jl: at max(k, k_int) * radius <= pi/2 the exterior
        // and interior K' and H come from the Calderon identities
        h->calderon = fluid && calderon_applies(mesh->h, *problem);
        h->par = HPar{problem->k, fluid ? problem->k / problem->soundspeed_contrast : problem->k,
                      fluid ? problem->density_contrast : 1.0};
        try {
            h->lane = make_lane(ctx);
            hm_build(h);
        } catch (...) {
            hm_free(h);
            delete h;
            throw;
        }
        *out = h;
    });
}

int bem_hmatrix_destroy(bem_hmatrix h) {
    if (!h) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(h->ctx->mu);
    hm_free(h);
    delete h;
    return BEM_OK;
}

int bem_hmatrix_get_info(bem_hmatrix h, bem_hmatrix_info* out) {
    if (!h || !out) return BEM_ERR_ARGUMENT;
    *out = h->info;
    return BEM_OK;
}

int bem_hmatrix_matvec(bem_hmatrix h, int32_t nrhs, const double* x, double* y) {
    return bem_hmatrix_apply(h, 0, nrhs, x, y);
}

int bem_hmatrix_apply(bem_hmatrix h, int32_t which, int32_t nrhs, const double* x, double* y) {
    if (!h || nrhs <= 0 || !x || !y || (which != 0 && which != 1)) return BEM_ERR_ARGUMENT;
    bem_ctx ctx = h->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&] {
        if (which == 1 && h->fluid)
            throw BemError(BEM_ERR_ARGUMENT, "a fluid body's right-hand side is not an operator");
        const HSys S = hm_system(h);
        const int64_t n = S.n;
        hipStream_t s = h->lane->s;
        z64* X = (z64*)ctx->alloc(sizeof(z64) * n * kHRhs);
        z64* Y = nullptr;
        z64* pw = nullptr;
        try {
            Y = (z64*)ctx->alloc(sizeof(z64) * n * kHRhs);
            pw = (z64*)ctx->alloc(sizeof(z64) * std::max<size_t>(1, S.work_size(kHRhs)));
            std::vector<z64> hx((size_t)n * kHRhs), hy((size_t)n * kHRhs);
            for (int c0 = 0; c0 < nrhs; c0 += kHRhs) {
                const int nr = std::min(kHRhs, nrhs - c0);
                for (int64_t pos = 0; pos < n; ++pos)
                    for (int c = 0; c < nr; ++c) {
                        const double* src = x + 2 * ((size_t)(c0 + c) * n + h->sysperm[pos]);
                        hx[pos * nr + c] = z64{src[0], src[1]};
                    }
                HIPCHECK(hipMemcpyAsync(X, hx.data(), sizeof(z64) * n * nr, hipMemcpyHostToDevice, s));
                if (which == 0) S.apply(X, nr, Y, pw);
                else hm_apply(h, h->B, X, nr, Y);
                HIPCHECK(hipGetLastError());
                HIPCHECK(hipMemcpyAsync(hy.data(), Y, sizeof(z64) * n * nr, hipMemcpyDeviceToHost, s));
                HIPCHECK(hipStreamSynchronize(s));
                for (int64_t pos = 0; pos < n; ++pos)
                    for (int c = 0; c < nr; ++c) {
                        double* dst = y + 2 * ((size_t)(c0 + c) * n + h->sysperm[pos]);
                        dst[0] = hy[pos * nr + c].x;
                        dst[1] = hy[pos * nr + c].y;
                    }
            }
        } catch (...) {
            hipStreamSynchronize(s);
            ctx->release(X);
            ctx->release(Y);
            ctx->release(pw);
            throw;
        }
        ctx->release(X);
        ctx->release(Y);
        ctx->release(pw);
    });
}

int bem_hmatrix_solve(bem_hmatrix h, int32_t nrhs, const double* angles, bem_solution* out) {
    if (!h || nrhs <= 0 || !angles || !out) return BEM_ERR_ARGUMENT;
    bem_ctx ctx = h->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&] {
        for (int r = 0; r < 2 * nrhs; ++r)
            if (!std::isfinite(angles[r])) throw BemError(BEM_ERR_ARGUMENT, "incidence angles must be finite");
        const int64_t n = h->nsys;
        const int nt = (int)h->nt;
        hipStream_t s = h->lane->s;
        auto* sol = new bem_solution_s();
        sol->ctx = ctx;
        sol->mesh = h->mesh;
        h->mesh->refs++;
        sol->k = h->prob.k;
        sol->prob = h->prob;
        sol->nrhs = nrhs;
        sol->nt = nt;
        sol->angles.assign(angles, angles + 2 * nrhs);
        sol->diag.resize(nrhs);
        std::vector<void*> tmp;
        try {
            const double t0 = now_ms();
            sol->p = (z64*)ctx->alloc(sizeof(z64) * (size_t)nt * nrhs);
            sol->q = (z64*)ctx->alloc(sizeof(z64) * (size_t)nt * nrhs);
            std::vector<double> dirs(3 * nrhs);
            for (int r = 0; r < nrhs; ++r) direction(angles[2 * r], angles[2 * r + 1], &dirs[3 * r]);
            double* dd = (double*)ctx->alloc(sizeof(double) * 3 * nrhs);
            tmp.push_back(dd);
            HIPCHECK(hipMemcpyAsync(dd, dirs.data(), sizeof(double) * 3 * nrhs, hipMemcpyHostToDevice, s));
            z64* B = (z64*)ctx->alloc(sizeof(z64) * n * kHRhs);
            tmp.push_back(B);
            z64* X = (z64*)ctx->alloc(sizeof(z64) * n * kHRhs);
            tmp.push_back(X);
            std::vector<bem_diagnostics> dg;
            for (int c0 = 0; c0 < nrhs; c0 += kHRhs) {
                const int nr = std::min(kHRhs, nrhs - c0);
                if (h->fluid) {
                    // Mueller: b = [p_inc; q_inc] (full.jl:442-447)
                    launch_hfluid_rhs(s, h->mesh->d, h->prob.k, nr, dd + 3 * c0, h->dsysperm, B);
                } else {
                    // b = B q_inc (the right-hand side of the scattered-field equation)
                    launch_hqinc(s, h->mesh->d, h->prob.k, nr, dd + 3 * c0, h->perm, X);
                    hm_apply(h, h->B, X, nr, B);
                }
                dg.assign(nr, bem_diagnostics{});
                hm_solve_recycled(h, nr, B, X, dg);
                // traces of this chunk: columns c0 .. c0 + nr of P, Q
                if (h->fluid)
                    launch_hfluid_traces(s, h->mesh->d, h->prob.k, nr, dd + 3 * c0, h->dsysperm, X,
                                         sol->p + (size_t)c0 * nt, sol->q + (size_t)c0 * nt);
                else
                    launch_htraces(s, h->mesh->d, h->prob.k, nr, dd + 3 * c0, h->perm, X, sol->p + (size_t)c0 * nt,
                                   sol->q + (size_t)c0 * nt);
                HIPCHECK(hipGetLastError());
                for (int c = 0; c < nr; ++c) {
                    sol->diag[c0 + c] = dg[c];
                    sol->diag[c0 + c].derivative_evaluation = h->fluid ? (h->calderon ? 2 : 1) : 0;
                }
            }
            sol->hp.resize((size_t)nt * nrhs);
            sol->hq.resize((size_t)nt * nrhs);
            HIPCHECK(hipMemcpyAsync(sol->hp.data(), sol->p, sizeof(z64) * nt * nrhs, hipMemcpyDeviceToHost, s));
            HIPCHECK(hipMemcpyAsync(sol->hq.data(), sol->q, sizeof(z64) * nt * nrhs, hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            const double ms = now_ms() - t0;
            for (auto& d : sol->diag) {
                d.assembly_ms = h->info.assembly_ms;
                d.factor_ms = h->info.aca_ms;
                d.solve_ms = ms;
            }
        } catch (...) {
            hipStreamSynchronize(s);
            for (void* p : tmp) ctx->release(p);
            ctx->release(sol->p);
            ctx->release(sol->q);
            mesh_release(sol->mesh);
            delete sol;
            throw;
        }
        for (void* p : tmp) ctx->release(p);
        *out = sol;
    });
}

int bem_debug_lu(bem_ctx ctx, int64_t n64, const double* A, int32_t flags, double* LUout, int64_t* ipiv_out,
                 const double* b, double* x) {
    if (!ctx || !A || n64 <= 0 || n64 > 65536 || (flags & ~31) != 0) return BEM_ERR_ARGUMENT;
    if ((flags & 4) && ((flags & 8) || !(flags & 1))) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&] {
        const int n = (int)n64;
        const int single = flags & 1;
        const bool adjoint = (flags & 2) != 0;
        hipStream_t s = ctx->lane0()->s;
        const size_t esz = single ? sizeof(z32) : sizeof(z64);
        std::vector<char> host((size_t)n * n * esz);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                const double re = A[2 * (i + (size_t)j * n)], im = A[2 * (i + (size_t)j * n) + 1];
                if (single) ((z32*)host.data())[(size_t)i * n + j] = z32{(float)re, (float)im};
                else ((z64*)host.data())[(size_t)i * n + j] = z64{re, im};
            }
        bem_operator_s op;
        op.ctx = ctx;
        op.lane = ctx->lane0();
        op.n = n;
        op.single = single != 0;
        std::vector<void*> bufs;
        auto al = [&](size_t bytes) { void* p = ctx->alloc(bytes); bufs.push_back(p); return p; };
        try {
            op.LU = al(esz * (size_t)n * n);
            op.ipiv = (int*)al(sizeof(int) * n);
            op.perm = (int*)al(sizeof(int) * n);
            op.inv = al(esz * 2 * 4096 * (size_t)((n + 63) / 64));
            HIPCHECK(hipMemcpyAsync(op.LU, host.data(), host.size(), hipMemcpyHostToDevice, s));
            LuTune t = lu_tune_for(ctx, n, single != 0);
            if (flags & 4) t.split = 1;
            if (flags & 8) t.split = 0;
            if (flags & 16) t.graph = 1;
            const LuTune* force = (flags & 28) ? &t : nullptr;
            if (single) factor_into<float>(&op, op.LU, op.ipiv, op.perm, op.inv, force);
            else factor_into<double>(&op, op.LU, op.ipiv, op.perm, op.inv, force);
            HIPCHECK(hipMemcpyAsync(host.data(), op.LU, host.size(), hipMemcpyDeviceToHost, s));
            std::vector<int> hp(n);
            HIPCHECK(hipMemcpyAsync(hp.data(), op.ipiv, sizeof(int) * n, hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            if (LUout)
                for (int i = 0; i < n; ++i)
                    for (int j = 0; j < n; ++j) {
                        double re, im;
                        if (single) { z32 v = ((z32*)host.data())[(size_t)i * n + j]; re = v.x; im = v.y; }
                        else { z64 v = ((z64*)host.data())[(size_t)i * n + j]; re = v.x; im = v.y; }
                        LUout[2 * (i + (size_t)j * n)] = re;
                        LUout[2 * (i + (size_t)j * n) + 1] = im;
                    }
            if (ipiv_out)
                for (int i = 0; i < n; ++i) ipiv_out[i] = hp[i] + 1;
            if (b && x) {
                op.vb = (z64*)al(sizeof(z64) * n);
                op.vx = (z64*)al(sizeof(z64) * n);
                op.wv = al(sizeof(z64) * n);
                op.wv2 = al(sizeof(z64) * n);
                op.wv3 = al(sizeof(z64) * n);
                op.red = (double*)al(sizeof(double) * 5 * kRedBlocks);
                op.precision = single ? 1 : 0;
                HIPCHECK(hipMemcpyAsync(op.vb, b, sizeof(z64) * n, hipMemcpyHostToDevice, s));
                if (!adjoint) {
                    if (!correction(&op, op.vb, op.vx, 0, nullptr)) throw BemError(BEM_ERR_ARGUMENT, "nonfinite b");
                } else if (single) {
                    launch_prep_rhs<float>(s, n, op.vb, nullptr, 1.0, (z32*)op.wv);
                    lu_solve_h<float>(s, (const z32*)op.LU, n, n, op.perm, (const z32*)op.inv, (const z32*)op.wv,
                                      (z32*)op.wv2, (z32*)op.wv3, op.lane->ws, &op.lane->pl);
                    launch_finish<float>(s, n, (const z32*)op.wv3, 1.0, nullptr, op.vx, 0, nullptr);
                } else {
                    lu_solve_h<double>(s, (const z64*)op.LU, n, n, op.perm, (const z64*)op.inv, op.vb,
                                       (z64*)op.wv2, op.vx, op.lane->ws, &op.lane->pl);
                }
                int expired = 0;
                HIPCHECK(hipMemcpyAsync(x, op.vx, sizeof(z64) * n, hipMemcpyDeviceToHost, s));
                HIPCHECK(hipMemcpyAsync(&expired, op.lane->ws.status + 2, sizeof(int), hipMemcpyDeviceToHost, s));
                HIPCHECK(hipStreamSynchronize(s));
                tri_check(op.lane, expired);
            }
        } catch (...) {
            hipStreamSynchronize(s);
            drop_graphs(op.lane, op.LU);
            for (void* p : bufs) ctx->release(p);
            for (void* p : op.heap) ctx->release(p);
            op.heap.clear();
            op.LU = op.inv = op.wv = op.wv2 = op.wv3 = op.split = nullptr;
            op.ipiv = op.perm = nullptr;
            op.vb = op.vx = nullptr;
            op.red = nullptr;
            throw;
        }
        drop_graphs(op.lane, op.LU);
        for (void* p : bufs) ctx->release(p);
        for (void* p : op.heap) ctx->release(p);
        op.heap.clear();
    });
}


// ---------------------------------------------------------------------------
// axisymmetric BEM
// ---------------------------------------------------------------------------
int bem_axi_operators(bem_ctx ctx, int64_t n_nodes, const double* rho, const double* z, double k, int32_t m0,
                      int32_t n_modes, double* K, double* V) {
    if (!ctx || !rho || !z || !K || !V || m0 < 0 || n_modes < 1 || n_modes > 256 || !(std::isfinite(k) && k > 0))
        return BEM_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&] {
        axi_validate(n_nodes, rho, z);
        auto* h = new bem_axi_s();
        h->ctx = ctx;
        std::vector<void*> tmp;
        try {
            axi_upload(h, n_nodes, rho, z);
            const int n = h->n;
            hipStream_t s = ctx->lane0()->s;
            const size_t nn = (size_t)n * n;
            z64* dK = (z64*)ctx->alloc(sizeof(z64) * nn * kAxiModes);
            tmp.push_back(dK);
            z64* dV = (z64*)ctx->alloc(sizeof(z64) * nn * kAxiModes);
            tmp.push_back(dV);
            const AxiRules R = axi_rules();
            std::vector<z64> hk(nn * kAxiModes), hv(nn * kAxiModes);
            for (int c0 = 0; c0 < n_modes; c0 += kAxiModes) {
                const int nm = std::min(kAxiModes, n_modes - c0);
                launch_axi_ops(s, axi_panels(h), R, k, m0 + c0, nm, 0, dK, dV, nullptr, nullptr);
                HIPCHECK(hipGetLastError());
                HIPCHECK(hipMemcpyAsync(hk.data(), dK, sizeof(z64) * nn * nm, hipMemcpyDeviceToHost, s));
                HIPCHECK(hipMemcpyAsync(hv.data(), dV, sizeof(z64) * nn * nm, hipMemcpyDeviceToHost, s));
                HIPCHECK(hipStreamSynchronize(s));
                // mode c0 + c, column-major n x n, interleaved complex
                for (int c = 0; c < nm; ++c)
                    for (int i = 0; i < n; ++i)
                        for (int j = 0; j < n; ++j) {
                            const size_t src = ((size_t)c * n + i) * n + j;
                            const size_t dst = 2 * (((size_t)(c0 + c) * n + j) * n + i);
                            K[dst] = hk[src].x;
                            K[dst + 1] = hk[src].y;
                            V[dst] = hv[src].x;
                            V[dst + 1] = hv[src].y;
                        }
            }
        } catch (...) {
            for (void* q : tmp) ctx->release(q);
            axi_free(h);
            throw;
        }
        for (void* q : tmp) ctx->release(q);
        axi_free(h);
    });
}

int bem_axi_solve(bem_ctx ctx, int64_t n_nodes, const double* rho, const double* z, const bem_problem* problem,
                  double incidence_angle, int32_t m_max, bem_axi* out) {
    if (!ctx || !rho || !z || !problem || !out) return BEM_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&] {
        const bem_problem& pr = *problem;
        if (!(std::isfinite(pr.k) && pr.k > 0)) throw BemError(BEM_ERR_ARGUMENT, "k must be finite and positive");
        if (!std::isfinite(incidence_angle)) throw BemError(BEM_ERR_ARGUMENT, "incidence angles must be finite");
        if (m_max < 0 || m_max > 255) throw BemError(BEM_ERR_ARGUMENT, "m_max must be in 0..255");
        const bool fluid = pr.boundary == BEM_FLUID;
        if (pr.boundary != BEM_RIGID && pr.boundary != BEM_PRESSURE_RELEASE && !fluid)
            throw BemError(BEM_ERR_ARGUMENT, "unknown boundary condition");
        if (fluid && !(std::isfinite(pr.density_contrast) && pr.density_contrast > 0 &&
                       std::isfinite(pr.soundspeed_contrast) && pr.soundspeed_contrast > 0))
            throw BemError(BEM_ERR_ARGUMENT, "fluid density and soundspeed contrasts must be finite and positive");
        axi_validate(n_nodes, rho, z);
        const double t0 = now_ms();
        auto* h = new bem_axi_s();
        h->ctx = ctx;
        h->prob = pr;
        h->beta = incidence_angle;
        const bool axial = incidence_angle == 0.0;
        if (axial) h->modes = {0};
        else
            for (int m = 0; m <= m_max; ++m) h->modes.push_back(m);
        std::vector<void*> tmp;
        auto al = [&](size_t b) { void* q = ctx->alloc(b); tmp.push_back(q); return q; };
        Lane* L = ctx->lane0();
        hipStream_t s = L->s;
        try {
            axi_upload(h, n_nodes, rho, z);
            const int n = h->n, nm_all = (int)h->modes.size();
            const int ns = fluid ? 2 * n : n;
            const size_t nn = (size_t)n * n;
            const double k = pr.k, g = fluid ? pr.density_contrast : 1.0;
            const double ki = fluid ? k / pr.soundspeed_contrast : k;
            const AxiPanels P = axi_panels(h);
            const AxiRules R = axi_rules();
            const int chunk = std::min(kAxiModes, nm_all);
            z64* Ke = (z64*)al(sizeof(z64) * nn * chunk);
            z64* Ve = (z64*)al(sizeof(z64) * nn * chunk);
            z64 *Ki = nullptr, *Vi = nullptr, *RK = nullptr, *RV = nullptr;
            if (fluid) {
                Ki = (z64*)al(sizeof(z64) * nn * chunk);
                Vi = (z64*)al(sizeof(z64) * nn * chunk);
            }
            if (axial && !fluid) {
                RK = (z64*)al(sizeof(z64) * nn);
                RV = (z64*)al(sizeof(z64) * nn);
            }
            z64* A = (z64*)al(sizeof(z64) * (size_t)ns * ns);
            z64* b = (z64*)al(sizeof(z64) * ns);
            z64* x = (z64*)al(sizeof(z64) * ns);
            z64* r = (z64*)al(sizeof(z64) * ns);
            z64* w1 = (z64*)al(sizeof(z64) * ns);
            z64* w2 = (z64*)al(sizeof(z64) * n);
            z64* vin = (z64*)al(sizeof(z64) * 2 * n);
            bem_operator_s op;
            op.ctx = ctx;
            op.lane = L;
            op.n = ns;
            op.single = false;
            op.LU = al(sizeof(z64) * (size_t)ns * ns);
            op.ipiv = (int*)al(sizeof(int) * ns);
            op.perm = (int*)al(sizeof(int) * ns);
            op.inv = al(sizeof(z64) * 2 * 4096 * (size_t)((ns + 63) / 64));
            h->dp = (z64*)ctx->alloc(sizeof(z64) * (size_t)n * nm_all);
            h->dd = (z64*)ctx->alloc(sizeof(z64) * (size_t)n * nm_all);
            h->hp.resize((size_t)n * nm_all);
            h->hd.resize((size_t)n * nm_all);
            const z64 one{1.0, 0.0}, zero{0.0, 0.0}, half{0.5, 0.0}, mone{-1.0, 0.0};
            std::vector<std::complex<double>> pinc, dinc;
            std::vector<z64> hb(ns), hv2(2 * n), hx(ns), hr(ns), hw(n);
            double t_asm = 0;
            for (int c0 = 0; c0 < nm_all; c0 += chunk) {
                const int nm = std::min(chunk, nm_all - c0);
                const double ta = now_ms();
                launch_axi_ops(s, P, R, k, h->modes[c0], nm, axial && !fluid, Ke, Ve, RK, RV);
                if (fluid) launch_axi_ops(s, P, R, ki, h->modes[c0], nm, 0, Ki, Vi, nullptr, nullptr);
                HIPCHECK(hipGetLastError());
                HIPCHECK(hipStreamSynchronize(s));
                t_asm += now_ms() - ta;
                for (int c = 0; c < nm; ++c) {
                    const int mi = c0 + c, m = h->modes[mi];
                    const z64 *Kc = Ke + nn * c, *Vc = Ve + nn * c;
                    axi_incident(h, k, m, pinc, dinc);
                    for (int j = 0; j < n; ++j) {
                        hv2[j] = z64{pinc[j].real(), pinc[j].imag()};
                        hv2[n + j] = z64{dinc[j].real(), dinc[j].imag()};
                    }
                    HIPCHECK(hipMemcpyAsync(vin, hv2.data(), sizeof(z64) * 2 * n, hipMemcpyHostToDevice, s));
                    // the system and its right-hand side (axisymmetric.jl:683-729,995-1057,1185-1300)
                    if (pr.boundary == BEM_RIGID) {
                        launch_axi_block(s, n, half, mone, Kc, zero, nullptr, A, n);
                        if (axial) launch_axi_rowsum(s, n, RV, b);
                        else launch_matvec<double>(s, Vc, n, n, n, vin + n, b, nullptr, nullptr);
                    } else if (pr.boundary == BEM_PRESSURE_RELEASE) {
                        launch_axi_block(s, n, zero, one, Vc, zero, nullptr, A, n);
                        // K p - p / 2 with p = -p_inc; axial: -rowsum(RK)
                        if (axial) launch_axi_rowsum(s, n, RK, w2);
                        else launch_matvec<double>(s, Kc, n, n, n, vin, w2, nullptr, nullptr);
                        HIPCHECK(hipMemcpyAsync(hw.data(), w2, sizeof(z64) * n, hipMemcpyDeviceToHost, s));
                        HIPCHECK(hipStreamSynchronize(s));
                        for (int j = 0; j < n; ++j) {
                            const std::complex<double> v = -std::complex<double>(hw[j].x, hw[j].y) + 0.5 * pinc[j];
                            hb[j] = z64{v.real(), v.imag()};
                        }
                        HIPCHECK(hipMemcpyAsync(b, hb.data(), sizeof(z64) * n, hipMemcpyHostToDevice, s));
                    } else {
                        const z64 *Kic = Ki + nn * c, *Vic = Vi + nn * c;
                        launch_axi_block(s, n, half, mone, Kc, zero, nullptr, A, ns);
                        launch_axi_block(s, n, zero, one, Vc, zero, nullptr, A + n, ns);
                        launch_axi_block(s, n, half, one, Kic, zero, nullptr, A + (size_t)n * ns, ns);
                        launch_axi_block(s, n, zero, z64{-g, 0.0}, Vic, zero, nullptr, A + (size_t)n * ns + n, ns);
                        // -[(1/2 I + Ki) p_inc - g Vi dpdn_inc]
                        launch_matvec<double>(s, Kic, n, n, n, vin, w2, nullptr, nullptr);
                        launch_matvec<double>(s, Vic, n, n, n, vin + n, w1, nullptr, nullptr);
                        HIPCHECK(hipMemcpyAsync(hw.data(), w2, sizeof(z64) * n, hipMemcpyDeviceToHost, s));
                        HIPCHECK(hipMemcpyAsync(hx.data(), w1, sizeof(z64) * n, hipMemcpyDeviceToHost, s));
                        HIPCHECK(hipStreamSynchronize(s));
                        for (int j = 0; j < n; ++j) {
                            hb[j] = zero;
                            const std::complex<double> v = -(0.5 * pinc[j] + std::complex<double>(hw[j].x, hw[j].y) -
                                                              g * std::complex<double>(hx[j].x, hx[j].y));
                            hb[n + j] = z64{v.real(), v.imag()};
                        }
                        HIPCHECK(hipMemcpyAsync(b, hb.data(), sizeof(z64) * ns, hipMemcpyHostToDevice, s));
                    }
                    HIPCHECK(hipGetLastError());
                    // FP64 LU of a copy (A stays for the residual), solve, residual
                    HIPCHECK(hipMemcpyAsync(op.LU, A, sizeof(z64) * (size_t)ns * ns, hipMemcpyDeviceToDevice, s));
                    factor_into<double>(&op, op.LU, op.ipiv, op.perm, op.inv);
                    lu_solve<double>(s, (const z64*)op.LU, ns, ns, op.perm, (const z64*)op.inv, b, w1, x, L->ws,
                                     &L->pl);
                    launch_matvec<double>(s, A, ns, ns, ns, x, r, b, nullptr);
                    int expired = 0;
                    HIPCHECK(hipMemcpyAsync(hx.data(), x, sizeof(z64) * ns, hipMemcpyDeviceToHost, s));
                    HIPCHECK(hipMemcpyAsync(hr.data(), r, sizeof(z64) * ns, hipMemcpyDeviceToHost, s));
                    HIPCHECK(hipMemcpyAsync(hb.data(), b, sizeof(z64) * ns, hipMemcpyDeviceToHost, s));
                    HIPCHECK(hipMemcpyAsync(&expired, L->ws.status + 2, sizeof(int), hipMemcpyDeviceToHost, s));
                    HIPCHECK(hipStreamSynchronize(s));
                    tri_check(L, expired);
                    double rn = 0, bn = 0;
                    for (int q = 0; q < ns; ++q) {
                        rn += hr[q].x * hr[q].x + hr[q].y * hr[q].y;
                        bn += hb[q].x * hb[q].x + hb[q].y * hb[q].y;
                    }
                    const double rel = bn > 0 ? std::sqrt(rn / bn) : std::sqrt(rn);
                    if (!std::isfinite(rel)) throw BemError(BEM_ERR_NUMERIC, "non-finite solution");
                    h->max_residual = std::max(h->max_residual, rel);
                    // scattered traces of this mode
                    for (int j = 0; j < n; ++j) {
                        z64 pj, dj;
                        if (pr.boundary == BEM_RIGID) {
                            pj = hx[j];
                            dj = z64{-dinc[j].real(), -dinc[j].imag()};
                        } else if (pr.boundary == BEM_PRESSURE_RELEASE) {
                            pj = z64{-pinc[j].real(), -pinc[j].imag()};
                            dj = hx[j];
                        } else {
                            pj = hx[j];
                            dj = hx[n + j];
                        }
                        h->hp[(size_t)mi * n + j] = pj;
                        h->hd[(size_t)mi * n + j] = dj;
                    }
                }
            }
            HIPCHECK(hipMemcpyAsync(h->dp, h->hp.data(), sizeof(z64) * h->hp.size(), hipMemcpyHostToDevice, s));
            HIPCHECK(hipMemcpyAsync(h->dd, h->hd.data(), sizeof(z64) * h->hd.size(), hipMemcpyHostToDevice, s));
            HIPCHECK(hipStreamSynchronize(s));
            h->assembly_ms = t_asm;
            h->solve_ms = now_ms() - t0 - t_asm;
        } catch (...) {
            hipStreamSynchronize(s);
            for (void* q : tmp) ctx->release(q);
            axi_free(h);
            throw;
        }
        for (void* q : tmp) ctx->release(q);
        *out = h;
    });
}

int bem_axi_info(bem_axi sol, int32_t* n_panels, int32_t* n_modes, double* max_relative_residual,
                 double* assembly_ms, double* solve_ms) {
    if (!sol) return BEM_ERR_ARGUMENT;
    if (n_panels) *n_panels = sol->n;
    if (n_modes) *n_modes = (int32_t)sol->modes.size();
    if (max_relative_residual) *max_relative_residual = sol->max_residual;
    if (assembly_ms) *assembly_ms = sol->assembly_ms;
    if (solve_ms) *solve_ms = sol->solve_ms;
    return BEM_OK;
}

int bem_axi_traces(bem_axi sol, int32_t mode, double* p_scat, double* dpdn_scat) {
    if (!sol || mode < 0 || mode >= (int32_t)sol->modes.size()) return BEM_ERR_ARGUMENT;
    const int n = sol->n;
    for (int j = 0; j < n; ++j) {
        const z64 p = sol->hp[(size_t)mode * n + j], d = sol->hd[(size_t)mode * n + j];
        if (p_scat) {
            p_scat[2 * j] = p.x;
            p_scat[2 * j + 1] = p.y;
        }
        if (dpdn_scat) {
            dpdn_scat[2 * j] = d.x;
            dpdn_scat[2 * j + 1] = d.y;
        }
    }
    return BEM_OK;
}

int bem_axi_far_field(bem_axi sol, int64_t n_dir, const double* angles, double* out) {
    if (!sol || n_dir <= 0 || n_dir > 65535 || !angles || !out) return BEM_ERR_ARGUMENT;
    bem_ctx ctx = sol->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&] {
        for (int64_t q = 0; q < 2 * n_dir; ++q)
            if (!std::isfinite(angles[q])) throw BemError(BEM_ERR_ARGUMENT, "directions must be finite");
        hipStream_t s = ctx->lane0()->s;
        double* da = (double*)ctx->alloc(sizeof(double) * 2 * n_dir);
        z64* dout = nullptr;
        try {
            dout = (z64*)ctx->alloc(sizeof(z64) * n_dir);
            HIPCHECK(hipMemcpyAsync(da, angles, sizeof(double) * 2 * n_dir, hipMemcpyHostToDevice, s));
            launch_axi_far(s, axi_panels(sol), axi_rules(), sol->prob.k, (int)sol->modes.size(), sol->dp, sol->dd,
                           (int)n_dir, da, dout);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipMemcpyAsync(out, dout, sizeof(z64) * n_dir, hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
        } catch (...) {
            hipStreamSynchronize(s);
            ctx->release(da);
            ctx->release(dout);
            throw;
        }
        ctx->release(da);
        ctx->release(dout);
    });
}

int bem_axi_destroy(bem_axi sol) {
    if (!sol) return BEM_ERR_ARGUMENT;
    bem_ctx ctx = sol->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    axi_free(sol);
    return BEM_OK;
}

}  // extern "C"
